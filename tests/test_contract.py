"""Contract v2 serialization (e-hgp_amd/contract.py): the CriticalEvent
records built from a catalogue equal the reference's contract fixtures
record for record, canonical SHA-256 event ids included
(tests/fixtures/contracts/*.json critical_catalog).  The catalogue fed in
here is the oracle's; the GPU test feeds the product's."""

import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

import oracle
from helpers import PKG  # noqa: F401  (puts the package on the path)
import contract

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "contracts.json"), encoding="utf-8") as fh:
    CONTRACTS = json.load(fh)


def catalog_view(records):
    """The oracle's records in the layout of mhgp.Catalog (sizes 2..4)."""
    recs = [r for r in records if len(r.support) >= 2]
    ids = np.full((len(recs), 4), 0xFFFFFFFF, dtype=np.uint32)
    for i, r in enumerate(recs):
        ids[i, : len(r.support)] = r.support
    return SimpleNamespace(ids=ids, size=np.array([len(r.support) for r in recs]),
                           rank=np.array([r.rank for r in recs]), interior=[r.interior for r in recs])


@pytest.mark.parametrize("case", CONTRACTS, ids=lambda c: c["name"])
def test_event_records_equal_contract_fixture(case):
    pts = np.asarray(case["points"], dtype=float)
    cat = catalog_view(oracle.catalog(pts, case["k_max"]))
    mine = contract.critical_events(pts, cat, case["k_max"])
    want = sorted(case["critical_events"], key=lambda r: r["event_id"])
    assert [r["event_id"] for r in mine] == [r["event_id"] for r in want]
    assert mine == want


def test_canonical_id_is_domain_separated_and_version_free():
    a = contract.canonical_id("CriticalEvent", {"x": 1, "schema_version": "2.0.0"})
    assert a == contract.canonical_id("CriticalEvent", {"x": 1})
    assert a != contract.canonical_id("MergeNode", {"x": 1})
    assert contract.canonical_json({"b": 1, "a": [1, 2]}) == '{"a":[1,2],"b":1}'
