"""Contract v2 output of the GPU path: the product's catalogue serialised
with e-hgp_amd/contract.py equals the reference's contract fixtures' critical
events record for record (canonical SHA-256 event ids included), and its
forests and vertical maps serialise to complete records."""

import json
import os

import numpy as np
import pytest

from helpers import mhgp
import contract

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "contracts.json"), encoding="utf-8") as fh:
    CONTRACTS = json.load(fh)


@pytest.mark.parametrize("case", CONTRACTS, ids=lambda c: c["name"])
def test_product_events_equal_contract_fixture(case):
    pts = np.asarray(case["points"], dtype=float)
    print(f"contract case {case['name']}", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(case["k_max"])
        tw = ctx.tower(case["k_max"]) if case["profile"] == "hgp_reduced" else None
    mine = contract.critical_events(pts, cat, case["k_max"])
    assert mine == sorted(case["critical_events"], key=lambda r: r["event_id"])
    if tw is not None:
        from fractions import Fraction
        level_of = {-1: Fraction(0)}
        by_key = {(tuple(e["minimal_support_ids"]), tuple(e["interior_ids"])):
                  Fraction(int(e["squared_level_exact"]["numerator"]), int(e["squared_level_exact"]["denominator"]))
                  for e in mine}
        for i in range(len(cat.size)):
            level_of[int(cat.level_class[i])] = by_key[(tuple(int(x) for x in cat.ids[i, : cat.size[i]]),
                                                        tuple(cat.interior[i]))]
        forests, vmaps = contract.forests_and_vertical_maps(tw, level_of, mine)
        assert sum(len(f["nodes"]) for f in forests) == len(tw["order"])
        for f, ref in zip(forests, case["forests"]):
            assert len(f["nodes"]) == len(ref["nodes"]) and len(f["root_ids"]) == len(ref["roots"])
        assert [len(v["assignments"]) for v in vmaps] == [len(v["assignments"]) for v in case["vertical_maps"]]
