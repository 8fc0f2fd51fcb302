"""Vertical maps of the exhaustive Gamma oracle against the reference's
fixtures: every contract fixture's vertical_maps (full_pi0 and hgp_reduced),
the vertical_q1_growth_target regression (the target root is read in its
state at the assignment level), and the naturality of the maps on random
clouds."""

import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle import tower as otower

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "contracts.json"), encoding="utf-8") as fh:
    CONTRACTS = json.load(fh)
with open(os.path.join(HERE, "golden", "regressions.json"), encoding="utf-8") as fh:
    REG = json.load(fh)


def _frac(v):
    return Fraction(int(v[0]), int(v[1])) if isinstance(v, list) else Fraction(v["numerator"], v["denominator"])


@pytest.mark.parametrize("case", [c for c in CONTRACTS if c.get("vertical_maps")], ids=lambda c: c["name"])
def test_gamma_vertical_maps_equal_contract_fixture(case):
    pts = np.asarray(case["points"], dtype=float)
    forests, maps = otower.gamma_forests(pts, case["k_max"], case["profile"], vertical=True)
    for vm in case["vertical_maps"]:
        src = case["forests"][vm["source_order"] - 1]["nodes"]
        tgt = case["forests"][vm["target_order"] - 1]["nodes"]

        def key(n):
            return (n["kind"], _frac(n["level"]), tuple(n["covered"]))

        want = sorted((key(src[a["source"]]), key(tgt[a["target"]])) for a in vm["assignments"])
        # the fixture's assignment level is the source node's own level
        assert all(_frac(a["level"]) == _frac(src[a["source"]]["level"]) for a in vm["assignments"])
        ms, mt = forests[vm["source_order"]], forests[vm["target_order"]]
        got = sorted(((n["kind"], n["level"], n["covered"]), (mt[t]["kind"], mt[t]["level"], mt[t]["covered"]))
                     for n, t in zip(ms, maps[vm["source_order"]]))
        assert got == want


def test_gamma_vertical_growth_target():
    fx = REG["vertical_q1_growth_target"]
    w = fx["vertical_witness"]
    pts = np.asarray(fx["points"], dtype=float)
    forests, maps = otower.gamma_forests(pts, fx["k_max"], fx["profile"], vertical=True)
    at = _frac(w["at_squared_level"])
    src = [i for i, n in enumerate(forests[w["source_order"]]) if n["level"] == at]
    assert len(src) == 1 and list(forests[w["source_order"]][src[0]]["covered"]) == w["source_covered_point_ids"]
    tgt = forests[w["target_order"]][maps[w["source_order"]][src[0]]]
    assert tgt["level"] == _frac(w["target_birth_squared_level"])
    assert list(tgt["covered"]) == w["target_birth_covered_point_ids"]
    # replayed at the assignment level the target covers more than at its birth
    assert set(w["target_replayed_covered_point_ids"]) > set(tgt["covered"])


def test_gamma_vertical_naturality_random():
    rng = np.random.default_rng(12)
    for trial in range(6):
        n = int(rng.integers(5, 9))
        pts = oracle.canonicalize(rng.integers(-30, 30, size=(n, 3)) + rng.random((n, 3)) * 1e-3)
        k = int(rng.integers(2, n + 1))
        forests, maps = otower.gamma_forests(pts, k, "hgp_reduced", vertical=True)
        for o, targets in maps.items():
            parent = {}
            for i, nd in enumerate(forests[o - 1]):
                for c in nd["children"]:
                    parent[c] = i
            src_parent = {}
            for i, nd in enumerate(forests[o]):
                for c in nd["children"]:
                    src_parent[c] = i
            for i, t in enumerate(targets):
                assert forests[o - 1][t]["level"] <= forests[o][i]["level"]
                if i in src_parent:
                    x, up = t, targets[src_parent[i]]
                    while x != up and x in parent:
                        x = parent[x]
                    assert x == up
