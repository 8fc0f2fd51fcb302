"""Degeneracy and regression fixtures of the reference, and scale
invariance, through the GPU path (C ABI).

* tests/fixtures/degeneracies/one_ulp_boundaries.json: a fifth point one ulp
  inside, on, or one ulp outside the sphere of a regular tetrahedron, and a
  fourth point one ulp below, on, or above the plane of three others.
* tests/fixtures/regressions/morse_rank_window.json: a sphere with two
  interior, two shell and one exterior point is a rank-4 event at level 1.
* morsehgp3d/tests/fixtures/spatial/relevant_gp_extra_shell_above_smax.json:
  a pair whose diametral sphere carries a third point is an extra-shell
  diagnostic, and the direct tower is refused for such a catalogue.
* Scale invariance: a cloud scaled by 2^-120 or 2^100 gives the same
  catalogue and tower, levels multiplied by exactly 4^e.
"""

import json
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "regressions.json"), encoding="utf-8") as fh:
    REG = json.load(fh)


def _hexf(s):
    return float.fromhex(s) if "0x" in s else float(s)


def _records(cat):
    out = []
    for i in range(len(cat.size)):
        m = int(cat.size[i])
        out.append((tuple(int(x) for x in cat.ids[i, :m]), cat.interior[i], int(cat.rank[i]),
                    bool(cat.degenerate[i])))
    return sorted(out)


def _oracle_records(canon, k):
    return sorted((r.support, r.interior, r.rank, r.degenerate)
                  for r in oracle.catalog(canon, k) if len(r.support) >= 2)


@pytest.mark.parametrize("variant", ["inside", "shell", "outside"])
def test_one_ulp_quasi_cospherical(variant):
    fx = REG["one_ulp_boundaries"]["quasi_cospherical"]
    var = next(v for v in fx["variants"] if v["name"] == variant)
    sup = [[_hexf(c) for c in p] for p in fx["support_points"]]
    mv = [_hexf(fx["moving_xy"][0]), _hexf(fx["moving_xy"][1]), _hexf(var["z"])]
    pts = np.asarray(sup + [mv])
    print(f"one-ulp cospherical {variant}", flush=True)
    canon = oracle.canonicalize(pts)
    ids = [int(np.flatnonzero((canon == np.asarray(p)).all(1))[0]) for p in sup]
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        inter, nin, nsh = ctx.closed_ball([sorted(ids)[:4]])
        cat = ctx.supports(REG["one_ulp_boundaries"]["k_max"])
    want = {"interior": (1, 4), "shell": (0, 5), "exterior": (0, 4)}[var["expected_relation"]]
    assert (int(nin[0]), int(nsh[0])) == want
    assert _records(cat) == _oracle_records(canon, REG["one_ulp_boundaries"]["k_max"])
    if variant == "shell":
        assert cat.info.diagnostic_count > 0


@pytest.mark.parametrize("variant", ["below", "exact", "above"])
def test_one_ulp_quasi_coplanar(variant):
    fx = REG["one_ulp_boundaries"]["quasi_coplanar"]
    var = next(v for v in fx["variants"] if v["name"] == variant)
    fixed = [[_hexf(c) for c in p] for p in fx["fixed_points"]]
    mv = [_hexf(fx["moving_xy"][0]), _hexf(fx["moving_xy"][1]), _hexf(var["z"])]
    pts = np.asarray(fixed + [mv])
    print(f"one-ulp coplanar {variant}", flush=True)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(REG["one_ulp_boundaries"]["k_max"])
    mine = _records(cat)
    assert mine == _oracle_records(canon, REG["one_ulp_boundaries"]["k_max"])
    has_tet = any(len(s) == 4 for s, _, _, _ in mine)
    if var["expected_affine_dimension"] == 2:
        assert not has_tet  # four coplanar points span no sphere of their own


def test_morse_rank_window():
    fx = REG["morse_rank_window"]
    pts = np.asarray([[float(Fraction(c)) for c in p] for p in fx["points"]])
    print("morse rank window", flush=True)
    canon = oracle.canonicalize(pts)
    src = [int(np.flatnonzero((canon == pts[i]).all(1))[0]) for i in range(len(pts))]
    exp = fx["expected"]
    sup = tuple(sorted(src[i] for i in exp["minimal_support_source_ids"]))
    inter = tuple(sorted(src[i] for i in exp["interior_source_ids"]))
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(4)
    recs = {(s, i): (r, d) for s, i, r, d in _records(cat)}
    assert recs[(sup, inter)] == (exp["closed_rank"], False)
    k = [j for j in range(len(cat.size)) if tuple(int(x) for x in cat.ids[j, :int(cat.size[j])]) == sup][0]
    assert cat.level[k] == float(Fraction(exp["squared_level"]))
    assert _records(cat) == _oracle_records(canon, 4)


def test_extra_shell_diagnostic_refuses_the_tower():
    fx = REG["relevant_gp_extra_shell_above_smax"]
    pts = np.asarray([[struct.unpack(">d", bytes.fromhex(w))[0] for w in p] for p in fx["points_binary64"]])
    print("extra-shell diagnostic", flush=True)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(2)  # the fixture's k_max
        assert cat.info.diagnostic_count >= 1
        deg = [(tuple(int(x) for x in cat.ids[j, :int(cat.size[j])])) for j in range(len(cat.size)) if cat.degenerate[j]]
        assert tuple(fx["candidate_ids"]) in deg
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.tower(2)
        assert err.value.code == mhgp.E_UNSUPPORTED
    assert _records(cat) == _oracle_records(canon, 2)


def _tower_sig(t):
    off = t["child_offset"]
    return [(int(t["order"][i]), int(t["kind"][i]), int(t["level_class"][i]),
             tuple(int(c) for c in t["children"][off[i]:off[i + 1]])) for i in range(len(t["order"]))]


@pytest.mark.parametrize("e", [-120, 100])
def test_scale_invariance(e):
    pts = workloads.jittered_lattice(110, side=5, seed=10)
    print(f"scale invariance 2^{e}", flush=True)
    out = []
    for scale in (0, e):
        with mhgp.Context(0) as ctx:
            ctx.upload(np.ldexp(pts, scale))
            cat = ctx.supports(10)
            t = ctx.tower(10)
            out.append((cat, t))
    (c0, t0), (c1, t1) = out
    assert np.array_equal(c0.ids, c1.ids) and np.array_equal(c0.level_class, c1.level_class)
    assert _records(c0) == _records(c1)
    assert np.array_equal(np.ldexp(c0.level, 2 * e), c1.level)
    assert _tower_sig(t0) == _tower_sig(t1)
    assert np.array_equal(np.ldexp(t0["level"], 2 * e), t1["level"])
