"""The reference's predicate fixtures against the product's predicate layers.

Inputs: morsehgp3d/tests/fixtures/predicates/*.json (binary64 words, copied
as data into tests/golden/predicates.json).  The expected outcome of each
case is derived twice, independently of the product: in exact rationals
(Python ``fractions``) from the closed forms of the reference's definitions
(circumcentre in the affine hull and barycentric well-centring,
exact/center.hpp:263-340 and exact/support.hpp:41-135; closed sphere side,
support.hpp:814; exact level order, level_order.hpp:110), and from the
outcome the fixture's file name states (inside / outside / boundary,
positive / negative / zero, less / greater / equal, obtuse, coplanar ...),
which must agree with the rationals.  Then:

* the exact host path (host/exact.cpp) must return that outcome, always;
* the certified FP64 filters (csrc/predicates.hpp) and the double-double
  levels (csrc/level_dd.hpp) must return it or "undecided", never the
  opposite (one-ulp, subnormal and cancellation cases included).

Fixture kinds outside this backend's path (plane intersections, affine
incidences, power bisectors, query barycentrics) are not replayed.
"""

import ctypes
import json
import os
import struct
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from helpers import PKG

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "predicates.json"), encoding="utf-8") as fh:
    FIX = json.load(fh)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("exact") / "libexactcheck.so"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-fPIC", "-shared", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-x", "c++", "-I", os.path.join(PKG, "csrc"), "-I/opt/rocm/include",
           os.path.join(HERE, "native", "exact_check.cpp"), os.path.join(PKG, "host", "exact.cpp"),
           "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    L = ctypes.CDLL(str(out))
    p, i, s = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
    L.ex_analyze.argtypes = [p, i, p, p]
    L.ex_level_vs.argtypes = [p, i, s, s]
    L.ex_side.argtypes = [p, i]
    L.ex_level_cmp.argtypes = [p, i, i]
    L.ex_compare_rationals.argtypes = [s, s, s, s]
    L.ex_cmp_dist.argtypes = [p]
    L.f64_well.argtypes = [p, i]
    L.f64_side.argtypes = [p, i]
    L.dd_level_order.argtypes = [p, i, i]
    return L


def _f(word):
    return struct.unpack(">d", bytes.fromhex(word))[0]


def _arr(points):
    return np.ascontiguousarray([[_f(w) for w in p] for p in points], dtype=np.float64)


def _Q(points):
    return [[Fraction(_f(w)) for w in p] for p in points]


def _sub(a, b):
    return [x - y for x, y in zip(a, b)]


def _dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def _solve(G, r):
    """Exact Gaussian elimination; None when G is singular."""
    n = len(r)
    M = [list(G[i]) + [r[i]] for i in range(n)]
    for c in range(n):
        piv = next((k for k in range(c, n) if M[k][c] != 0), None)
        if piv is None:
            return None
        M[c], M[piv] = M[piv], M[c]
        for k in range(n):
            if k != c and M[k][c] != 0:
                f = M[k][c] / M[c][c]
                M[k] = [x - f * y for x, y in zip(M[k], M[c])]
    return [M[k][n] / M[k][k] for k in range(n)]


def circumsphere(P):
    """(dependent, centre, squared radius, barycentric coordinates) of the
    circumsphere of P with centre in the affine hull of P."""
    if len(P) == 1:
        return False, P[0], Fraction(0), [Fraction(1)]
    v = [_sub(p, P[0]) for p in P[1:]]
    G = [[_dot(a, b) for b in v] for a in v]
    lam = _solve(G, [_dot(a, a) / 2 for a in v])
    if lam is None:
        return True, None, None, None
    c = [P[0][k] + sum(l * a[k] for l, a in zip(lam, v)) for k in range(3)]
    return False, c, _dot(_sub(c, P[0]), _sub(c, P[0])), [1 - sum(lam)] + lam


def _named(name, table):
    hits = {out for key, out in table.items() if key in name}
    assert len(hits) <= 1, name
    return hits.pop() if hits else None


ANALYSIS = [k for k, v in FIX.items()
            if v.get("predicate") in ("circumcenter_support", "circumcenter_support_analysis",
                                      "binary64_circumcenter_support_analysis")]


@pytest.mark.parametrize("name", ANALYSIS)
def test_support_analysis(lib, name):
    pts = FIX[name]["points"]
    m = len(pts)
    dep, c, level, bary = circumsphere(_Q(pts))
    well = (not dep) and all(b > 0 for b in bary)
    # what the file name promises
    want = _named(name, {"coplanar": "dep", "collinear": "dep", "duplicate": "dep", "dependent": "dep",
                         "obtuse": "out", "exterior": "out", "negative": "out",
                         "right": "edge", "boundary": "edge",
                         "acute": "in", "well_centered": "in", "positive": "in"})
    if want == "dep":
        assert dep, name
    elif want == "out":
        assert not dep and min(bary) < 0, name
    elif want == "edge":
        assert not dep and min(bary) == 0, name
    elif want == "in":
        assert well, name
    a = _arr(pts)
    d, w = ctypes.c_int(), ctypes.c_int()
    orient = lib.ex_analyze(a.ctypes.data, m, ctypes.byref(d), ctypes.byref(w))
    if m == 2 and dep:  # a duplicate pair: the catalogue never forms one (uploads reject duplicates)
        return
    assert bool(d.value) == dep, name
    assert bool(w.value) == well, name
    if not dep:
        assert lib.ex_level_vs(a.ctypes.data, m, str(level.numerator).encode(),
                               str(level.denominator).encode()) == 0, name
        fw = lib.f64_well(a.ctypes.data, m)
        assert fw == 0 or (fw > 0) == well, name
    if m == 4 and not dep:
        v = [_sub(p, _Q(pts)[0]) for p in _Q(pts)[1:]]
        det = _dot(v[0], [v[1][1] * v[2][2] - v[1][2] * v[2][1], v[1][2] * v[2][0] - v[1][0] * v[2][2],
                          v[1][0] * v[2][1] - v[1][1] * v[2][0]])
        assert orient == (det > 0) - (det < 0), name


ORIENT = [k for k, v in FIX.items() if v.get("predicate") == "orientation_3d"]


@pytest.mark.parametrize("name", ORIENT)
def test_orientation(lib, name):
    P = _Q(FIX[name]["points"])
    v = [_sub(p, P[0]) for p in P[1:]]
    det = _dot(v[0], [v[1][1] * v[2][2] - v[1][2] * v[2][1], v[1][2] * v[2][0] - v[1][0] * v[2][2],
                      v[1][0] * v[2][1] - v[1][1] * v[2][0]])
    a = _arr(FIX[name]["points"])
    d, w = ctypes.c_int(), ctypes.c_int()
    orient = lib.ex_analyze(a.ctypes.data, 4, ctypes.byref(d), ctypes.byref(w))
    assert orient == (det > 0) - (det < 0) and bool(d.value) == (det == 0)
    # the tetrahedron filter: certified or undecided, never wrong
    fw = lib.f64_well(a.ctypes.data, 4)
    if det == 0:
        assert fw <= 0


SIDE = [k for k, v in FIX.items() if v.get("predicate") == "support_sphere_side"]


@pytest.mark.parametrize("name", SIDE)
def test_support_sphere_side(lib, name):
    sup = FIX[name]["support"]
    q = FIX[name]["point"]
    dep, c, level, _ = circumsphere(_Q(sup))
    assert not dep
    dist = _dot(_sub(_Q([q])[0], c), _sub(_Q([q])[0], c))
    exact = (dist > level) - (dist < level)
    want = _named(name, {"inside": -1, "outside": 1, "boundary": 0})
    assert want is None or want == exact, name
    a = _arr(list(sup) + [q])
    assert lib.ex_side(a.ctypes.data, len(sup)) == exact, name
    f = lib.f64_side(a.ctypes.data, len(sup))
    assert f in (0, 9) or f == exact, name


RATIONAL_SPHERE = [k for k, v in FIX.items() if v.get("predicate") == "sphere_side"]


@pytest.mark.parametrize("name", RATIONAL_SPHERE)
def test_rational_sphere_side(lib, name):
    # |p - c|^2 against the exact level, compared by the product's rational order
    fx = FIX[name]
    ce = fx["center_exact"]
    den = int(ce["denominator"])
    c = [Fraction(int(ce[k + "_numerator"]), den) for k in "xyz"]
    L = Fraction(int(fx["squared_level_exact"]["numerator"]), int(fx["squared_level_exact"]["denominator"]))
    p = _Q([fx["point"]])[0]
    dist = _dot(_sub(p, c), _sub(p, c))
    exact = (dist > L) - (dist < L)
    want = _named(name, {"inside": -1, "outside": 1, "boundary": 0})
    assert want is None or want == exact, name
    got = lib.ex_compare_rationals(str(dist.numerator).encode(), str(dist.denominator).encode(),
                                   str(L.numerator).encode(), str(L.denominator).encode())
    assert got == exact, name


LEVELS = [k for k, v in FIX.items() if v.get("predicate") == "compare_support_levels"]


@pytest.mark.parametrize("name", LEVELS)
def test_compare_support_levels(lib, name):
    left, right = FIX[name]["left_support"], FIX[name]["right_support"]
    _, _, l1, _ = circumsphere(_Q(left))
    _, _, l2, _ = circumsphere(_Q(right))
    exact = (l1 > l2) - (l1 < l2)
    want = _named(name, {"less": -1, "greater": 1, "equal": 0})
    assert want is None or want == exact, name
    a = _arr(list(left) + list(right))
    assert lib.ex_level_cmp(a.ctypes.data, len(left), len(right)) == exact, name
    dd = lib.dd_level_order(a.ctypes.data, len(left), len(right))  # +1: left < right
    assert dd == 0 or dd == -exact, name


EXACT_LEVELS = [k for k, v in FIX.items() if v.get("predicate") == "compare_exact_levels"]


@pytest.mark.parametrize("name", EXACT_LEVELS)
def test_compare_exact_levels(lib, name):
    lft, rgt = FIX[name]["left_squared_level_exact"], FIX[name]["right_squared_level_exact"]
    l1 = Fraction(int(lft["numerator"]), int(lft["denominator"]))
    l2 = Fraction(int(rgt["numerator"]), int(rgt["denominator"]))
    exact = (l1 > l2) - (l1 < l2)
    want = _named(name, {"minus_one": -1, "plus_one": 1, "equal": 0})
    assert want is None or want == exact, name
    got = lib.ex_compare_rationals(lft["numerator"].encode(), lft["denominator"].encode(),
                                   rgt["numerator"].encode(), rgt["denominator"].encode())
    assert got == exact, name


DIST = [k for k, v in FIX.items() if v.get("predicate") == "compare_squared_distances"]


@pytest.mark.parametrize("name", DIST)
def test_compare_squared_distances(lib, name):
    q, a, b = _Q(FIX[name]["points"])
    da, db = _dot(_sub(q, a), _sub(q, a)), _dot(_sub(q, b), _sub(q, b))
    exact = (da > db) - (da < db)
    if "equal" in name:
        assert exact == 0
    arr = _arr(FIX[name]["points"])
    assert lib.ex_cmp_dist(arr.ctypes.data) == exact, name


def test_every_replayed_kind_has_cases():
    assert len(ANALYSIS) >= 30 and len(SIDE) >= 8 and len(LEVELS) >= 8 and len(EXACT_LEVELS) >= 5
    assert len(RATIONAL_SPHERE) >= 7 and len(DIST) >= 3 and len(ORIENT) >= 3
