"""Double-double level evaluation (e-hgp_amd/csrc/level_dd.hpp), built for
the host: every interval contains the exact squared circumradius, computed
with Python fractions from the same binary64 coordinates, and the certified
order of two levels never contradicts the exact order."""

import ctypes
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from helpers import PKG

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ddlib(tmp_path_factory):
    out = tmp_path_factory.mktemp("dd") / "libddcheck.so"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-fPIC", "-shared", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-x", "c++", "-I", os.path.join(PKG, "csrc"), "-I/opt/rocm/include",
           os.path.join(HERE, "native", "dd_check.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    lib = ctypes.CDLL(str(out))
    p = ctypes.c_void_p
    lib.dd_level_check.argtypes = [p, p, p, p, ctypes.c_int, p]
    lib.dd_order_check.argtypes = [p, p]
    return lib


def _level(lib, pts, ids):
    X = np.ascontiguousarray(pts[:, 0])
    Y = np.ascontiguousarray(pts[:, 1])
    Z = np.ascontiguousarray(pts[:, 2])
    i = np.asarray(ids, dtype=np.uint32)
    out = np.zeros(3)
    lib.dd_level_check(X.ctypes.data, Y.ctypes.data, Z.ctypes.data, i.ctypes.data, len(ids), out.ctypes.data)
    return out


def _exact(pts, ids):
    P = [[Fraction(float(c)) for c in pts[j]] for j in ids]
    u = P[0]
    vs = [[p[k] - u[k] for k in range(3)] for p in P[1:]]
    dot = lambda a, b: sum(x * y for x, y in zip(a, b))
    if len(vs) == 1:
        return dot(vs[0], vs[0]) / 4
    if len(vs) == 2:
        a, b = vs
        c = [b[k] - a[k] for k in range(3)]
        D = dot(a, a) * dot(b, b) - dot(a, b) ** 2
        return dot(a, a) * dot(b, b) * dot(c, c) / (4 * D)
    a, b, d = vs
    cross = lambda x, y: [x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]]
    bxd, dxa, axb = cross(b, d), cross(d, a), cross(a, b)
    V = dot(a, bxd)
    N = [dot(a, a) * bxd[k] + dot(b, b) * dxa[k] + dot(d, d) * axb[k] for k in range(3)]
    return dot(N, N) / (4 * V * V)


def _contains(l, exact):
    hi, lo, err = (Fraction(float(x)) for x in l)
    return abs(exact - (hi + lo)) <= err


@pytest.mark.parametrize("m", [2, 3, 4])
def test_interval_contains_exact_level(ddlib, m):
    rng = np.random.default_rng(m)
    for trial in range(300):
        pts = rng.random((m, 3)) * (10.0 ** rng.integers(-3, 4)) + rng.normal(size=3) * 50
        if trial % 3 == 0 and m >= 3:
            # nearly dependent: last point close to the affine hull of the others
            w = rng.random(m - 1)
            w /= w.sum()
            pts[-1] = w @ pts[:-1] + rng.normal(size=3) * 1e-7
        l = _level(ddlib, pts, list(range(m)))
        if not np.isfinite(l[2]):
            continue
        assert _contains(l, _exact(pts, list(range(m)))), (trial, pts.tolist(), l.tolist())
        assert l[2] <= abs(l[0]) * 1e-20 or trial % 3 == 0


def test_order_certifies_only_true_order(ddlib):
    rng = np.random.default_rng(5)
    checked = 0
    for trial in range(200):
        base = rng.random((4, 3)) * 4
        pts = np.concatenate([base, base + np.array([7.0, 0.0, 0.0])])
        # second copy translated by an exact power-of-two-friendly offset: equal
        # exact levels for the same support shape; nudge one coordinate by k ulp
        k = trial % 3
        pts[7, 2] = np.nextafter(pts[7, 2], np.inf) if k == 1 else pts[7, 2]
        for m in (2, 3, 4):
            a, b = list(range(m)), list(range(4, 4 + m))
            la, lb = _level(ddlib, pts, a), _level(ddlib, pts, b)
            ea, eb = _exact(pts, a), _exact(pts, b)
            o = ddlib.dd_order_check(la.ctypes.data, lb.ctypes.data)
            if ea < eb:
                assert o in (0, 1)
            elif ea > eb:
                assert o in (0, -1)
            else:
                assert o == 0
            if ea != eb and o != 0:
                checked += 1
    assert checked > 50
