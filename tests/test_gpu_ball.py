"""Closed-ball partition (mhgp_closed_ball) against the oracle's exact
sphere-side predicate, point by point (reference semantics:
lbvh_closed_ball, morsehgp3d/src/cpu/spatial/lbvh.cpp:1210-1340)."""

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def _expected(n, sup):
    inside, shell = [], len(sup)
    for x in range(n):
        if x in sup:
            continue
        sd = oracle.support_side(list(sup), x)
        if sd < 0:
            inside.append(x)
        elif sd == 0:
            shell += 1
    return inside, shell


def _check(pts, supports, cap=12):
    canon = oracle.canonicalize(pts)
    n = len(canon)
    oracle.catalog(canon, 1, 2)  # loads the cloud for support_side
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        inter, nin, nsh = ctx.closed_ball(supports, cap)
    for q, sup in enumerate(supports):
        inside, shell = _expected(n, sup)
        assert nin[q] == len(inside), (q, sup)
        assert nsh[q] == shell, (q, sup)
        # the cap smallest interior ids, whether or not the list is complete
        got = [int(x) for x in inter[q] if x != 0xFFFFFFFF]
        assert got == inside[:cap], (q, sup)


def _random_supports(rng, n, count):
    out = []
    for i in range(count):
        m = 1 + i % 4
        out.append([int(x) for x in rng.choice(n, size=m, replace=False)])
    return out


def test_closed_ball_uniform_random_supports():
    pts = workloads.uniform_cloud(180, seed=31)
    print("closed ball: uniform n=180, 48 random supports", flush=True)
    _check(pts, _random_supports(np.random.default_rng(1), 180, 48))


def test_closed_ball_catalogue_spheres():
    """Critical spheres of the catalogue: small interiors, listed in full."""
    pts = workloads.uniform_cloud(150, seed=32)
    print("closed ball: catalogue spheres n=150 k=4", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(4)
    sups = [[int(x) for x in cat.ids[i, : cat.size[i]]] for i in range(0, len(cat.size), max(1, len(cat.size) // 60))]
    _check(pts, sups)


def test_closed_ball_exact_shell_points():
    """Integer grid: many points exactly on the query spheres (host settlement)."""
    g = np.arange(4, dtype=np.float64)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    canon = oracle.canonicalize(pts)
    idx = {tuple(p): i for i, p in enumerate(canon.tolist())}
    sup = [
        [idx[(0.0, 0.0, 0.0)], idx[(2.0, 0.0, 0.0)]],                       # diametral, 4 more on it
        [idx[(0.0, 0.0, 0.0)], idx[(2.0, 0.0, 0.0)], idx[(0.0, 2.0, 0.0)]],  # square's circumcircle
        [idx[(0.0, 0.0, 0.0)], idx[(2.0, 0.0, 0.0)], idx[(0.0, 2.0, 0.0)], idx[(0.0, 0.0, 2.0)]],
        [idx[(1.0, 1.0, 1.0)]],
    ]
    print("closed ball: 4x4x4 grid, exact shells", flush=True)
    _check(pts, sup)


def test_closed_ball_truncated_lists_are_the_smallest_ids():
    """Interiors far above the cap on an integer grid, with more undecided
    shell points than the device's list holds (queries then move to the host
    whole): every list is the cap smallest interior ids, as the oracle's."""
    g = np.arange(7, dtype=np.float64)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(7)
    sups = []
    while len(sups) < 400:
        m = 2 + len(sups) % 3
        sup = [int(x) for x in rng.choice(len(pts), size=m, replace=False)]
        a = pts[sup[1:]] - pts[sup[0]]
        if m == 3 and not np.any(np.cross(a[0], a[1]) != 0):
            continue
        if m == 4 and abs(np.linalg.det(a)) < 0.5:
            continue
        sups.append(sup)
    print("closed ball: 7x7x7 grid, 400 supports, cap 5", flush=True)
    _check(pts, sups, cap=5)
