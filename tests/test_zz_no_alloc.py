"""mhgp_closed_ball_noalloc (no device allocation, fixed device scratch)
against mhgp_closed_ball on the same queries: identical interior ids,
interior counts and shell counts, on two calls in a row with the second one
smaller (the scratch keeps the first call's contents), then a call beyond the
scratch, which must be refused without harming the context."""

import numpy as np
import pytest

from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def _independent(pts, sup):
    if len(sup) < 3:
        return True
    a = pts[sup[1:]] - pts[sup[0]]
    if len(sup) == 3:
        return bool(np.any(np.cross(a[0], a[1]) != 0))
    return abs(np.linalg.det(a)) > 1e-9


def _supports(rng, pts, count):
    out = []
    while len(out) < count:
        m = 1 + len(out) % 4
        sup = [int(x) for x in rng.choice(len(pts), size=m, replace=False)]
        if _independent(pts, sup):  # the grid holds collinear and coplanar sets
            out.append(sup)
    return out


def _same(ctx, supports, cap):
    a = ctx.closed_ball(supports, cap)
    b = ctx.closed_ball(supports, cap, noalloc=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    return a


def test_noalloc_matches_closed_ball_twice_second_smaller():
    rng = np.random.default_rng(5)
    # integer grid: many spheres pass exactly through further grid points,
    # which sends points to the exact host settlement
    g = np.stack(np.meshgrid(np.arange(9), np.arange(9), np.arange(9), indexing="ij"), -1).reshape(-1, 3)
    grid = g.astype(np.float64)
    print("noalloc: 729-point integer grid, 3000 then 700 queries", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(grid)
        first = _supports(rng, grid, 3000)
        _, nin, nsh = _same(ctx, first, 12)
        assert (nsh > [len(s) for s in first]).any(), "no exact shell point was exercised"
        _same(ctx, _supports(rng, grid, 700), 12)
    pts = workloads.uniform_cloud(2000, seed=44)
    print("noalloc: uniform n=2000, 5000 then 64 queries, cap 0 and 16", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        _same(ctx, _supports(rng, pts, 5000), 16)
        _same(ctx, _supports(rng, pts, 64), 0)


def test_noalloc_refuses_beyond_scratch_and_context_survives():
    limit = mhgp.lib().mhgp_closed_ball_noalloc_scratch_bytes()
    assert limit > 0
    pts = workloads.uniform_cloud(300, seed=45)
    rng = np.random.default_rng(6)
    too_many = limit // 64 + 1  # each query needs more than 64 bytes
    print(f"noalloc: {too_many} queries against {limit} bytes of scratch", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        big = [[int(rng.integers(300))] for _ in range(too_many)]
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.closed_ball(big, 4, noalloc=True)
        assert err.value.code == mhgp.E_CAPACITY
        _same(ctx, _supports(rng, pts, 200), 8)
