"""Canonical cloud, exact Morton LBVH snapshot and k-NN on the GPU vs the oracle."""

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def _clouds():
    rng = np.random.default_rng(123)
    yield "single", np.array([[0.5, -0.0, 2.0]])
    yield "pair", np.array([[1.0, 2.0, 3.0], [-1.0, 2.0, 3.0]])
    yield "uniform17", rng.random((17, 3))
    yield "uniform1000", workloads.uniform_cloud(1000, seed=3)
    yield "lattice3000", workloads.jittered_lattice(3000, side=16, seed=4)
    grid = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(6)), -1).reshape(-1, 3) * 0.25
    yield "exact_grid216", grid.astype(float)
    # runs of equal Morton codes (bin width 2^-21 of the extent): midpoint splits
    run = [[i * 1e-9, (i % 3) * 1e-9, 0.0] for i in range(7)]
    run += [[0.5 + i * 1e-9, 0.5, 0.5] for i in range(6)]
    tiny = np.array(run + [[1.0, 1.0, 1.0], [0.25, 0.75, 0.1]], dtype=float)
    yield "equal_codes", tiny
    yield "mixture2000", workloads.gaussian_mixture(2000, blobs=5, seed=5)


@pytest.mark.parametrize("name,pts", list(_clouds()), ids=lambda v: v if isinstance(v, str) else "")
def test_canonical_and_snapshot_match_oracle(name, pts):
    print(f"case {name} n={len(pts)}", flush=True)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        xyz, src = ctx.canonical_points()
        assert np.array_equal(xyz, canon)
        assert np.array_equal(np.asarray(pts)[src] + 0.0, canon)
        leaves, nodes, counters = ctx.lbvh_snapshot()
        ref = oracle.lbvh_snapshot(canon)
        assert np.array_equal(leaves, ref.leaves)
        assert np.array_equal(nodes, ref.nodes)
        assert np.array_equal(counters, ref.counters)


def test_knn_matches_oracle():
    for name, pts in _clouds():
        if len(pts) < 12 or len(pts) > 3000:
            continue
        print(f"knn case {name}", flush=True)
        canon = oracle.canonicalize(pts)
        with mhgp.Context(0) as ctx:
            ctx.upload(pts)
            for k in (1, 11, 32):
                if k >= len(pts):
                    continue
                ids = ctx.knn(k)
                ref, _ = oracle.knn(canon, k)
                assert np.array_equal(ids.astype(np.int64), ref), (name, k)


def test_rejects_duplicates_and_nonfinite():
    with mhgp.Context(0) as ctx:
        with pytest.raises(mhgp.MhgpError) as e:
            ctx.upload([[0, 0, 0], [1, 1, 1], [0, -0.0, 0]])
        assert e.value.code == mhgp.E_DUPLICATE
        with pytest.raises(mhgp.MhgpError) as e:
            ctx.upload([[0, 0, 0], [np.nan, 1, 1]])
        assert e.value.code == mhgp.E_NONFINITE
        ctx.upload([[0, 0, 0], [1, 1, 1]])  # context stays usable after argument errors
        assert ctx.n == 2


def test_lattice_50k_snapshot_invariants():
    pts = workloads.jittered_lattice(50_000, seed=0)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        leaves, nodes, counters = ctx.lbvh_snapshot()
    ref = oracle.lbvh_snapshot(canon)
    assert np.array_equal(leaves, ref.leaves)
    assert np.array_equal(nodes, ref.nodes)
    # published snapshot size of the reference at n = 50 000: 176 n - 80 bytes
    assert leaves.nbytes + nodes.nbytes == 176 * 50_000 - 80


def test_device_snapshot_is_certified_exactly():
    """The device snapshot passes the exact host re-certification (the
    reference's import_certified_snapshot role) at 50 000 points and on a
    cloud with equal-code runs."""
    for label, pts in (("lattice50k", workloads.jittered_lattice(50_000, side=37, seed=0)),
                       ("collisions", np.concatenate([workloads.uniform_cloud(200, seed=3),
                                                      0.5 + np.arange(40)[:, None] * 1e-13 * np.ones((1, 3))]))):
        print(f"certify device snapshot {label}", flush=True)
        with mhgp.Context(0) as ctx:
            ctx.upload(pts)
            xyz, _ = ctx.canonical_points()
            leaves, nodes, counters = ctx.lbvh_snapshot()
        mhgp.lbvh_certify(xyz, leaves, nodes, counters)
