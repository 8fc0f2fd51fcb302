"""Exact Borůvka EMST on the GPU vs the oracle's complete-graph Kruskal."""

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["uniform500", "lattice2000", "grid_ties", "mixture1500", "square_tail"])
def test_emst_matches_kruskal(name):
    print(f"emst case {name}", flush=True)
    if name == "uniform500":
        pts = workloads.uniform_cloud(500, seed=11)
    elif name == "lattice2000":
        pts = workloads.jittered_lattice(2000, side=13, seed=12)
    elif name == "grid_ties":  # exact equal squared lengths everywhere
        g = np.stack(np.meshgrid(np.arange(7), np.arange(7), np.arange(7)), -1).reshape(-1, 3)
        pts = g.astype(float) * 0.5
    elif name == "mixture1500":
        pts = workloads.gaussian_mixture(1500, blobs=4, seed=13)
    else:
        pts = np.array([[-1, -1, 0], [-1, 1, 0], [1, -1, 0], [1, 1, 0], [5, 0, 0]], dtype=float)
    canon = oracle.canonicalize(pts)
    ref = oracle.emst(canon)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        e = ctx.emst().astype(np.int64)
    assert np.array_equal(e, ref)


def test_emst_50k_lattice_weight():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree

    pts = workloads.jittered_lattice(50_000, seed=0)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        xyz, _ = ctx.canonical_points()
        e = ctx.emst().astype(np.int64)
        nb = ctx.knn(32).astype(np.int64)
    n = len(xyz)
    assert len(e) == n - 1 and np.all(e[:, 0] < e[:, 1])
    w = np.sqrt(((xyz[e[:, 0]] - xyz[e[:, 1]]) ** 2).sum(1))
    assert np.all(np.diff(w ** 2) >= -1e-12)  # kappa order
    rows = np.repeat(np.arange(n), 32)
    d = np.sqrt(((xyz[rows] - xyz[nb.ravel()]) ** 2).sum(1))
    g = coo_matrix((d, (rows, nb.ravel())), shape=(n, n)).tocsr()
    mst = minimum_spanning_tree(g)
    assert abs(mst.sum() - w.sum()) <= 1e-9 * w.sum()
    # spanning: union-find over the edges
    parent = np.arange(n)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in e:
        ra, rb = find(a), find(b)
        assert ra != rb
        parent[ra] = rb
