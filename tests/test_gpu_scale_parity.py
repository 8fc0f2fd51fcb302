"""Parity at the headline scale: the 50 000-point jittered lattice
(37^3 sites, sigma = 0.1 spacing) at K_max = 10.

* Catalogue: for 200 seeded random anchors, the device records whose least
  PointId is the anchor equal the oracle's exhaustive exact enumeration among
  the points within the anchor's certified enumeration radius (supports,
  interiors, ranks, degeneracy; FP64 levels within 1e-12 of the exact
  rational; level classes ordered and tied exactly as the exact levels).
* Catalogue, whole cloud: the Euler identity sum_k (-1)^index multiplicity = 1
  of every complete order window k = 1..8.
* EMST: the device tree equals the exact kappa-Kruskal forest of the 32-nearest
  -neighbour graph (scipy's k-d tree supplies the candidate edges).
"""

from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction
from math import comb

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu

N, SIDE, K = 50_000, 37, 10


@pytest.fixture(scope="module")
def lattice():
    pts = workloads.jittered_lattice(N, side=SIDE, sigma=0.1, seed=0)
    print("scale parity: jittered lattice n=50000 k=10, catalogue", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        edges = ctx.emst()
        cat = ctx.supports_raw(K)
        radii = ctx.radii()
    return pts, oracle.canonicalize(pts), edges, cat, radii


def test_catalogue_anchors_match_exhaustive_oracle(lattice):
    pts, canon, _, cat, radii = lattice
    ids, meta = cat["ids"], cat["meta"]
    size, rank = meta & 0xFF, (meta >> 8) & 0xFF
    nint, deg = (meta >> 16) & 0xFF, (meta >> 24) & 0xFF
    anchors = np.random.default_rng(2024).choice(N, size=200, replace=False)
    print(f"scale parity: {len(anchors)} anchors against the exhaustive oracle", flush=True)
    oracle.anchor_records(canon, K, int(anchors[0]), float(radii[anchors[0]]))  # loads the cloud
    with ThreadPoolExecutor(max_workers=16) as pool:
        refs = list(pool.map(lambda a: oracle.anchor_records(canon, K, int(a), float(radii[a])), anchors))
    by_anchor = {}
    for i in np.nonzero(np.isin(ids[:, 0], anchors))[0]:
        by_anchor.setdefault(int(ids[i, 0]), []).append(int(i))
    classes, exact_levels, total = [], [], 0
    for a, ref in zip(anchors, refs):
        a = int(a)
        rows = by_anchor.get(a, [])
        mine = {}
        for i in rows:
            m = int(size[i])
            key = (tuple(int(x) for x in ids[i, :m]), tuple(int(x) for x in cat["interior"][i, : nint[i]]),
                   int(rank[i]), bool(deg[i]))
            mine[key] = i
        theirs = {(r.support, r.interior, r.rank, r.degenerate): r.level for r in ref}
        assert set(mine) == set(theirs), f"anchor {a}"
        for key, i in mine.items():
            lev = theirs[key]
            # every record of the anchor lies inside its certified radius
            assert lev <= Fraction(float(radii[a])) ** 2 / 4, f"anchor {a}"
            assert abs(cat["level"][i] - float(lev)) <= 1e-12 * float(lev), f"anchor {a} {key}"
            classes.append(int(cat["level_class"][i]))
            exact_levels.append(lev)
        total += len(mine)
    assert total > 200 * 300, total
    # level classes: sorted by class, exact levels ascend, ties exactly within a class
    order = np.argsort(np.asarray(classes), kind="stable")
    for x, y in zip(order[:-1], order[1:]):
        if classes[x] == classes[y]:
            assert exact_levels[x] == exact_levels[y]
        else:
            assert exact_levels[x] < exact_levels[y]
    print(f"scale parity: {total} records of {len(anchors)} anchors equal", flush=True)


def test_catalogue_euler_identity_complete_windows(lattice):
    _, _, _, cat, _ = lattice
    meta = cat["meta"]
    m = (meta & 0xFF).astype(np.int64)
    s = ((meta >> 8) & 0xFF).astype(np.int64)
    deg = (meta >> 24) & 0xFF
    assert not deg.any(), "the jittered lattice has no extra-shell degeneracy"
    nint = s - m
    binom = np.array([[comb(a, b) for b in range(4)] for a in range(4)], dtype=np.int64)
    for k in range(1, 9):  # complete windows: ranks k..k+3 <= s_max = 11
        sel = (nint < k) & (k <= s)
        idx = s[sel] - k
        total = int(np.sum((-1) ** idx * binom[m[sel] - 1, idx]))
        total += N if k == 1 else 0
        assert total == 1, (k, total)


def test_emst_equals_exact_kruskal_of_knn_graph(lattice):
    from scipy.spatial import cKDTree

    _, canon, edges, _, _ = lattice
    _, nb = cKDTree(canon).query(canon, k=33)
    cand = np.stack([np.repeat(np.arange(N), 32), nb[:, 1:].ravel()], 1)
    ref = oracle.emst_graph(canon, cand)
    assert len(ref) == N - 1, "the 32-NN graph is connected"
    assert np.array_equal(edges.astype(np.int64), ref)
