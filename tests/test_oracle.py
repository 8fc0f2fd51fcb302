"""Pins the CPU oracle against the reference's own fixtures and closed forms."""

import json
import struct
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import workloads
from oracle import tower

HERE = os.path.dirname(os.path.abspath(__file__))

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name), encoding="utf-8") as fh:
        return json.load(fh)


def _frac(pair):
    return Fraction(int(pair[0]), int(pair[1]))


@pytest.mark.parametrize("case", _load("contracts.json"), ids=lambda c: c["name"])
def test_catalog_matches_contract_fixture(case):
    pts = oracle.canonicalize(case["points"])
    assert np.array_equal(pts, np.asarray(case["points"]))  # fixtures are stored canonically
    cat = oracle.catalog(pts, case["k_max"])
    assert not any(r.degenerate for r in cat)
    mine = sorted((r.support, r.interior, r.rank, r.level) for r in cat)
    ref = sorted(
        (tuple(e["support"]), tuple(e["interior"]), e["rank"], _frac(e["level"]))
        for e in case["catalog"]
    )
    assert mine == ref
    for e in case["catalog"]:
        assert e["shell"] == e["support"]
        assert e["birth_order"] == (e["rank"] if e["rank"] <= case["k_eff"] else None)


@pytest.mark.parametrize("case", _load("contracts.json"), ids=lambda c: c["name"])
def test_gamma_forests_match_contract_fixture(case):
    pts = oracle.canonicalize(case["points"])
    forests = tower.gamma_forests(pts, case["k_max"], case["profile"])
    for f in case["forests"]:
        ref = [
            {"order": f["order"], "kind": n["kind"], "level": _frac(n["level"]),
             "children": tuple(n["children"]), "covered": tuple(n["covered"])}
            for n in f["nodes"]
        ]
        assert tower.signatures(ref, True) == tower.signatures(forests[f["order"]], True)


@pytest.mark.parametrize("case", _load("selected_n12_n14.json"), ids=lambda c: c["name"])
def test_catalog_event_count_selected_clouds(case):
    pts = oracle.canonicalize(np.asarray(case["points"], dtype=float))
    cat = oracle.catalog(pts, case["k_max"])
    events = [r for r in cat if not r.degenerate]
    assert len(events) == case["expected_event_count"]
    assert (len(events) == len(cat)) == case["relevant_gp_complete"]
    assert max(r.rank for r in events) <= case["s_max"]


def test_direct_tower_equals_exhaustive_gamma():
    rng = np.random.default_rng(7)
    for _ in range(6):
        n = int(rng.integers(5, 9))
        k = int(rng.integers(1, n + 1))
        pts = oracle.canonicalize(rng.integers(-40, 40, size=(n, 3)) + rng.random((n, 3)) * 1e-3)
        g = tower.gamma_forests(pts, k, "hgp_reduced")
        d = tower.direct_forests(pts, k)
        for order in range(1, min(k, n) + 1):
            assert tower.signatures(d[order]) == tower.signatures(g[order])


def test_regular_tetrahedron_supports():
    # reference test_hierarchy_higher_support_stream.cpp:205-251: 6 pairs,
    # 4 triangles and 1 tetrahedron are the non-trivial supports.
    pts = oracle.canonicalize([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]])
    cat = oracle.catalog(pts, 10)
    sizes = sorted(len(r.support) for r in cat)
    assert sizes == [1] * 4 + [2] * 6 + [3] * 4 + [4]
    tet = [r for r in cat if len(r.support) == 4][0]
    assert tet.level == 3 and tet.rank == 4


def test_emst_known_answers():
    # test_hierarchy_emst.cpp:231-258: d^2 = 3 edge
    assert oracle.emst([[0, 0, 0], [1, 1, 1]]).tolist() == [[0, 1]]
    # test_hierarchy_boruvka.cpp:443-511: square with tail, kappa tie-breaks
    pts = oracle.canonicalize([[-1, -1, 0], [-1, 1, 0], [1, -1, 0], [1, 1, 0], [5, 0, 0]])
    assert oracle.emst(pts).tolist() == [[0, 1], [0, 2], [1, 3], [2, 4]]


def test_emst_weight_matches_prim():
    rng = np.random.default_rng(3)
    pts = oracle.canonicalize(rng.random((150, 3)))
    e = oracle.emst(pts)
    w = float(np.sum(np.sum((pts[e[:, 0]] - pts[e[:, 1]]) ** 2, axis=1) ** 0.5))
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    inn = np.zeros(len(pts), bool); inn[0] = True
    best = d[0].copy(); tot = 0.0
    for _ in range(len(pts) - 1):
        best[inn] = np.inf
        j = int(np.argmin(best)); tot += best[j]; inn[j] = True
        best = np.minimum(best, d[j])
    assert abs(tot - w) < 1e-9 * tot


def test_knn_matches_numpy():
    rng = np.random.default_rng(5)
    pts = oracle.canonicalize(rng.integers(0, 9, size=(120, 3)) + rng.random((120, 3)) * 1e-2)
    ids, ties = oracle.knn(pts, 11)
    d = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    order = np.lexsort((np.broadcast_to(np.arange(len(pts)), d.shape), d), axis=1)[:, :11]
    assert np.array_equal(ids, order)
    assert not ties.any()


def test_lbvh_snapshot_structure():
    rng = np.random.default_rng(11)
    pts = oracle.canonicalize(rng.normal(size=(300, 3)))
    snap = oracle.lbvh_snapshot(pts)
    n = len(pts)
    assert snap.counters[0] == n and snap.counters[1] == 2 * n - 1
    codes = snap.leaves[:, 0]
    assert np.all((codes[1:] > codes[:-1]) | ((codes[1:] == codes[:-1]) & (snap.leaves[1:, 1] > snap.leaves[:-1, 1])))
    assert sorted(snap.leaves[:, 1].tolist()) == list(range(n))
    nodes = snap.nodes
    root = nodes[-1]
    assert root[8] == 0 and root[9] == n
    inv = np.uint64(2**64 - 1)
    for i, nd in enumerate(nodes):
        if nd[6] == inv:
            assert nd[9] - nd[8] == 1
            pid = snap.leaves[nd[8], 1]
            assert all(nd[a] == pid for a in range(6))
            continue
        l, r = nodes[nd[6]], nodes[nd[7]]
        assert nd[6] < i and nd[7] < i  # postorder
        assert l[8] == nd[8] and l[9] == r[8] and r[9] == nd[9]
        lo = pts[list(range(n))]
        for a in range(3):
            ids = snap.leaves[nd[8]:nd[9], 1].astype(int)
            vals = pts[ids, a]
            assert pts[int(nd[a]), a] == vals.min() and pts[int(nd[3 + a]), a] == vals.max()
            assert int(nd[a]) == int(ids[vals == vals.min()].min())
            assert int(nd[3 + a]) == int(ids[vals == vals.max()].min())


def test_morton_quantisation_closed_form():
    # points on an exact dyadic grid: bins are x * 2^21 / extent
    pts = oracle.canonicalize([[0, 0, 0], [1, 2, 4], [0.5, 0.25, 1], [0.75, 1.5, 3]])
    snap = oracle.lbvh_snapshot(pts)
    def inter(q):
        c = 0
        for b in range(21):
            c |= ((q[0] >> b) & 1) << (3 * b + 2) | ((q[1] >> b) & 1) << (3 * b + 1) | ((q[2] >> b) & 1) << (3 * b)
        return c
    for pid, p in enumerate(pts):
        q = [min(int(p[a] / [1, 2, 4][a] * 2**21), 2**21 - 1) for a in range(3)]
        assert int(snap.codes_by_id[pid]) == inter(q)


@pytest.mark.parametrize("seed,n", [(3, 40), (4, 56)])
def test_anchor_records_equal_exhaustive_catalogue(seed, n):
    """The anchored enumeration (used at scale) against the exhaustive
    catalogue: every record of every anchor, with a radius covering the cloud,
    at K = 3 and K = 10; with a small radius it can only lose records whose
    circumradius exceeds half of it, never invent or misclassify one."""
    pts = oracle.canonicalize(workloads.uniform_cloud(n, seed=seed))
    for k in (3, 10):
        ref = [r for r in oracle.catalog(pts, k) if len(r.support) >= 2]
        full = {(r.support, r.interior, r.rank, r.degenerate, r.level) for r in ref}
        got = set()
        for a in range(n):
            got |= {(r.support, r.interior, r.rank, r.degenerate, r.level)
                    for r in oracle.anchor_records(pts, k, a, 4.0)}
        assert got == full
        part = set()
        for a in range(n):
            part |= {(r.support, r.interior, r.rank, r.degenerate, r.level)
                     for r in oracle.anchor_records(pts, k, a, 0.5)}
        assert part <= full
        assert part == {r for r in full if r[4] <= Fraction(1, 16)}


def test_anchor_records_planted_mistake_is_caught():
    """A planted wrong interior (one record's interior id changed) is seen by
    the comparison the scale-parity test makes."""
    pts = oracle.canonicalize(workloads.uniform_cloud(30, seed=8))
    recs = oracle.anchor_records(pts, 10, 0, 4.0)
    ref = [r for r in oracle.catalog(pts, 10) if len(r.support) >= 2 and r.support[0] == 0]
    key = lambda r: (r.support, r.interior, r.rank, r.degenerate)
    assert {key(r) for r in recs} == {key(r) for r in ref}
    bad = [r for r in recs if r.interior]
    assert bad, "need a record with an interior"
    planted = {key(r) for r in recs if r is not bad[0]} | {(bad[0].support, bad[0].interior[:-1] + (29,),
                                                             bad[0].rank, bad[0].degenerate)}
    assert planted != {key(r) for r in ref}


def test_emst_graph_equals_complete_graph_kruskal():
    from scipy.spatial import cKDTree

    for n, seed in ((40, 1), (300, 2)):
        pts = oracle.canonicalize(workloads.uniform_cloud(n, seed=seed))
        full = oracle.emst(pts)
        u, v = np.triu_indices(n, 1)
        assert np.array_equal(oracle.emst_graph(pts, np.stack([u, v], 1)), full)
        _, nb = cKDTree(pts).query(pts, k=min(n, 17))
        cand = np.stack([np.repeat(np.arange(n), nb.shape[1] - 1), nb[:, 1:].ravel()], 1)
        assert np.array_equal(oracle.emst_graph(pts, cand), full)
    # a graph without the tree's edges gives a different (heavier) forest
    pts = oracle.canonicalize(workloads.uniform_cloud(60, seed=3))
    full = oracle.emst(pts)
    u, v = np.triu_indices(60, 1)
    keep = ~((u == full[0, 0]) & (v == full[0, 1]))
    assert not np.array_equal(oracle.emst_graph(pts, np.stack([u[keep], v[keep]], 1)), full)


def test_knn_partition_oracle_pins():
    """The exact top-k partition oracle against the reference's
    gpu_fp64_tie_split fixture (two points one ulp apart, a rational query
    half way: an exact tie that FP64 splits) and closed forms."""
    with open(os.path.join(HERE, "golden", "regressions.json"), encoding="utf-8") as fh:
        fx = json.load(fh)["gpu_fp64_tie_split"]
    pts = np.array([[float.fromhex(x) if x.startswith("0x") else struct.unpack(">d", bytes.fromhex(x))[0]
                     for x in p] for p in fx["points_binary64"]])
    qd = fx["query"]
    num = [int(qd["x_numerator"]), int(qd["y_numerator"]), int(qd["z_numerator"])]
    r = oracle.knn_partition(pts, num, int(qd["denominator"]), fx["requested_rank"])
    cut = fx["expected_exact_cutoff"]
    assert r["cutoff_sq"] == Fraction(int(cut["numerator"]), int(cut["denominator"]))
    assert r["shell"] == fx["expected_exact_shell_ids"] and r["strict"] == []
    assert r["choice"] == [0]
    # the centre of a unit cube of an integer grid: eight corners at 3/4
    g = np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    c = oracle.canonicalize(g.astype(float))
    r = oracle.knn_partition(c, [1, 1, 1], 2, 3)
    assert r["cutoff_sq"] == Fraction(3, 4) and len(r["shell"]) == 8 and r["strict"] == []
    assert r["choice"] == r["shell"][:3]
    # a grid point itself: strictly below it the point, then six at distance 1
    r = oracle.knn_partition(c, [1, 1, 1], 1, 4)
    assert r["cutoff_sq"] == 1 and len(r["strict"]) == 1 and len(r["shell"]) == 6
    # excluding the point moves every neighbour one rank up
    pid = r["strict"][0]
    r2 = oracle.knn_partition(c, [1, 1, 1], 1, 4, exclusions=[pid])
    assert r2["cutoff_sq"] == 1 and r2["strict"] == [] and r2["shell"] == r["shell"]
