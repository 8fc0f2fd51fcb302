"""hgp_reduced tower of the backend vs the exhaustive Gamma and direct oracles."""

import numpy as np
import pytest

import oracle
from oracle import tower as otower
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def product_forests(ctx, k_max, cat):
    """Nodes per order with exact oracle levels attached through level classes."""
    t = ctx.tower(k_max)
    return t


def _exact_level_of_class(cat, canon, k_max):
    ref = {(r.support, r.interior): r.level for r in oracle.catalog(canon, k_max) if len(r.support) >= 2}
    out = {-1: 0}
    for i in range(len(cat.size)):
        m = int(cat.size[i])
        key = (tuple(int(x) for x in cat.ids[i, :m]), cat.interior[i])
        out[int(cat.level_class[i])] = ref[key]
    return out


def _signatures(t, levels):
    nodes = []
    off = t["child_offset"]
    kinds = {0: "birth", 1: "birth", 2: "merge"}
    for i in range(len(t["order"])):
        kids = tuple(int(x) for x in t["children"][off[i]:off[i + 1]])
        nodes.append({"order": int(t["order"][i]), "kind": kinds[int(t["kind"][i])],
                      "level": levels[int(t["level_class"][i])], "children": kids, "covered": None})
    by_order = {}
    for i, nd in enumerate(nodes):
        by_order.setdefault(nd["order"], []).append(i)
    out = {}
    for k, idx in by_order.items():
        remap = {old: new for new, old in enumerate(idx)}
        sub = []
        for old in idx:
            nd = dict(nodes[old])
            nd["children"] = tuple(remap[c] for c in nd["children"])
            sub.append(nd)
        out[k] = otower.signatures(sub)
    return out


def _node_sigs(nodes):
    """Per node (local index) its signature (order, kind, level, children)."""
    memo = {}

    def sig(i):
        if i not in memo:
            nd = nodes[i]
            memo[i] = (nd["order"], nd["kind"], nd["level"], tuple(sorted(sig(c) for c in nd["children"])))
        return memo[i]

    return [sig(i) for i in range(len(nodes))]


def _product_vertical(t, levels):
    """Vertical assignments of the product as (source, target) signatures,
    with the invariants every vertical map satisfies: the target's level is at
    most the source's, and a child's target lies below its parent's target."""
    kinds = {0: "birth", 1: "birth", 2: "merge"}
    off = t["child_offset"]
    n = len(t["order"])
    order = [int(x) for x in t["order"]]
    local, per = {}, {}
    for i in range(n):
        per.setdefault(order[i], []).append(i)
    nodes = {}
    for k, idx in per.items():
        remap = {old: new for new, old in enumerate(idx)}
        for old in idx:
            local[old] = (k, remap[old])
        nodes[k] = [{"order": k, "kind": kinds[int(t["kind"][old])], "level": levels[int(t["level_class"][old])],
                     "children": tuple(remap[int(c)] for c in t["children"][off[old]:off[old + 1]])} for old in idx]
    sigs = {k: _node_sigs(v) for k, v in nodes.items()}
    parent = {}
    for i in range(n):
        for c in t["children"][off[i]:off[i + 1]]:
            parent[int(c)] = i
    vert = t["vertical"]
    pairs = []
    for i in range(n):
        if order[i] == 1:
            assert vert[i] == -1
            continue
        tg = int(vert[i])
        assert order[tg] == order[i] - 1
        assert t["level_class"][tg] <= t["level_class"][i]
        if i in parent:  # naturality: the parent's target is an ancestor-or-self of the child's
            x, up = tg, int(vert[parent[i]])
            while x != up and x in parent:
                x = parent[x]
            assert x == up
        pairs.append((sigs[order[i]][local[i][1]], sigs[order[tg]][local[tg][1]]))
    return sorted(pairs)


def _check(pts, k_max, label, gamma=False):
    print(f"tower case {label} n={len(pts)} k={k_max}", flush=True)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(k_max)
        t = ctx.tower(k_max)
    levels = _exact_level_of_class(cat, canon, k_max)
    mine = _signatures(t, levels)
    pairs = _product_vertical(t, levels)
    if gamma:
        ref, maps = otower.gamma_forests(canon, k_max, "hgp_reduced", vertical=True)
    else:
        ref = otower.direct_forests(canon, k_max)
    for k in range(1, min(k_max, len(pts)) + 1):
        assert mine.get(k, []) == otower.signatures(ref[k]), (label, k)
    if gamma:
        want = []
        for k, targets in maps.items():
            ss, ts = _node_sigs(ref[k]), _node_sigs(ref[k - 1])
            want += [(ss[i], ts[tg]) for i, tg in enumerate(targets)]
        assert pairs == sorted(want), (label, "vertical maps")


def test_tower_matches_gamma_small():
    rng = np.random.default_rng(31)
    for trial in range(4):
        n = int(rng.integers(6, 10))
        pts = rng.integers(-50, 50, size=(n, 3)) + rng.random((n, 3)) * 1e-3
        _check(pts, int(rng.integers(1, n + 1)), f"gamma{trial}", gamma=True)


def test_tower_matches_direct_oracle():
    _check(workloads.uniform_cloud(30, seed=41), 10, "uniform30")
    _check(workloads.jittered_lattice(40, side=4, seed=42), 6, "lattice40")


def test_tower_contract_overlap_case():
    import json, os
    with open(os.path.join(os.path.dirname(__file__), "golden", "contracts.json")) as fh:
        cases = {c["name"]: c for c in json.load(fh)}
    for name in ("overlap-k2", "k1-emst"):
        case = cases[name]
        _check(np.asarray(case["points"], dtype=float), case["k_max"], name, gamma=True)


def test_tower_shared_levels_translated_copies():
    """Three exact translates of one cluster: every sphere has twins of the
    same exact level, so saddles share levels and are contracted as groups."""
    rng = np.random.default_rng(5)
    a = np.round(rng.random((16, 3)) * 2**20) / 2**20  # translation by 4 stays exact
    pts = np.concatenate([a, a + np.array([4.0, 0, 0]), a + np.array([0, 4.0, 0])])
    _check(pts, 8, "translated-copies")


def test_tower_lattice90():
    """Enough minima per order that most saddles only grow a component."""
    _check(workloads.jittered_lattice(90, side=5, seed=44), 5, "lattice90")


def test_plain_reference_configuration_uniform256_k3():
    """The plain CPU reference configuration (256 points uniform in [0, 1)^3,
    K_max 3): EMST edges, catalogue records and the hgp_reduced forests equal
    the exact oracle's.  The oracle's catalogue is its per-anchor exhaustive
    enumeration with radius 4, which encloses every well-centred sphere of
    points of the unit cube; the forests are its direct construction."""
    pts = workloads.uniform_cloud(256, seed=0)
    canon = oracle.canonicalize(pts)
    print("plain reference configuration: uniform n=256 k=3, oracle", flush=True)
    ref_edges = np.asarray(oracle.emst(canon), dtype=np.int64)
    recs = []
    for a in range(len(canon)):
        recs += oracle.anchor_records(canon, 3, a, 4.0)
    ref = otower.direct_forests(canon, 3, events=recs)
    print("plain reference configuration: GPU", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        edges = ctx.emst()
        cat = ctx.supports(3)
        t = ctx.tower(3)
    assert np.array_equal(np.asarray(edges, dtype=np.int64), ref_edges)
    exact = {(r.support, r.interior): (r.rank, r.level) for r in recs}
    mine = {}
    for i in range(len(cat.size)):
        key = (tuple(int(x) for x in cat.ids[i, : cat.size[i]]), cat.interior[i])
        mine[key] = (int(cat.rank[i]), int(cat.level_class[i]))
    assert set(mine) == set(exact)
    levels = {-1: 0}
    for key, (rank, cls) in mine.items():
        assert rank == exact[key][0], key
        assert levels.setdefault(cls, exact[key][1]) == exact[key][1], key  # a class is one exact level
    sigs = _signatures(t, levels)
    for k in (1, 2, 3):
        assert sigs.get(k, []) == otower.signatures(ref[k]), k


def test_vertical_growth_target_regression():
    """tests/fixtures/regressions/vertical_q1_growth_target.json: the order-3
    component at level 4 maps onto the order-2 root born at level 1 covering
    {0, 1, 2}, evaluated in its state at level 4 (it has grown by then)."""
    import json, os
    from fractions import Fraction
    with open(os.path.join(os.path.dirname(__file__), "golden", "regressions.json")) as fh:
        fx = json.load(fh)["vertical_q1_growth_target"]
    w = fx["vertical_witness"]
    pts = np.asarray(fx["points"], dtype=float)
    _check(pts, fx["k_max"], "vertical-q1-growth", gamma=True)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(fx["k_max"])
        t = ctx.tower(fx["k_max"])
    levels = _exact_level_of_class(cat, canon, fx["k_max"])
    at = Fraction(w["at_squared_level"]["numerator"], w["at_squared_level"]["denominator"])
    born = Fraction(w["target_birth_squared_level"]["numerator"], w["target_birth_squared_level"]["denominator"])
    src = [i for i in range(len(t["order"])) if t["order"][i] == w["source_order"]
           and levels[int(t["level_class"][i])] == at]
    assert src, "no source node at the witnessed level"
    tgt = int(t["vertical"][src[0]])
    assert t["order"][tgt] == w["target_order"] and levels[int(t["level_class"][tgt])] == born


def test_order_one_cuts_equal_emst_components():
    """Strict and closed cuts of the order-one forest: the components equal
    the union-find of the exact EMST edges of level (d^2 / 4) below or at the
    cut (reference K1CompactForest::cut_strict / cut_closed)."""
    from fractions import Fraction
    pts = workloads.uniform_cloud(80, seed=52)
    canon = oracle.canonicalize(pts)
    print("tower cuts: uniform n=80 k=3", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(3)
        t = ctx.tower(3)
        levels = _exact_level_of_class(cat, canon, 3)
        edges = oracle.emst(canon)
        d2 = [sum((Fraction(float(canon[u][a])) - Fraction(float(canon[v][a]))) ** 2 for a in range(3)) / 4
              for u, v in edges]
        classes = sorted(set(int(c) for c in t["level_class"][t["order"] == 1]))
        for cls in [-1] + classes[:: max(1, len(classes) // 12)]:
            lev = levels[cls]
            for closed in (True, False):
                roots, lab = ctx.tower_cut(1, cls, closed, node_count=len(t["order"]))
                parent = list(range(len(canon)))

                def find(x):
                    while parent[x] != x:
                        parent[x] = parent[parent[x]]
                        x = parent[x]
                    return x

                for (u, v), w in zip(edges, d2):
                    if (w <= lev) if closed else (w < lev):
                        parent[find(int(u))] = find(int(v))
                if not closed and lev == 0:
                    assert (lab == -1).all() and len(roots) == 0
                    continue
                groups = {}
                for p in range(len(canon)):
                    groups.setdefault(find(p), set()).add(p)
                mine = {}
                for p in range(len(canon)):
                    mine.setdefault(int(lab[p]), set()).add(p)
                assert sorted(map(sorted, groups.values())) == sorted(map(sorted, mine.values())), (cls, closed)
                assert sorted(mine) == sorted(int(r) for r in roots)
