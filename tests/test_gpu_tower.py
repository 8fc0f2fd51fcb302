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


def _check(pts, k_max, label, gamma=False):
    print(f"tower case {label} n={len(pts)} k={k_max}", flush=True)
    canon = oracle.canonicalize(pts)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(k_max)
        t = ctx.tower(k_max)
    levels = _exact_level_of_class(cat, canon, k_max)
    mine = _signatures(t, levels)
    ref = otower.gamma_forests(canon, k_max, "hgp_reduced") if gamma else otower.direct_forests(canon, k_max)
    for k in range(1, min(k_max, len(pts)) + 1):
        assert mine.get(k, []) == otower.signatures(ref[k]), (label, k)


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
