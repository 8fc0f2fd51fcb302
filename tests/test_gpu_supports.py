"""Critical-sphere catalogue on the GPU vs the exhaustive oracle catalogue."""

from fractions import Fraction
from math import comb

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def _product_records(cat):
    out = []
    for i in range(len(cat.size)):
        m = int(cat.size[i])
        sup = tuple(int(x) for x in cat.ids[i, :m])
        out.append((sup, cat.interior[i], int(cat.rank[i]), bool(cat.degenerate[i])))
    return out


def _compare(pts, k_max, label):
    print(f"catalog case {label} n={len(pts)} k={k_max}", flush=True)
    canon = oracle.canonicalize(pts)
    ref = [r for r in oracle.catalog(canon, k_max) if len(r.support) >= 2]
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(k_max)
    mine = _product_records(cat)
    ref_set = sorted((r.support, r.interior, r.rank, r.degenerate) for r in ref)
    assert sorted(mine) == ref_set, label
    # exact level classes agree with the oracle's exact rational levels
    lev = {(r.support, r.interior): r.level for r in ref}
    exact = [lev[(s, i)] for (s, i, _, _) in mine]
    for a in range(len(mine) - 1):
        assert exact[a] <= exact[a + 1], "canonical order must follow exact levels"
        assert (cat.level_class[a] == cat.level_class[a + 1]) == (exact[a] == exact[a + 1])
    for a in range(len(mine)):
        assert abs(cat.level[a] - float(exact[a])) <= 1e-9 * float(exact[a])
    return cat


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_catalog_uniform_small(seed):
    n = [24, 40, 60][seed - 1]
    pts = workloads.uniform_cloud(n, seed=seed)
    for k in (1, 3, 10):
        _compare(pts, k, f"uniform{n}-s{seed}")


def test_catalog_lattice_small():
    pts = workloads.jittered_lattice(64, side=5, seed=9)
    _compare(pts, 10, "lattice64")
    pts = workloads.jittered_lattice(110, side=5, seed=10)
    _compare(pts, 10, "lattice110")


def test_catalog_mixture_small():
    pts = workloads.gaussian_mixture(70, blobs=3, snap_bits=None, seed=4)
    _compare(pts, 10, "mixture70")


def test_catalog_contract_clouds():
    import json, os
    with open(os.path.join(os.path.dirname(__file__), "golden", "selected_n12_n14.json")) as fh:
        cases = json.load(fh)
    for case in cases:
        pts = np.asarray(case["points"], dtype=float)
        cat = _compare(pts, case["k_max"], case["name"])
        # singletons are implicit in the product catalogue
        assert cat.info.event_count + len(pts) == case["expected_event_count"]


def _euler(cat, n, k):
    """Sum over order-k critical points of (-1)^index * multiplicity."""
    total = n if k == 1 else 0  # singletons: index 0 at order 1
    for i in range(len(cat.size)):
        if cat.degenerate[i]:
            continue
        m, s = int(cat.size[i]), int(cat.rank[i])
        nint = s - m
        if not (nint < k <= s):
            continue
        idx = s - k
        total += (-1) ** idx * comb(m - 1, idx)
    return total


def test_catalog_euler_characteristic_lattice_5k():
    pts = workloads.jittered_lattice(5000, side=18, seed=21)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        cat = ctx.supports(10)
    assert cat.info.diagnostic_count == 0
    for k in range(1, 9):  # complete windows: ranks k..k+3 <= s_max = 11
        assert _euler(cat, len(pts), k) == 1, k


def test_capacity_error_is_clean_and_context_survives():
    """Anchors whose candidate region exceeds the global-memory path's bound
    report MHGP_E_CAPACITY (no fault, no poisoning); the same context then
    serves an ordinary cloud."""
    pts = workloads.gaussian_mixture(50_000, seed=0)
    print("capacity case: snapped mixture n=50000 k=10", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.supports(10, fetch=False)
        assert err.value.code == mhgp.E_CAPACITY, err.value
        small = workloads.uniform_cloud(400, seed=9)
        ctx.upload(small)
        cat = ctx.supports(4)
    with mhgp.Context(0) as fresh:
        fresh.upload(small)
        ref = fresh.supports(4)
    assert np.array_equal(cat.ids, ref.ids) and np.array_equal(cat.level_class, ref.level_class)
