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
        shell, witness, exterior = ctx.diagnostics(len(cat.size))
    mine = _product_records(cat)
    ref_set = sorted((r.support, r.interior, r.rank, r.degenerate) for r in ref)
    assert sorted(mine) == ref_set, label
    # shell diagnostics: shell size, least extra-shell point, exterior count
    diag = {(r.support, r.interior): (r.shell_count, r.shell_witness, len(canon) - r.rank - (r.shell_count - len(r.support)))
            for r in ref}
    for i, (sup, inter, _, _) in enumerate(mine):
        assert (int(shell[i]), int(witness[i]), int(exterior[i])) == diag[(sup, inter)], (label, sup)
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


@pytest.mark.parametrize("seed,n,bits", [(5, 100, 2), (6, 110, 3), (7, 110, 12)])
def test_catalog_snapped_mixture_small(seed, n, bits):
    """Snapped anisotropic mixtures with noise (the large-cloud stress
    distribution at oracle size): coarse grids force exact cospherical and
    coplanar ties, extra-shell diagnostics and host settlement."""
    pts = workloads.gaussian_mixture(n, blobs=3, noise_fraction=0.15, box=2.0, snap_bits=bits, seed=seed)
    cat = _compare(pts, 10, f"snapped{n}-b{bits}")
    assert cat.info.status == 0
    if bits <= 3:
        assert cat.info.diagnostic_count > 0 or cat.info.fallback_count > 0


@pytest.mark.parametrize("path", ["region", "slabs"])
def test_catalog_big_paths_small(path, monkeypatch):
    """The big-anchor paths (probe + LDS region kernels; global-memory slabs)
    against the exhaustive oracle: the test hook MHGP_TEST_ANCHOR_PATH sends
    every anchor of these small clouds through them."""
    monkeypatch.setenv("MHGP_TEST_ANCHOR_PATH", path)
    _compare(workloads.uniform_cloud(60, seed=31), 10, f"uniform60-{path}")
    _compare(workloads.jittered_lattice(64, side=5, seed=32), 10, f"lattice64-{path}")
    pts = workloads.gaussian_mixture(100, blobs=3, noise_fraction=0.15, box=2.0, snap_bits=3, seed=33)
    _compare(pts, 10, f"snapped100-{path}")
    _compare(workloads.uniform_cloud(40, seed=34), 4, f"uniform40-k4-{path}")


def test_big_regions_give_a_conditional_catalogue():
    """Anchors whose candidate region exceeds the enumeration bound are left
    open: the catalogue reports status conditional with their count and
    largest region (never exact), the tower is refused, and the same context
    then serves an ordinary cloud exactly."""
    pts = workloads.gaussian_mixture(50_000, seed=0)
    print("conditional case: snapped mixture n=50000 k=10", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        info = ctx.supports(10, fetch=False)
        assert info.status == 1
        assert info.anchors_open > 0 and info.largest_open_region > 2048
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.tower(10)
        assert err.value.code == mhgp.E_UNSUPPORTED
        small = workloads.uniform_cloud(400, seed=9)
        ctx.upload(small)
        cat = ctx.supports(4)
        assert cat.info.status == 0 and cat.info.anchors_open == 0
    with mhgp.Context(0) as fresh:
        fresh.upload(small)
        ref = fresh.supports(4)
    assert np.array_equal(cat.ids, ref.ids) and np.array_equal(cat.level_class, ref.level_class)


def test_region_limit_is_an_enumeration_budget():
    """mhgp_ctx_set_region_limit: with the limit at 640 the anisotropic
    mixture's anchors whose candidate regions exceed it stay open (status
    conditional, largest open region within (640, 2048]), and every record
    of that catalogue is a record of the unlimited one.  Limits outside
    640..2048 are refused."""
    rng = workloads.SplitMix64(1012).numpy_rng()
    rng.integers(3)
    pts = workloads.gaussian_mixture(6467, blobs=int(rng.integers(5, 21)), noise_fraction=0.15, box=10.0,
                                     snap_bits=None, seed=1012)
    print("region limit: anisotropic mixture n=6467 k=10", flush=True)

    def records(cat):
        return {(tuple(int(x) for x in cat.ids[i, : cat.size[i]]), cat.interior[i], int(cat.rank[i]))
                for i in range(len(cat.size))}

    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        full = ctx.supports(10)
        assert full.info.status == 0
        for bad in (639, 2049):
            with pytest.raises(mhgp.MhgpError) as err:
                ctx.set_region_limit(bad)
            assert err.value.code == mhgp.E_ARGUMENT
        ctx.set_region_limit(640)
        ctx.upload(pts)
        part = ctx.supports(10)
    assert part.info.status == 1 and part.info.anchors_open > 0
    assert 640 < part.info.largest_open_region <= 2048
    mine, ref = records(part), records(full)
    assert mine < ref


def test_context_capacity_refuses_and_recovers():
    """A context created with a device capacity refuses a catalogue whose
    record queues would exceed it (MHGP_E_CAPACITY, no poisoning); after an
    upload the same context serves a small cloud within the bound."""
    pts = workloads.jittered_lattice(20_000, side=28, sigma=0.1, seed=3)
    print("capacity: 20k lattice under a 64 MiB context", flush=True)
    with mhgp.Context(0, capacity_bytes=64 << 20) as ctx:
        ctx.upload(pts)
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.supports(10, fetch=False)
        assert err.value.code == mhgp.E_CAPACITY
        small = workloads.uniform_cloud(300, seed=2)
        ctx.upload(small)
        cat = ctx.supports(4)
        assert ctx.device_bytes() <= 64 << 20
    with mhgp.Context(0) as fresh:
        fresh.upload(small)
        ref = fresh.supports(4)
    assert np.array_equal(cat.ids, ref.ids) and np.array_equal(cat.level_class, ref.level_class)


def test_reused_context_one_point_cloud_has_an_empty_catalogue():
    """A context that held a cloud and then takes a one-point cloud returns an
    empty catalogue (no records left over from the earlier cloud)."""
    with mhgp.Context(0) as ctx:
        ctx.upload(workloads.uniform_cloud(200, seed=4))
        assert ctx.supports(4).info.event_count > 0
        ctx.upload(np.array([[1.0, 2.0, 3.0]]))
        cat = ctx.supports(4)
        assert cat.info.event_count == 0 and len(cat.size) == 0
