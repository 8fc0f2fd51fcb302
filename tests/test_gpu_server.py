"""The serving loop returns, for every request in a mixed stream, the same
tower as the plain per-context path does for that cloud alone."""

import numpy as np
import pytest

from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def _plain(pts, k_max):
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        ctx.emst()
        ctx.supports(k_max, fetch=False)
        return ctx.tower(k_max)


def _same(a, b):
    for key in ("order", "kind", "level_class", "level", "child_offset", "children"):
        assert np.array_equal(a[key], b[key]), key


def _clouds():
    return [
        (workloads.uniform_cloud(300, seed=11), 4),
        (workloads.jittered_lattice(2000, side=13, sigma=0.1, seed=3), 10),
        (workloads.uniform_cloud(900, seed=12), 10),
        (workloads.gaussian_mixture(1500, blobs=5, snap_bits=None, seed=4), 10),
        (workloads.uniform_cloud(64, seed=13), 3),
    ]


@pytest.mark.parametrize("slots", [1, 3])
def test_server_matches_plain_path(slots):
    clouds = _clouds()
    print(f"server slots={slots} sizes={[len(p) for p, _ in clouds]}", flush=True)
    plain = [_plain(p, k) for p, k in clouds]
    with mhgp.Server(0, slots=slots) as srv:
        tickets = [srv.submit(p, k) for p, k in clouds]
        for t, ref in zip(tickets, plain):
            r = srv.wait(t)
            assert r.code == 0 and r.status == 0
            assert r.latency_ms >= r.service_ms >= 0 and r.queue_ms >= 0
            _same(srv.fetch(t, r), ref)
            srv.release(t)


def test_server_memory_cap_serialises_and_reports_errors():
    """A cap below one request still serves it (alone); a duplicate point is
    reported on its own ticket without disturbing the others."""
    good = workloads.uniform_cloud(400, seed=21)
    bad = good.copy()
    bad[7] = bad[3]
    with mhgp.Server(0, slots=2, memory_cap_bytes=1) as srv:
        t1 = srv.submit(good, 5)
        t2 = srv.submit(bad, 5)
        t3 = srv.submit(good, 5)
        r1, r2, r3 = srv.wait(t1), srv.wait(t2), srv.wait(t3)
        assert r1.code == 0 and r3.code == 0
        assert r2.code == mhgp.E_DUPLICATE
        _same(srv.fetch(t1, r1), srv.fetch(t3, r3))
        _same(srv.fetch(t1, r1), _plain(good, 5))


def _grid_cloud():
    """3 x 3 x 3 integer grid: many cospherical quadruples, so the catalogue
    holds extra-shell diagnostics and the direct tower does not exist."""
    g = np.arange(3, dtype=np.float64)
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)


def test_degenerate_cloud_refused_alike():
    pts = _grid_cloud()
    print("server degenerate grid n=27 k=3", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        info = ctx.supports(3, fetch=False)
        assert info.diagnostic_count > 0
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.tower(3)
        assert err.value.code == mhgp.E_UNSUPPORTED
    with mhgp.Server(0, slots=2) as srv:
        t_bad = srv.submit(pts, 3)
        t_good = srv.submit(workloads.uniform_cloud(200, seed=5), 3)
        assert srv.wait(t_bad).code == mhgp.E_UNSUPPORTED
        r = srv.wait(t_good)
        assert r.code == 0
        _same(srv.fetch(t_good, r), _plain(workloads.uniform_cloud(200, seed=5), 3))
