"""Batched processing (mhgp_batch_run): groups of clouds side by side with one
persistent anchor launch per group.  Every cloud's result (code, status,
catalogue size) and tower equal those of the same cloud on the plain path,
including a cloud that fails (duplicate points), a one-point cloud (no anchor
pass) and clouds of different K_max, over several groups."""

import numpy as np
import pytest

from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu


def _plain(pts, k):
    with mhgp.Context(0) as ctx:
        try:
            ctx.upload(pts)
            ctx.emst()
            info = ctx.supports(k, fetch=False)
            if info.status != 0:
                return 0, 1, info.event_count, None
            return 0, 0, info.event_count, ctx.tower(k)
        except mhgp.MhgpError as e:
            return e.code, 0, None, None


@pytest.mark.skip(reason="mhgp_batch_run is switched off (returns MHGP_E_UNSUPPORTED): a process running it "
                         "ended with a host segmentation fault whose cause has not been found")
def test_batch_equals_plain_path():
    clouds, ks = [], []
    for i, n in enumerate([2000, 3500, 1200, 5000, 2500, 1800, 4000]):
        if i % 2:
            clouds.append(workloads.gaussian_mixture(n, blobs=5 + i, noise_fraction=0.15, box=10.0, snap_bits=None,
                                                     seed=300 + i))
        else:
            side = int(np.ceil((n * 1.3) ** (1 / 3)))
            clouds.append(workloads.jittered_lattice(n, side=side, seed=300 + i))
        ks.append(10 if i % 3 else 4)
    dup = workloads.uniform_cloud(500, seed=9)
    dup[17] = dup[3]
    clouds += [dup, np.array([[0.5, 0.25, 0.125]]), workloads.uniform_cloud(800, seed=10)]
    ks += [10, 10, 6]
    print(f"batch: {len(clouds)} clouds in groups of 4", flush=True)
    results, towers, info = mhgp.batch_run(clouds, ks, group=4)
    assert info.clouds == len(clouds) and info.groups == 3 and info.anchors > 0 and info.anchor_ms > 0
    for i, (pts, k) in enumerate(zip(clouds, ks)):
        code, status, events, tw = _plain(pts, k)
        r = results[i]
        assert (r.code, r.status) == (code, status), i
        if code != 0:
            continue
        assert r.event_count == events, i
        if tw is None:
            assert towers[i] is None
            continue
        for key in ("order", "kind", "level_class", "child_offset", "children"):
            assert np.array_equal(towers[i][key], tw[key]), (i, key)
    assert results[7].code == mhgp.E_DUPLICATE
