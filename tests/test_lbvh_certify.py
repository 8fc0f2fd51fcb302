"""Exact host re-certification of LBVH snapshots (mhgp_lbvh_certify, the
role of the reference's import_certified_snapshot, lbvh.cpp:562-829): the
oracle's snapshots are accepted, and every planted mistake (a code, the leaf
order, a split, a witness, a counter, a non-canonical cloud) is rejected.
Runs on the CPU: the certifier uses no device."""

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads


def _snap(pts):
    c = oracle.canonicalize(pts)
    s = oracle.lbvh_snapshot(c)
    return c, s.leaves.copy(), s.nodes.copy(), s.counters.copy()


@pytest.mark.parametrize("name,pts", [
    ("uniform500", workloads.uniform_cloud(500, seed=1)),
    ("lattice5k", workloads.jittered_lattice(5000, side=18, seed=2)),
    ("collisions", np.array([[0, 0, 0], [0, 0, 1e-9], [0, 1e-9, 0], [1, 1, 1], [1, 1, 1 - 1e-12], [0.5, 0.5, 0.5]])),
    ("flat", np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0.0]])),
    ("single", np.array([[1.0, 2.0, 3.0]])),
])
def test_oracle_snapshots_certify(name, pts):
    c, lv, nd, ct = _snap(pts)
    mhgp.lbvh_certify(c, lv, nd, ct)


def _rejects(c, lv, nd, ct, what):
    with pytest.raises(mhgp.MhgpError) as err:
        mhgp.lbvh_certify(c, lv, nd, ct)
    assert err.value.code == mhgp.E_ARGUMENT
    assert what in str(err.value), str(err.value)


def test_planted_mistakes_are_rejected():
    c, lv, nd, ct = _snap(workloads.uniform_cloud(400, seed=3))
    bad = lv.copy(); bad[5, 0] ^= np.uint64(1)
    _rejects(c, bad, nd, ct, "")
    bad = lv.copy(); bad[[5, 6]] = bad[[6, 5]]
    _rejects(c, bad, nd, ct, "order")
    inner = [i for i in range(len(nd)) if nd[i, 6] != np.uint64(2**64 - 1)]
    j = inner[len(inner) // 2]
    bad = nd.copy(); bad[j, 0] = bad[j, 3]
    _rejects(c, lv, bad, ct, "witness")
    bad = nd.copy(); bad[j, [6, 7]] = bad[j, [7, 6]]
    _rejects(c, lv, bad, ct, "split rule")
    for k in range(5):
        bad = ct.copy(); bad[k] += np.uint64(1)
        _rejects(c, lv, nd, bad, "")
    shuffled = c[::-1].copy()
    _rejects(shuffled, lv, nd, ct, "canonical")


def test_equal_code_run_uses_the_midpoint_split():
    """Eight points in one Morton cell: their codes are equal, the tree over
    the run must split at b + (e - b) / 2; a split one further is rejected."""
    base = np.array([[0.0, 0, 0], [1.0, 1, 1]])
    tiny = np.array([[0.5 + i * 1e-12, 0.5, 0.5] for i in range(8)])
    c, lv, nd, ct = _snap(np.concatenate([base, tiny]))
    assert ct[3] >= 1 and ct[4] >= 8
    mhgp.lbvh_certify(c, lv, nd, ct)
