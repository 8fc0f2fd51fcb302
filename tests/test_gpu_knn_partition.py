"""Exact top-k partitions of rational queries (mhgp_knn_partition) against
the exact oracle: cutoff, strict-below list, complete cutoff shell and the
canonical choice, with exclusions; the reference's gpu_fp64_tie_split
fixture (an exact tie that a binary64 proposal splits), ties on an integer
grid, and random rational queries."""

import json
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import mhgp, workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _check(ctx, canon, num, den, k, excl=None):
    got = ctx.knn_partition(num, den, k, excl, shell_cap=64)
    for q in range(len(num)):
        ref = oracle.knn_partition(canon, num[q], den[q], k, (excl or [[]] * len(num))[q])
        g = got[q]
        assert g["strict"] == ref["strict"], q
        assert g["shell_count"] == len(ref["shell"]) and g["shell"] == ref["shell"][:64], q
        assert g["choice"] == ref["choice"], q
        assert g["cutoff_id"] == ref["shell"][0], q
        c = float(ref["cutoff_sq"])
        assert abs(g["cutoff_sq"] - c) <= 1e-15 * c, q


def test_gpu_fp64_tie_split_fixture():
    with open(os.path.join(HERE, "golden", "regressions.json"), encoding="utf-8") as fh:
        fx = json.load(fh)["gpu_fp64_tie_split"]
    pts = np.array([[struct.unpack(">d", bytes.fromhex(x))[0] for x in p] for p in fx["points_binary64"]])
    qd = fx["query"]
    num = [[int(qd["x_numerator"]), int(qd["y_numerator"]), int(qd["z_numerator"])]]
    den = [int(qd["denominator"])]
    print("knn partition: gpu_fp64_tie_split fixture", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        canon = oracle.canonicalize(pts)
        r = ctx.knn_partition(num, den, fx["requested_rank"])[0]
        assert r["shell"] == fx["expected_exact_shell_ids"] and r["strict"] == []
        cut = fx["expected_exact_cutoff"]
        assert r["cutoff_sq"] == float(Fraction(int(cut["numerator"]), int(cut["denominator"])))
        _check(ctx, canon, num, den, 1)


def test_grid_ties_and_exclusions():
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij"), -1).reshape(-1, 3)
    pts = g.astype(np.float64) * 0.5
    canon = oracle.canonicalize(pts)
    rng = np.random.default_rng(3)
    # cube centres, face centres, edge midpoints and grid points themselves (den 4)
    num = rng.integers(0, 11, size=(40, 3))
    den = [4] * 40
    print("knn partition: 216-point grid, 40 half-grid queries, k 1..12", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        for k in (1, 3, 8, 12):
            _check(ctx, canon, num.tolist(), den, k)
        excl = [[int(x) for x in rng.choice(216, size=int(rng.integers(0, 20)), replace=False)] for _ in range(40)]
        _check(ctx, canon, num.tolist(), den, 6, excl)
        # an exclusion list longer than the device checks goes to the host scan
        long = [[int(x) for x in rng.choice(216, size=100, replace=False)] for _ in range(3)]
        _check(ctx, canon, num[:3].tolist(), den[:3], 5, long)


def test_random_rational_queries_and_errors():
    pts = workloads.uniform_cloud(600, seed=71)
    canon = oracle.canonicalize(pts)
    rng = np.random.default_rng(4)
    den = rng.integers(1, 2**20, size=30)
    num = (rng.random((30, 3)) * den[:, None]).astype(np.int64)
    print("knn partition: uniform n=600, 30 random rational queries", flush=True)
    with mhgp.Context(0) as ctx:
        ctx.upload(pts)
        _check(ctx, canon, num.tolist(), den.tolist(), 11)
        _check(ctx, canon, num.tolist(), den.tolist(), 32)
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.knn_partition(num[:1].tolist(), [0], 3)
        assert err.value.code == mhgp.E_ARGUMENT
        with pytest.raises(mhgp.MhgpError) as err:
            ctx.knn_partition(num[:1].tolist(), [5], 33)
        assert err.value.code == mhgp.E_ARGUMENT
        # the context stays usable after argument errors
        _check(ctx, canon, num[:2].tolist(), den[:2].tolist(), 4)
