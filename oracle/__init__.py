"""CPU oracle for the MorseHGP3D hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path never imports it.

The compute lives in ``liboracle.so`` (built from ``oracle.c`` and
``bigint.c`` by ``oracle/Makefile``): exact dyadic-integer restatements of the
reference's canonical cloud, Morton LBVH, brute-force k-NN, complete-graph
EMST and exhaustive critical catalogue.  ``tower.py`` adds the small-n
exhaustive Gamma hierarchy and the direct Morse tower in exact rationals.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        i64, i32 = ctypes.c_int64, ctypes.c_int
        p = ctypes.c_void_p
        _lib.or_canonical_order.argtypes = [p, i64, p]
        _lib.or_lbvh_snapshot.argtypes = [p, i64, p, p, p, p]
        _lib.or_knn.argtypes = [p, i64, i64, p, p]
        _lib.or_emst.argtypes = [p, i64, p]
        _lib.or_catalog_build.argtypes = [p, i64, i32, i32]
        _lib.or_catalog_build.restype = i64
        _lib.or_catalog_get.argtypes = [i64, p, p, p, p, p, p, p, p, p, ctypes.c_char_p,
                                        ctypes.c_char_p, i32]
        _lib.or_catalog_level_cmp.argtypes = [i64, i64]
        _lib.or_support_side.argtypes = [p, i32, i64]
        _lib.or_support_info.argtypes = [p, i32, p, ctypes.c_char_p, ctypes.c_char_p, i32]
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise OracleError(f"oracle {what} failed with code {rc}")


def _pts(points) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    return a


def canonical_order(points) -> np.ndarray:
    """perm[point_id] = source index (reference point_cloud.cpp:40-95)."""
    a = _pts(points)
    perm = np.zeros(len(a), dtype=np.int64)
    _check(lib().or_canonical_order(a.ctypes.data, len(a), perm.ctypes.data), "canonical_order")
    return perm


def canonicalize(points) -> np.ndarray:
    a = _pts(points)
    a = np.where(a == 0.0, 0.0, a)  # -0 -> +0
    return np.ascontiguousarray(a[canonical_order(a)])


@dataclass
class LbvhSnapshot:
    leaves: np.ndarray  # (n, 2) code, point id
    nodes: np.ndarray  # (2n-1, 10) lo[3], hi[3], left, right, begin, end
    counters: np.ndarray  # point_count, node_count, max_depth, collision groups, max collision
    codes_by_id: np.ndarray


def lbvh_snapshot(canonical_points) -> LbvhSnapshot:
    a = _pts(canonical_points)
    n = len(a)
    leaves = np.zeros((n, 2), dtype=np.uint64)
    nodes = np.zeros((2 * n - 1, 10), dtype=np.uint64)
    counters = np.zeros(5, dtype=np.uint64)
    codes = np.zeros(n, dtype=np.uint64)
    _check(lib().or_lbvh_snapshot(a.ctypes.data, n, leaves.ctypes.data, nodes.ctypes.data,
                                  counters.ctypes.data, codes.ctypes.data), "lbvh_snapshot")
    return LbvhSnapshot(leaves, nodes, counters, codes)


def knn(canonical_points, k: int):
    a = _pts(canonical_points)
    out = np.zeros((len(a), k), dtype=np.int64)
    ties = np.zeros(len(a), dtype=np.uint8)
    _check(lib().or_knn(a.ctypes.data, len(a), k, out.ctypes.data, ties.ctypes.data), "knn")
    return out, ties.astype(bool)


def emst(canonical_points) -> np.ndarray:
    a = _pts(canonical_points)
    n = len(a)
    edges = np.zeros((max(n - 1, 0), 2), dtype=np.int64)
    _check(lib().or_emst(a.ctypes.data, n, edges.ctypes.data), "emst")
    return edges


@dataclass(frozen=True)
class CatalogRecord:
    support: tuple
    interior: tuple
    rank: int
    level: Fraction
    degenerate: bool
    shell_count: int
    shell_witness: int

    @property
    def key(self):
        return (self.support, self.interior, self.rank, self.level, self.degenerate)


def catalog(canonical_points, k_max: int, max_support: int = 4) -> list:
    """Exhaustive catalogue: every well-centred support of size 1..4 whose
    closed rank is <= s_max, plus relevant extra-shell degeneracies."""
    a = _pts(canonical_points)
    L = lib()
    count = L.or_catalog_build(a.ctypes.data, len(a), k_max, max_support)
    if count < 0:
        raise OracleError(f"catalogue failed with code {count}")
    out = []
    m = ctypes.c_int(); rank = ctypes.c_int(); nint = ctypes.c_int(); deg = ctypes.c_int()
    shell = ctypes.c_int(); wit = ctypes.c_int64(); lev = ctypes.c_double()
    sup = np.zeros(4, dtype=np.int64)
    interior = np.zeros(16, dtype=np.int64)
    num = ctypes.create_string_buffer(4096)
    den = ctypes.create_string_buffer(4096)
    for i in range(count):
        _check(L.or_catalog_get(i, ctypes.byref(m), sup.ctypes.data, ctypes.byref(rank),
                                ctypes.byref(nint), interior.ctypes.data, ctypes.byref(deg),
                                ctypes.byref(shell), ctypes.byref(wit), ctypes.byref(lev),
                                num, den, 4096), "catalog_get")
        out.append(CatalogRecord(
            support=tuple(int(x) for x in sup[: m.value]),
            interior=tuple(sorted(int(x) for x in interior[: nint.value])),
            rank=rank.value,
            level=Fraction(int(num.value), int(den.value)),
            degenerate=bool(deg.value),
            shell_count=shell.value,
            shell_witness=int(wit.value),
        ))
    return out


_anchor_cloud = {"key": None, "arr": None}


def anchor_catalog(canonical_points, k_max: int, anchor: int, radius: float) -> int:
    """Exact count of the critical supports whose least PointId is `anchor`,
    enumerated exhaustively among the points within `radius` of it."""
    L = lib()
    if not hasattr(L.or_anchor_catalog, "_typed"):
        L.or_anchor_prepare.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.or_anchor_catalog.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_double]
        L.or_anchor_catalog.restype = ctypes.c_int64
        L.or_anchor_catalog._typed = True
    a = _pts(canonical_points)
    key = (a.ctypes.data, len(a))
    if _anchor_cloud["key"] != key:
        _check(L.or_anchor_prepare(a.ctypes.data, len(a)), "anchor_prepare")
        _anchor_cloud["key"], _anchor_cloud["arr"] = key, a
    r = L.or_anchor_catalog(int(k_max), int(anchor), float(radius))
    if r < 0:
        raise OracleError(f"anchor catalogue failed with code {r}")
    return int(r)


def support_info(ids) -> tuple:
    """(well_centred, level) of a support on the cloud of the last catalog() call."""
    arr = np.ascontiguousarray(np.asarray(ids, dtype=np.int64))
    well = ctypes.c_int()
    num = ctypes.create_string_buffer(4096)
    den = ctypes.create_string_buffer(4096)
    rc = lib().or_support_info(arr.ctypes.data, len(arr), ctypes.byref(well), num, den, 4096)
    if rc == 2:
        return None
    return bool(well.value), Fraction(int(num.value), int(den.value))


def support_side(ids, x: int) -> int:
    arr = np.ascontiguousarray(np.asarray(ids, dtype=np.int64))
    return int(lib().or_support_side(arr.ctypes.data, len(arr), int(x)))


def anchor_records(canonical_points, k_max: int, anchor: int, radius: float) -> list:
    """Exhaustive exact records whose least PointId is `anchor` among the
    points within `radius` of it (complete for circumradii <= radius / 2).
    Reentrant in the C library: threads may call it concurrently once the
    cloud is prepared (the call releases the GIL)."""
    L = lib()
    if not hasattr(L.or_anchor_records, "_typed"):
        L.or_anchor_prepare.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        p = ctypes.c_void_p
        L.or_anchor_records.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
                                        p, p, p, p, p, p, p, p, p, ctypes.c_int]
        L.or_anchor_records.restype = ctypes.c_int64
        L.or_anchor_records._typed = True
    a = _pts(canonical_points)
    key = (a.ctypes.data, len(a))
    if _anchor_cloud["key"] != key:
        _check(L.or_anchor_prepare(a.ctypes.data, len(a)), "anchor_prepare")
        _anchor_cloud["key"], _anchor_cloud["arr"] = key, a
    cap, strcap = 2048, 160
    while True:
        m = np.zeros(cap, np.int32); sup = np.zeros((cap, 4), np.int64); rank = np.zeros(cap, np.int32)
        nint = np.zeros(cap, np.int32); inter = np.zeros((cap, 10), np.int64); deg = np.zeros(cap, np.int32)
        shell = np.zeros(cap, np.int32)
        num = ctypes.create_string_buffer(cap * strcap); den = ctypes.create_string_buffer(cap * strcap)
        cnt = L.or_anchor_records(int(k_max), int(anchor), float(radius), cap, m.ctypes.data, sup.ctypes.data,
                                  rank.ctypes.data, nint.ctypes.data, inter.ctypes.data, deg.ctypes.data,
                                  shell.ctypes.data, num, den, strcap)
        if cnt < 0:
            raise OracleError(f"anchor records failed with code {cnt}")
        if cnt <= cap:
            break
        cap = int(cnt) + 16
    out = []
    for i in range(cnt):
        k = int(m[i])
        nu = num.raw[i * strcap:(i + 1) * strcap].split(b"\0", 1)[0]
        de = den.raw[i * strcap:(i + 1) * strcap].split(b"\0", 1)[0]
        out.append(CatalogRecord(
            support=tuple(int(x) for x in sup[i, :k]),
            interior=tuple(int(x) for x in inter[i, : int(nint[i])]),
            rank=int(rank[i]),
            level=Fraction(int(nu), int(de)),
            degenerate=bool(deg[i]),
            shell_count=int(shell[i]),
            shell_witness=-1,
        ))
    return out


def emst_graph(canonical_points, cand_edges) -> np.ndarray:
    """Exact minimum spanning forest under kappa = (d^2, u, v) of the given
    candidate edges (rows (u, v)), in kappa order."""
    L = lib()
    if not hasattr(L.or_emst_graph, "_typed"):
        L.or_emst_graph.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_void_p]
        L.or_emst_graph.restype = ctypes.c_int64
        L.or_emst_graph._typed = True
    a = _pts(canonical_points)
    c = np.ascontiguousarray(np.asarray(cand_edges, dtype=np.int64).reshape(-1, 2))
    out = np.zeros((max(len(a) - 1, 1), 2), dtype=np.int64)
    k = L.or_emst_graph(a.ctypes.data, len(a), c.ctypes.data, len(c), out.ctypes.data)
    if k < 0:
        raise OracleError(f"emst_graph failed with code {k}")
    return out[:k]


def knn_partition(canonical_points, num, den: int, k: int, exclusions=()) -> dict:
    """Exact top-k partition of the rational query num / den (pure Python,
    exact rationals; small clouds), following the reference TopKPartition
    (morsehgp3d/src/cpu/spatial/brute_force.cpp:320-370): cutoff = exact k-th
    smallest squared distance over the eligible points, strict = the points
    below it by (distance, id), shell = every eligible point at it, choice =
    strict plus the smallest shell ids up to k, ascending."""
    a = _pts(canonical_points)
    q = [Fraction(int(x), int(den)) for x in num]
    ex = set(int(x) for x in exclusions)
    d = []
    for i, p in enumerate(a):
        if i in ex:
            continue
        d.append((sum((Fraction(float(p[t])) - q[t]) ** 2 for t in range(3)), i))
    if not 1 <= k <= len(d):
        raise OracleError("requested rank out of range")
    d.sort()
    cut = d[k - 1][0]
    strict = [i for v, i in d if v < cut]
    shell = sorted(i for v, i in d if v == cut)
    choice = sorted(strict + shell[: k - len(strict)])
    return {"cutoff_sq": cut, "strict": strict, "shell": shell, "choice": choice}
