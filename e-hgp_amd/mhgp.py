"""Python binding of libmhgp.so, the MI355X MorseHGP3D backend (ctypes over the C ABI).

The compute runs in the HIP library; this module only moves arrays across
the C ABI declared in include/mhgp.h.  There is no CPU fallback: if the
library is missing or no GPU is visible, every call raises.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmhgp.so")
_lib = None


class MhgpError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"mhgp error {code}: {message}")
        self.code = code


E_ARGUMENT, E_DUPLICATE, E_NONFINITE, E_DEVICE, E_POISONED, E_STATE = -1, -2, -3, -4, -5, -6
E_CAPACITY, E_UNSUPPORTED, E_INTERNAL = -7, -8, -9


class CatalogInfo(ctypes.Structure):
    _fields_ = [("event_count", ctypes.c_int64), ("diagnostic_count", ctypes.c_int64),
                ("candidate_count", ctypes.c_int64), ("fallback_count", ctypes.c_int64),
                ("anchors_widened", ctypes.c_int64), ("s_max", ctypes.c_int32),
                ("status", ctypes.c_int32), ("anchors_open", ctypes.c_int64),
                ("largest_open_region", ctypes.c_int64)]


class TowerInfo(ctypes.Structure):
    _fields_ = [("node_count", ctypes.c_int64), ("child_count", ctypes.c_int64),
                ("descent_steps", ctypes.c_int64), ("orders", ctypes.c_int32),
                ("status", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("anchor_ms", ctypes.c_double), ("anchor_launches", ctypes.c_int64),
                ("anchor_bytes", ctypes.c_int64), ("arm_ms", ctypes.c_double),
                ("arm_launches", ctypes.c_int64), ("arm_bytes", ctypes.c_int64),
                ("gathered_points", ctypes.c_int64), ("events", ctypes.c_int64),
                ("anchor_flops", ctypes.c_int64), ("anchor_flops32", ctypes.c_int64)]


class RequestResult(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("status", ctypes.c_int32),
                ("event_count", ctypes.c_int64), ("node_count", ctypes.c_int64),
                ("child_count", ctypes.c_int64), ("queue_ms", ctypes.c_double),
                ("service_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("clouds", ctypes.c_int64), ("groups", ctypes.c_int64), ("anchors", ctypes.c_int64),
                ("anchor_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MhgpError(-9, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        p, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.mhgp_last_error.restype = ctypes.c_char_p
        L.mhgp_ctx_create.argtypes = [i32, i64, ctypes.POINTER(p)]
        L.mhgp_ctx_destroy.argtypes = [p]
        L.mhgp_ctx_device_bytes.argtypes = [p]
        L.mhgp_ctx_device_bytes.restype = i64
        L.mhgp_cloud_upload.argtypes = [p, p, i64]
        L.mhgp_cloud_upload_device.argtypes = [p, p, i64]
        L.mhgp_point_count.argtypes = [p]
        L.mhgp_point_count.restype = i64
        L.mhgp_canonical_points.argtypes = [p, p, p]
        L.mhgp_lbvh_snapshot.argtypes = [p, p, p, p]
        L.mhgp_knn.argtypes = [p, i32, p]
        L.mhgp_lbvh_certify.argtypes = [p, i64, p, p, p]
        L.mhgp_emst.argtypes = [p, p]
        L.mhgp_knn_partition.argtypes = [p, p, p, i64, i32, p, p, ctypes.c_uint32, p, p, p, p, p, p, p]
        L.mhgp_supports.argtypes = [p, i32, ctypes.POINTER(CatalogInfo)]
        L.mhgp_supports_fetch.argtypes = [p, p, p, p, p, p]
        L.mhgp_supports_radii.argtypes = [p, p]
        L.mhgp_supports_diagnostics.argtypes = [p, p, p, p]
        L.mhgp_tower.argtypes = [p, i32, ctypes.POINTER(TowerInfo)]
        L.mhgp_tower_fetch.argtypes = [p, p, p, p, p, p, p]
        L.mhgp_tower_vertical.argtypes = [p, p]
        L.mhgp_tower_cut.argtypes = [p, i32, i64, i32, p, p, p]
        L.mhgp_ctx_stats.argtypes = [p, ctypes.POINTER(Stats), i32]
        L.mhgp_ctx_count_flops.argtypes = [p, i32]
        L.mhgp_ctx_set_region_limit.argtypes = [p, i64]
        L.mhgp_server_set_region_limit.argtypes = [p, i64]
        L.mhgp_closed_ball.argtypes = [p, p, p, i64, ctypes.c_uint32, p, p, p]
        L.mhgp_closed_ball_noalloc.argtypes = [p, p, p, i64, ctypes.c_uint32, p, p, p]
        L.mhgp_closed_ball_noalloc_scratch_bytes.argtypes = []
        L.mhgp_closed_ball_noalloc_scratch_bytes.restype = i64
        L.mhgp_server_create.argtypes = [i32, i32, i64, ctypes.POINTER(p)]
        L.mhgp_server_submit.argtypes = [p, p, i64, i32, ctypes.POINTER(i64)]
        L.mhgp_server_wait.argtypes = [p, i64, ctypes.POINTER(RequestResult)]
        L.mhgp_server_fetch.argtypes = [p, i64, p, p, p, p, p, p]
        L.mhgp_server_release.argtypes = [p, i64]
        L.mhgp_server_destroy.argtypes = [p]
        L.mhgp_batch_run.argtypes = [i32, i64, p, p, p, i32, i64, ctypes.POINTER(p)]
        L.mhgp_batch_info_get.argtypes = [p, ctypes.POINTER(BatchInfo)]
        L.mhgp_batch_result.argtypes = [p, i64, ctypes.POINTER(RequestResult)]
        L.mhgp_batch_fetch.argtypes = [p, i64, p, p, p, p, p, p]
        L.mhgp_batch_destroy.argtypes = [p]
        L.mhgp_batch_destroy.restype = None
        L.mhgp_server_destroy.restype = None
        _lib = L
    return _lib


EXPORTED = [
    "mhgp_last_error", "mhgp_ctx_create", "mhgp_ctx_destroy", "mhgp_ctx_device_bytes",
    "mhgp_cloud_upload", "mhgp_cloud_upload_device", "mhgp_point_count",
    "mhgp_canonical_points", "mhgp_lbvh_snapshot", "mhgp_lbvh_certify", "mhgp_knn", "mhgp_knn_partition", "mhgp_emst", "mhgp_supports",
    "mhgp_supports_fetch", "mhgp_supports_radii", "mhgp_supports_diagnostics", "mhgp_tower", "mhgp_tower_fetch", "mhgp_tower_vertical", "mhgp_tower_cut", "mhgp_ctx_stats", "mhgp_ctx_count_flops", "mhgp_ctx_set_region_limit",
    "mhgp_closed_ball", "mhgp_closed_ball_noalloc", "mhgp_closed_ball_noalloc_scratch_bytes",
    "mhgp_server_create", "mhgp_server_set_region_limit", "mhgp_server_submit", "mhgp_server_wait", "mhgp_server_fetch",
    "mhgp_server_release", "mhgp_server_destroy", "mhgp_batch_run", "mhgp_batch_info_get", "mhgp_batch_result",
    "mhgp_batch_fetch", "mhgp_batch_destroy",
]


def lbvh_certify(canonical_xyz, leaves, nodes, counters) -> None:
    """Exact host re-certification of a snapshot (raises MhgpError with the
    first violation); needs no GPU."""
    xyz = np.ascontiguousarray(np.asarray(canonical_xyz, dtype=np.float64).reshape(-1, 3))
    lv = np.ascontiguousarray(np.asarray(leaves, dtype=np.uint64))
    nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.uint64))
    ct = np.ascontiguousarray(np.asarray(counters, dtype=np.uint64))
    _check(lib().mhgp_lbvh_certify(_ptr(xyz), len(xyz), _ptr(lv), _ptr(nd), _ptr(ct)))


def _check(rc: int) -> None:
    if rc != 0:
        raise MhgpError(rc, lib().mhgp_last_error().decode())


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class Catalog:
    ids: np.ndarray        # (E, 4) uint32
    size: np.ndarray       # (E,) support size
    rank: np.ndarray       # (E,) closed rank
    interior: list         # list of tuples
    degenerate: np.ndarray
    level: np.ndarray      # FP64 nearest value
    level_class: np.ndarray
    info: CatalogInfo


class Context:
    """One cloud on one GPU (owns a HIP stream and the device state)."""

    def __init__(self, device: int = 0, capacity_bytes: int = 0):
        h = ctypes.c_void_p()
        _check(lib().mhgp_ctx_create(device, capacity_bytes, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            lib().mhgp_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self) -> int:
        return int(lib().mhgp_point_count(self._h))

    def device_bytes(self) -> int:
        return int(lib().mhgp_ctx_device_bytes(self._h))

    def stats(self, reset: bool = False) -> Stats:
        st = Stats()
        _check(lib().mhgp_ctx_stats(self._h, ctypes.byref(st), 1 if reset else 0))
        return st

    def count_flops(self, on: bool) -> None:
        """FP64 and FP32 operation counts of the anchor kernel (Stats.anchor_flops, anchor_flops32)."""
        _check(lib().mhgp_ctx_count_flops(self._h, 1 if on else 0))

    def set_region_limit(self, limit: int) -> None:
        """Enumeration budget: the largest candidate region enumerated (640..2048)."""
        _check(lib().mhgp_ctx_set_region_limit(self._h, int(limit)))

    def upload(self, points) -> None:
        a = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        _check(lib().mhgp_cloud_upload(self._h, _ptr(a), len(a)))

    def upload_device(self, dev_ptr: int, n: int) -> None:
        _check(lib().mhgp_cloud_upload_device(self._h, dev_ptr, n))

    def canonical_points(self):
        n = self.n
        xyz = np.zeros((n, 3), dtype=np.float64)
        src = np.zeros(n, dtype=np.uint32)
        _check(lib().mhgp_canonical_points(self._h, _ptr(xyz), _ptr(src)))
        return xyz, src

    def lbvh_snapshot(self):
        n = self.n
        leaves = np.zeros((n, 2), dtype=np.uint64)
        nodes = np.zeros((2 * n - 1, 10), dtype=np.uint64)
        counters = np.zeros(5, dtype=np.uint64)
        _check(lib().mhgp_lbvh_snapshot(self._h, _ptr(leaves), _ptr(nodes), _ptr(counters)))
        return leaves, nodes, counters

    def knn(self, k: int) -> np.ndarray:
        out = np.zeros((self.n, k), dtype=np.uint32)
        _check(lib().mhgp_knn(self._h, k, _ptr(out)))
        return out

    def knn_partition(self, num, den, k: int, exclusions=None, shell_cap: int = 64):
        """Exact top-k partitions of rational queries num[q] / den[q]
        (integer numerators, positive integer denominators).  Returns a list
        of dicts: cutoff_sq, cutoff_id, strict (ids by exact distance), shell
        (complete ids ascending), shell_count, choice (ascending)."""
        num = np.ascontiguousarray(np.asarray(num, dtype=np.int64).reshape(-1, 3))
        den = np.ascontiguousarray(np.asarray(den, dtype=np.int64).reshape(-1))
        q = len(num)
        ex = exclusions or [[] for _ in range(q)]
        off = np.zeros(q + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(e) for e in ex])
        ids = np.ascontiguousarray(np.asarray([x for e in ex for x in e] or [0], dtype=np.uint32))
        choice = np.zeros((max(q, 1), k), dtype=np.uint32)
        strict = np.zeros((max(q, 1), k), dtype=np.uint32)
        ns = np.zeros(max(q, 1), dtype=np.uint32)
        shell = np.zeros((max(q, 1), max(shell_cap, 1)), dtype=np.uint32)
        nsh = np.zeros(max(q, 1), dtype=np.uint32)
        cid = np.zeros(max(q, 1), dtype=np.uint32)
        csq = np.zeros(max(q, 1), dtype=np.float64)
        _check(lib().mhgp_knn_partition(self._h, _ptr(num), _ptr(den), q, k, _ptr(ids), _ptr(off), shell_cap,
                                        _ptr(choice), _ptr(strict), _ptr(ns), _ptr(shell), _ptr(nsh), _ptr(cid),
                                        _ptr(csq)))
        return [{"cutoff_sq": float(csq[i]), "cutoff_id": int(cid[i]),
                 "strict": [int(x) for x in strict[i, : ns[i]]], "shell_count": int(nsh[i]),
                 "shell": [int(x) for x in shell[i, : min(int(nsh[i]), shell_cap)]],
                 "choice": [int(x) for x in choice[i]]} for i in range(q)]

    def emst(self) -> np.ndarray:
        out = np.zeros((max(self.n - 1, 0), 2), dtype=np.uint32)
        _check(lib().mhgp_emst(self._h, _ptr(out)))
        return out

    def closed_ball(self, supports, cap: int = 16, noalloc: bool = False):
        """Interior ids (complete when the count <= cap), interior counts and
        shell counts of the circumspheres of the given supports (lists of 1..4
        canonical PointIds).  noalloc: the entry point that allocates no
        device memory (same results)."""
        q = len(supports)
        ids = np.full((max(q, 1), 4), 0xFFFFFFFF, dtype=np.uint32)
        sizes = np.zeros(max(q, 1), dtype=np.uint8)
        for i, sup in enumerate(supports):
            ids[i, : len(sup)] = sup
            sizes[i] = len(sup)
        inter = np.zeros((max(q, 1), max(cap, 1)), dtype=np.uint32)
        nin = np.zeros(max(q, 1), dtype=np.uint32)
        nsh = np.zeros(max(q, 1), dtype=np.uint32)
        fn = lib().mhgp_closed_ball_noalloc if noalloc else lib().mhgp_closed_ball
        _check(fn(self._h, _ptr(ids), _ptr(sizes), q, cap, _ptr(inter), _ptr(nin), _ptr(nsh)))
        return inter[:q, :cap], nin[:q], nsh[:q]

    def supports(self, k_max: int, fetch: bool = True):
        info = CatalogInfo()
        _check(lib().mhgp_supports(self._h, k_max, ctypes.byref(info)))
        if not fetch:
            return info
        e = info.event_count + info.diagnostic_count
        smax = max(info.s_max, 2)
        ids = np.zeros((e, 4), dtype=np.uint32)
        meta = np.zeros(e, dtype=np.uint32)
        interior = np.zeros((e, smax - 1), dtype=np.uint32)
        level = np.zeros(e, dtype=np.float64)
        cls = np.zeros(e, dtype=np.int64)
        _check(lib().mhgp_supports_fetch(self._h, _ptr(ids), _ptr(meta), _ptr(interior),
                                         _ptr(level), _ptr(cls)))
        size = meta & 0xFF
        rank = (meta >> 8) & 0xFF
        nint = (meta >> 16) & 0xFF
        deg = (meta >> 24) & 0xFF
        inter = [tuple(int(x) for x in interior[i, : nint[i]]) for i in range(e)]
        return Catalog(ids, size, rank, inter, deg.astype(bool), level, cls, info)

    def supports_raw(self, k_max: int):
        """The catalogue as flat arrays (no per-record Python objects): ids
        (E, 4), meta (E,), interior (E, s_max - 1), level (E,), level_class
        (E,) and the info block."""
        info = CatalogInfo()
        _check(lib().mhgp_supports(self._h, k_max, ctypes.byref(info)))
        e = info.event_count + info.diagnostic_count
        smax = max(info.s_max, 2)
        out = {"ids": np.zeros((e, 4), dtype=np.uint32), "meta": np.zeros(e, dtype=np.uint32),
               "interior": np.zeros((e, smax - 1), dtype=np.uint32), "level": np.zeros(e, dtype=np.float64),
               "level_class": np.zeros(e, dtype=np.int64), "info": info}
        _check(lib().mhgp_supports_fetch(self._h, _ptr(out["ids"]), _ptr(out["meta"]), _ptr(out["interior"]),
                                         _ptr(out["level"]), _ptr(out["level_class"])))
        return out

    def diagnostics(self, count: int):
        """Shell count, extra-shell witness (-1: none) and exterior count of
        each of the last catalogue's `count` records (fetch order)."""
        sc = np.zeros(max(count, 1), dtype=np.uint32)
        wit = np.zeros(max(count, 1), dtype=np.int64)
        ext = np.zeros(max(count, 1), dtype=np.int64)
        _check(lib().mhgp_supports_diagnostics(self._h, _ptr(sc), _ptr(wit), _ptr(ext)))
        return sc[:count], wit[:count], ext[:count]

    def radii(self) -> np.ndarray:
        """Per-PointId enumeration radius of the last catalogue."""
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        _check(lib().mhgp_supports_radii(self._h, _ptr(out)))
        return out[: self.n]

    def tower_cut(self, order: int, level_class: int, closed: bool = True, node_count: int = 0):
        """Roots of `order`'s components at the cut, and for order one the
        root of every PointId (-1: not yet born)."""
        roots = np.zeros(max(node_count, 1), dtype=np.int64)
        cnt = ctypes.c_int64()
        labels = np.zeros(max(self.n, 1), dtype=np.int64) if order == 1 else None
        _check(lib().mhgp_tower_cut(self._h, order, level_class, int(closed), _ptr(roots) if node_count else None,
                                    ctypes.byref(cnt), _ptr(labels) if labels is not None else None))
        return roots[: cnt.value] if node_count else None, labels[: self.n] if labels is not None else None

    def tower(self, k_max: int, fetch: bool = True):
        info = TowerInfo()
        _check(lib().mhgp_tower(self._h, k_max, ctypes.byref(info)))
        if not fetch:
            return info
        m = info.node_count
        order = np.zeros(m, dtype=np.int32)
        kind = np.zeros(m, dtype=np.int32)
        cls = np.zeros(m, dtype=np.int64)
        level = np.zeros(m, dtype=np.float64)
        off = np.zeros(m + 1, dtype=np.int64)
        kids = np.zeros(max(info.child_count, 1), dtype=np.int64)
        _check(lib().mhgp_tower_fetch(self._h, _ptr(order), _ptr(kind), _ptr(cls), _ptr(level),
                                      _ptr(off), _ptr(kids)))
        vert = np.zeros(max(m, 1), dtype=np.int64)
        _check(lib().mhgp_tower_vertical(self._h, _ptr(vert)))
        return {"order": order, "kind": kind, "level_class": cls, "level": level,
                "child_offset": off, "children": kids[: info.child_count], "vertical": vert[:m], "info": info}


def _tower_arrays(node_count: int, child_count: int):
    return (np.zeros(node_count, dtype=np.int32), np.zeros(node_count, dtype=np.int32),
            np.zeros(node_count, dtype=np.int64), np.zeros(node_count, dtype=np.float64),
            np.zeros(node_count + 1, dtype=np.int64), np.zeros(max(child_count, 1), dtype=np.int64))


class Server:
    """Continuous-batching server on one GPU: worker slots with their own
    streams, memory admission under a cap (default 80% of HBM)."""

    def __init__(self, device: int = 0, slots: int = 4, memory_cap_bytes: int = 0, region_limit: int = 0):
        h = ctypes.c_void_p()
        _check(lib().mhgp_server_create(device, slots, memory_cap_bytes, ctypes.byref(h)))
        self._h = h
        if region_limit:
            _check(lib().mhgp_server_set_region_limit(self._h, int(region_limit)))

    def close(self) -> None:
        if self._h:
            lib().mhgp_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, points: np.ndarray, k_max: int) -> int:
        xyz = np.ascontiguousarray(points, dtype=np.float64)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("points must be (n, 3)")
        t = ctypes.c_int64()
        _check(lib().mhgp_server_submit(self._h, _ptr(xyz), xyz.shape[0], k_max, ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int) -> RequestResult:
        r = RequestResult()
        _check(lib().mhgp_server_wait(self._h, ticket, ctypes.byref(r)))
        return r

    def fetch(self, ticket: int, result: RequestResult) -> dict:
        if result.code != 0:
            raise MhgpError(result.code, f"request {ticket} failed")
        order, kind, cls, level, off, kids = _tower_arrays(result.node_count, result.child_count)
        _check(lib().mhgp_server_fetch(self._h, ticket, _ptr(order), _ptr(kind), _ptr(cls),
                                       _ptr(level), _ptr(off), _ptr(kids)))
        return {"order": order, "kind": kind, "level_class": cls, "level": level,
                "child_offset": off, "children": kids[: result.child_count]}

    def release(self, ticket: int) -> None:
        _check(lib().mhgp_server_release(self._h, ticket))


def batch_run(clouds, k_max, group: int = 32, memory_cap_bytes: int = 0, device: int = 0):
    """Runs many clouds through mhgp_batch_run (groups side by side, one
    persistent anchor launch per group).  k_max: one value or one per cloud.
    Returns (results, towers, info): per cloud a RequestResult and its tower
    (dict as Context.tower without vertical maps, or None), and the batch's
    BatchInfo."""
    arrs = [np.ascontiguousarray(np.asarray(c, dtype=np.float64).reshape(-1, 3)) for c in clouds]
    count = len(arrs)
    ptrs = (ctypes.c_void_p * max(count, 1))(*[a.ctypes.data for a in arrs])
    n = np.ascontiguousarray([len(a) for a in arrs] or [0], dtype=np.int64)
    ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k_max, dtype=np.int32), (max(count, 1),)))
    h = ctypes.c_void_p()
    _check(lib().mhgp_batch_run(device, count, ctypes.cast(ptrs, ctypes.c_void_p), _ptr(n), _ptr(ks), group,
                                memory_cap_bytes, ctypes.byref(h)))
    try:
        info = BatchInfo()
        _check(lib().mhgp_batch_info_get(h, ctypes.byref(info)))
        results, towers = [], []
        for i in range(count):
            r = RequestResult()
            _check(lib().mhgp_batch_result(h, i, ctypes.byref(r)))
            results.append(r)
            if r.code != 0 or r.status != 0:
                towers.append(None)
                continue
            order, kind, cls, level, off, kids = _tower_arrays(r.node_count, r.child_count)
            _check(lib().mhgp_batch_fetch(h, i, _ptr(order), _ptr(kind), _ptr(cls), _ptr(level), _ptr(off),
                                          _ptr(kids)))
            towers.append({"order": order, "kind": kind, "level_class": cls, "level": level, "child_offset": off,
                           "children": kids[: r.child_count]})
        return results, towers, info
    finally:
        lib().mhgp_batch_destroy(h)
