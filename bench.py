"""Benchmark: p95 warm_e2e of the MorseHGP3D hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): a 50 000-point
jittered cubic lattice (37^3 sites, seeded subsample, Gaussian jitter
sigma = 0.1 x spacing), binary64 input, K_max = 10 (supports of closed rank
<= 11).  One step is one warm end-to-end run: host-to-device upload,
canonical PointIds and exact Morton LBVH, exact EMST, the critical-sphere
catalogue, the hgp_reduced tower for orders 1..10, and the tower copied back
into host memory.  The context (HIP runtime, allocator, device buffers) is
initialised by the warm-up steps, as warm_e2e prescribes.

Prints one JSON line (see the field notes in the comments below).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "e-hgp_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import mhgp  # noqa: E402
import workloads  # noqa: E402

METRIC = "p95 warm_e2e ms (50k pts, Kmax=10); served clouds/s + p99 vs plain path"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def run_step(ctx, pts, k_max):
    ctx.upload(pts)
    edges = ctx.emst()
    info = ctx.supports(k_max, fetch=False)
    tower = ctx.tower(k_max, fetch=True)
    return edges, info, tower


def cpu_baseline(pts, k_max, budget_s=20.0):
    """The exact CPU oracle on a bounded sample of anchors of the same cloud,
    scaled to the whole cloud (the oracle enumerates each anchor's supports
    exhaustively inside a fixed neighbourhood, single thread)."""
    import oracle

    canon = oracle.canonicalize(pts)
    n = len(canon)
    rng = np.random.default_rng(0)
    anchors = rng.choice(n, size=n, replace=False)
    done, t0 = 0, time.perf_counter()
    for a in anchors:
        oracle.anchor_catalog(canon, k_max, int(a), radius=3.0)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    elapsed = time.perf_counter() - t0
    per_anchor_ms = 1e3 * elapsed / done
    return {
        "value": per_anchor_ms * n,
        "unit": "ms",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} random anchors of the 50k lattice: exhaustive exact supports "
                  f"(anchor = least id) within radius 3 spacings, {per_anchor_ms:.1f} ms/anchor, "
                  f"scaled x{n}",
    }


def copy_calibration(mib: int) -> None:
    """A plain elementwise kernel of known traffic (reads and writes `mib` MiB
    of float32 once each), profiled beside the catalogue kernel to scale the
    FETCH_SIZE / WRITE_SIZE counters into bytes."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 3}))


def served_stream(k_values=(4, 10), count=48, rate_mult=4.0, slots=4, seed=7):
    """Poisson stream of clouds (sizes log-uniform in [2 000, 50 000]) through
    the continuous-batching server, against the same clouds run one after
    another on one context (the plain path).  The offered rate is rate_mult
    times the plain path's rate; served clouds/s is the number of clouds over
    the span from the first arrival to the last completion."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(2_000), np.log(50_000), size=count)).astype(int)
    ks = rng.choice(np.asarray(k_values), size=count)
    clouds = [workloads.jittered_lattice(int(m), side=int(np.ceil((m * 1.3) ** (1 / 3))), sigma=0.1,
                                         seed=1000 + i) for i, m in enumerate(sizes)]
    ctx = mhgp.Context(0)
    run_step(ctx, clouds[0], int(ks[0]))  # warm
    t0 = time.perf_counter()
    plain_lat = []
    for pts, k in zip(clouds, ks):
        a = time.perf_counter()
        run_step(ctx, pts, int(k))
        plain_lat.append(1e3 * (time.perf_counter() - a))
    plain_s = time.perf_counter() - t0
    ctx.close()
    plain_rate = count / plain_s
    rate = rate_mult * plain_rate
    gaps = rng.exponential(1.0 / rate, size=count)
    srv = mhgp.Server(0, slots=slots)
    warm = srv.submit(clouds[0], int(ks[0]))
    srv.wait(warm)
    srv.release(warm)
    t0 = time.perf_counter()
    tickets, due = [], t0
    for pts, k, g in zip(clouds, ks, gaps):
        due += g
        while time.perf_counter() < due:
            time.sleep(min(0.001, max(0.0, due - time.perf_counter())))
        tickets.append(srv.submit(pts, int(k)))
    lat = []
    for t in tickets:
        r = srv.wait(t)
        if r.code != 0:
            raise mhgp.MhgpError(r.code, "served request failed")
        lat.append(r.latency_ms)
        srv.release(t)
    served_s = time.perf_counter() - t0
    srv.close()
    return {
        "clouds": count, "slots": slots, "offered_rate_hz": round(rate, 2),
        "served_clouds_per_s": round(count / served_s, 3),
        "plain_clouds_per_s": round(plain_rate, 3),
        "served_p50_ms": round(float(np.percentile(lat, 50)), 2),
        "served_p99_ms": round(float(np.percentile(lat, 99)), 2),
        "plain_p99_ms": round(float(np.percentile(plain_lat, 99)), 2),
    }


def batched(count=256, slots=4, plain_subset=32):
    """256 clouds (sizes log-uniform in [1 000, 50 000], each a jittered
    lattice or a 5-20 blob anisotropic mixture) submitted at once to the
    server; the plain path runs the first clouds one after another."""
    sizes = workloads.batch_sizes(count)
    clouds = [workloads.batch_cloud(int(m), seed=2000 + i) for i, m in enumerate(sizes)]
    codes = {}
    ctx = mhgp.Context(0)
    t0 = time.perf_counter()
    for i, pts in enumerate(clouds[:plain_subset]):
        try:
            run_step(ctx, pts, 10)
        except mhgp.MhgpError:
            pass
        if (i + 1) % 8 == 0:
            print(f"batched, plain path: {i + 1}/{plain_subset} done", file=sys.stderr, flush=True)
    plain_s = time.perf_counter() - t0
    ctx.close()
    srv = mhgp.Server(0, slots=slots)
    t0 = time.perf_counter()
    tickets = [srv.submit(p, 10) for p in clouds]
    lat = []
    for i, t in enumerate(tickets):
        r = srv.wait(t)
        codes[int(r.code)] = codes.get(int(r.code), 0) + 1
        if (i + 1) % 32 == 0:
            print(f"batched: {i + 1}/{count} done", file=sys.stderr, flush=True)
        lat.append(r.latency_ms)
        srv.release(t)
    served_s = time.perf_counter() - t0
    srv.close()
    return {
        "clouds": count, "points": int(sizes.sum()), "slots": slots,
        "served_clouds_per_s": round(count / served_s, 3),
        "plain_clouds_per_s_first_%d" % plain_subset: round(plain_subset / plain_s, 3),
        "result_codes": {str(k): v for k, v in sorted(codes.items())},
        "served_total_s": round(served_s, 2),
    }


def traffic_from_profile():
    path = os.path.join(ROOT, "profiles", "anchor_kernel_c2_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path, encoding="utf-8") as fh:
        return json.load(fh).get("bytes_per_launch")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--kmax", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--served", action="store_true", help="add the served-stream leg")
    ap.add_argument("--batched", action="store_true", help="add the 256-cloud batch leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain kernel of known traffic (counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    pts = workloads.jittered_lattice(args.n, side=37, sigma=0.1, seed=0)
    ctx = mhgp.Context(0)
    for _ in range(args.warmup):
        run_step(ctx, pts, args.kmax)
    ctx.stats(reset=True)
    times = []
    last = None
    for s in range(args.steps):
        t0 = time.perf_counter()
        last = run_step(ctx, pts, args.kmax)
        times.append(1e3 * (time.perf_counter() - t0))
        print(f"step {s}: {times[-1]:.1f} ms", file=sys.stderr, flush=True)
    st = ctx.stats()
    edges, info, tower = last
    times_a = np.asarray(times)
    p95 = float(np.percentile(times_a, 95))
    anchor_ms = st.anchor_ms / max(st.anchor_launches, 1)
    achieved = (st.anchor_bytes / max(st.anchor_launches, 1)) / (anchor_ms * 1e-3)
    line = {
        "metric": METRIC,
        "value": round(p95, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(times_a.mean()), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "jittered lattice 37^3 -> 50000 points, sigma 0.1 spacing, K_max 10, warm_e2e",
            "n": args.n,
            "k_max": args.kmax,
            "events": int(info.event_count),
            "diagnostics": int(info.diagnostic_count),
            "host_fallback_predicates": int(info.fallback_count),
            "tower_nodes": int(tower["info"].node_count),
            "descent_steps": int(tower["info"].descent_steps),
            "p50_ms": round(float(np.percentile(times_a, 50)), 3),
            "p99_ms": round(float(np.percentile(times_a, 99)), 3),
            "max_ms": round(float(times_a.max()), 3),
            "samples_ms": [round(t, 3) for t in times],
        },
        "roofline": {
            "kernel": "k_anchor_lds (critical-sphere catalogue, one workgroup per anchor)",
            "bound": "hbm",
            "achieved": round(achieved / 1e9, 3),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 6),
            "traffic": traffic_from_profile(),
            "avg_launch_ms": round(anchor_ms, 3),
            "algorithmic_bytes_per_launch": int(st.anchor_bytes / max(st.anchor_launches, 1)),
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(pts, args.kmax),
    }
    ctx.close()
    if args.served:
        line["config"]["served"] = served_stream()
    if args.batched:
        line["config"]["batched"] = batched()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
