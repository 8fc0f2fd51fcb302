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
FP64_PEAK = 78.6e12  # MI355X FP64 vector, FLOP/s
# The VALU's FP32 rate without packed FP32 instructions (which this build
# disables at compile time) is the same 64 FMA per CU and clock as FP64: the
# published 157.3 TFLOP/s FP32 figure counts packed pairs.
VALU_PEAK = 78.6e12


def run_step(ctx, pts, k_max):
    ctx.upload(pts)
    edges = ctx.emst()
    info = ctx.supports(k_max, fetch=False)
    tower = ctx.tower(k_max, fetch=True)
    return edges, info, tower


C1_N, C1_K = 256, 3


def c1_cloud():
    """The plain-method CPU reference configuration: 256 points i.i.d. uniform
    in [0, 1)^3 (seeded), K_max = 3."""
    return workloads.uniform_cloud(C1_N, seed=0)


def cpu_baseline_c1(pts, k_max=C1_K):
    """The exact CPU oracle (one thread) over the whole pipeline of the
    256-point cloud: canonical PointIds, exact Morton LBVH, exact EMST, the
    exhaustive exact catalogue (every anchor, every support of 2..4 points,
    exact classification against every point) and the direct hgp_reduced
    tower by exact facet descents."""
    import oracle
    from oracle import tower as otower

    t0 = time.perf_counter()
    canon = oracle.canonicalize(pts)
    oracle.lbvh_snapshot(canon)
    edges = oracle.emst(canon)
    # every sphere through a point of the unit cube with centre inside it has
    # radius below sqrt(3): radius 4 encloses every support and its interior
    recs = []
    for a in range(len(canon)):
        recs += oracle.anchor_records(canon, k_max, a, 4.0)
    forests = otower.direct_forests(canon, k_max, events=recs)
    elapsed = time.perf_counter() - t0
    return {
        "value": round(1e3 * elapsed, 1),
        "unit": "ms",
        "cores": 1,
        "kind": "port",
        "sample": f"one complete run of the 256-point uniform cloud at K_max={k_max} (canonical ids, LBVH, "
                  f"EMST, exhaustive exact catalogue of {len(recs)} records, direct tower), single thread",
    }, (canon, edges, recs, forests)


def c1_gpu(pts, oracle_result, steps=20, warmup=3, k_max=C1_K):
    """The same 256-point cloud on the GPU (warm_e2e), checked against the
    oracle's EMST, catalogue and tower."""
    from oracle import tower as otower

    canon, ref_edges, recs, forests = oracle_result
    ctx = mhgp.Context(0)
    for _ in range(warmup):
        run_step(ctx, pts, k_max)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        edges, _, tw = run_step(ctx, pts, k_max)
        times.append(1e3 * (time.perf_counter() - t0))
    cat = ctx.supports(k_max)
    ctx.close()
    ok = np.array_equal(edges.astype(np.int64), ref_edges)
    exact = {(r.support, r.interior): r.level for r in recs}
    mine = {(tuple(int(x) for x in cat.ids[i, : cat.size[i]]), cat.interior[i]) for i in range(len(cat.size))}
    ok &= mine == set(exact)
    level_of = {-1: 0}
    for i in range(len(cat.size)):
        level_of[int(cat.level_class[i])] = exact[(tuple(int(x) for x in cat.ids[i, : cat.size[i]]), cat.interior[i])]
    off = tw["child_offset"]
    for k in range(1, k_max + 1):
        idx = [i for i in range(len(tw["order"])) if tw["order"][i] == k]
        remap = {old: new for new, old in enumerate(idx)}
        nodes = [{"order": k, "kind": "merge" if tw["kind"][i] == 2 else "birth",
                  "level": level_of[int(tw["level_class"][i])],
                  "children": tuple(sorted(remap[int(c)] for c in tw["children"][off[i]:off[i + 1]])),
                  "covered": None} for i in idx]
        ok &= otower.signatures(nodes) == otower.signatures(forests[k])
    t = np.asarray(times)
    return {"gpu_p50_ms": round(float(np.percentile(t, 50)), 3), "gpu_p95_ms": round(float(np.percentile(t, 95)), 3),
            "steps": steps, "equal_to_oracle": bool(ok)}


def c2_cpu_extrapolated(pts, radii, k_max, budget_s=8.0):
    """The exhaustive exact oracle on a bounded sample of anchors of the
    50 000-point cloud (each anchor's records among the points within its
    certified radius), scaled to every anchor: an extrapolation, not a run."""
    import oracle

    canon = oracle.canonicalize(pts)
    n = len(canon)
    rng = np.random.default_rng(0)
    done, t0 = 0, time.perf_counter()
    for a in rng.choice(n, size=n, replace=False):
        oracle.anchor_records(canon, k_max, int(a), float(radii[a]))
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    per_anchor_ms = 1e3 * (time.perf_counter() - t0) / done
    return {"value": round(per_anchor_ms * n, 1), "unit": "ms", "cores": 1, "kind": "port",
            "extrapolated": True,
            "sample": f"{done} random anchors of the 50k lattice, exhaustive exact records within the certified "
                      f"radius, {per_anchor_ms:.1f} ms per anchor, scaled x{n}"}


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


SERVE_REGION_LIMIT = 640  # enumeration budget of the served and batched legs (plain path alike)


def plain_request(ctx, pts, k_max):
    """One request on the plain path, as the server runs it: catalogue, then
    the tower unless the catalogue is conditional; returns (code, status)."""
    try:
        ctx.upload(pts)
        ctx.emst()
        info = ctx.supports(k_max, fetch=False)
        if info.status != 0:
            return 0, 1
        ctx.tower(k_max, fetch=True)
        return 0, 0
    except mhgp.MhgpError as e:
        return e.code, 0


def c5_cloud(size: int, seed: int):
    """The served-stream generators: a jittered lattice, a 5-20 blob
    anisotropic mixture with 15 % uniform noise, or the same snapped to the
    2^-12 grid (chosen by seed)."""
    rng = workloads.SplitMix64(seed).numpy_rng()
    kind = int(rng.integers(3))
    if kind == 0:
        side = int(np.ceil((size * 1.3) ** (1 / 3)))
        return workloads.jittered_lattice(size, side=side, sigma=0.1, seed=seed)
    return workloads.gaussian_mixture(size, blobs=int(rng.integers(5, 21)), noise_fraction=0.15, box=10.0,
                                      snap_bits=12 if kind == 2 else None, seed=seed)


def served_stream(k_values=(4, 10), count=32, lo=2_000, hi=200_000, rate_mult=4.0, slots=4, seed=7):
    """Poisson stream of clouds (sizes log-uniform in [lo, hi], K_max drawn
    from k_values, lattice / mixture / snapped-mixture generators) through the
    continuous-batching server, against the same clouds run one after another
    on one context (the plain path).  The offered rate is rate_mult times the
    plain path's rate; served clouds/s is the number of clouds over the span
    from the first arrival to the last completion.  Results per request:
    exact, conditional (catalogue budget stop, no tower) or refused tower
    (extra-shell catalogue)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(int)
    ks = rng.choice(np.asarray(k_values), size=count)
    clouds = [c5_cloud(int(m), 1000 + i) for i, m in enumerate(sizes)]
    ctx = mhgp.Context(0)
    ctx.set_region_limit(SERVE_REGION_LIMIT)
    plain_request(ctx, clouds[0], int(ks[0]))  # warm
    t0 = time.perf_counter()
    plain_lat, plain_out = [], []
    for i, (pts, k) in enumerate(zip(clouds, ks)):
        a = time.perf_counter()
        plain_out.append(plain_request(ctx, pts, int(k)))
        plain_lat.append(1e3 * (time.perf_counter() - a))
        print(f"served, plain path: cloud {i + 1}/{count} ({len(pts)} points, K {k}) {plain_lat[-1]:.0f} ms",
              file=sys.stderr, flush=True)
    plain_s = time.perf_counter() - t0
    ctx.close()
    plain_rate = count / plain_s
    rate = rate_mult * plain_rate
    gaps = rng.exponential(1.0 / rate, size=count)
    srv = mhgp.Server(0, slots=slots, region_limit=SERVE_REGION_LIMIT)
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
    lat, outcome = [], {}
    for i, t in enumerate(tickets):
        r = srv.wait(t)
        key = "exact" if r.code == 0 and r.status == 0 else "conditional" if r.code == 0 else f"code {r.code}"
        outcome[key] = outcome.get(key, 0) + 1
        if (r.code, r.status) != plain_out[i]:
            raise mhgp.MhgpError(r.code, "a served request ended unlike the same request on the plain path")
        lat.append(r.latency_ms)
        if (i + 1) % 4 == 0:
            print(f"served: {i + 1}/{count} done", file=sys.stderr, flush=True)
        srv.release(t)
    served_s = time.perf_counter() - t0
    srv.close()
    return {
        "clouds": count, "sizes": [lo, hi], "points": int(sizes.sum()), "slots": slots,
        "region_limit": SERVE_REGION_LIMIT,
        "offered_rate_hz": round(rate, 2),
        "served_clouds_per_s": round(count / served_s, 3),
        "plain_clouds_per_s": round(plain_rate, 3),
        "served_vs_plain": round(plain_s / served_s, 3),
        "served_p50_ms": round(float(np.percentile(lat, 50)), 2),
        "served_p95_ms": round(float(np.percentile(lat, 95)), 2),
        "served_p99_ms": round(float(np.percentile(lat, 99)), 2),
        "plain_p50_ms": round(float(np.percentile(plain_lat, 50)), 2),
        "plain_p99_ms": round(float(np.percentile(plain_lat, 99)), 2),
        "outcomes": outcome,
    }


def batched(count=256, slots=4, plain_subset=32):
    """256 clouds (sizes log-uniform in [1 000, 50 000], each a jittered
    lattice or a 5-20 blob anisotropic mixture, K_max 10) submitted at once to
    the server; the plain path runs the first clouds one after another on one
    context, and every served result among them must end as it does there."""
    sizes = workloads.batch_sizes(count)
    clouds = [workloads.batch_cloud(int(m), seed=2000 + i) for i, m in enumerate(sizes)]
    ctx = mhgp.Context(0)
    ctx.set_region_limit(SERVE_REGION_LIMIT)
    plain = []
    t0 = time.perf_counter()
    for i, pts in enumerate(clouds[:plain_subset]):
        plain.append(plain_request(ctx, pts, 10))
        if (i + 1) % 8 == 0:
            print(f"batched, plain path: {i + 1}/{plain_subset} done", file=sys.stderr, flush=True)
    plain_s = time.perf_counter() - t0
    ctx.close()
    srv = mhgp.Server(0, slots=slots, region_limit=SERVE_REGION_LIMIT)
    t0 = time.perf_counter()
    tickets = [srv.submit(p, 10) for p in clouds]
    outcomes = {}
    for i, t in enumerate(tickets):
        r = srv.wait(t)
        key = "exact" if r.code == 0 and r.status == 0 else "conditional" if r.code == 0 else f"code {r.code}"
        outcomes[key] = outcomes.get(key, 0) + 1
        if i < plain_subset and (r.code, r.status) != plain[i]:
            raise mhgp.MhgpError(r.code, "a served cloud ended unlike the same cloud on the plain path")
        if (i + 1) % 8 == 0:
            print(f"batched: {i + 1}/{count} done", file=sys.stderr, flush=True)
        srv.release(t)
    served_s = time.perf_counter() - t0
    srv.close()
    return {
        "clouds": count, "points": int(sizes.sum()), "slots": slots, "region_limit": SERVE_REGION_LIMIT,
        "served_clouds_per_s": round(count / served_s, 3),
        "plain_clouds_per_s_first_%d" % plain_subset: round(plain_subset / plain_s, 3),
        "outcomes": outcomes, "served_total_s": round(served_s, 2),
    }


def traffic_from_profile():
    path = os.path.join(ROOT, "profiles", "anchor_kernel_c2_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path, encoding="utf-8") as fh:
        return json.load(fh).get("bytes_per_launch")


def mixture_catalogues() -> None:
    """Catalogue time and record counts of three served-stream mixtures
    (6 467, 111 721 and 195 898 points; K_max 10, enumeration budget 640):
    the workload of the big-anchor paths' kernel statistics."""
    ctx = mhgp.Context(0)
    ctx.set_region_limit(SERVE_REGION_LIMIT)
    for n, seed in [(6467, 1012), (111721, 1005), (195898, 1017)]:
        pts = c5_cloud(n, seed)
        a = time.perf_counter()
        ctx.upload(pts)
        ctx.emst()
        info = ctx.supports(10, fetch=False)
        print(f"mixture n={n}: catalogue {1e3 * (time.perf_counter() - a):.0f} ms status {info.status} "
              f"records {info.event_count} open {info.anchors_open} largest {info.largest_open_region}", flush=True)
    ctx.close()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--kmax", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-cloud batch leg")
    ap.add_argument("--mixture-catalogues", action="store_true",
                    help="only time the catalogues of three served-stream mixtures (big-anchor paths)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain kernel of known traffic (counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.mixture_catalogues:
        mixture_catalogues()
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
    radii = ctx.radii()
    edges, info, tower = last
    # FP64 operations of the anchor kernel: one more run, untimed, with the
    # kernel's count switched on (the timed runs skip it); the work is the same
    ctx.count_flops(True)
    ctx.stats(reset=True)
    run_step(ctx, pts, args.kmax)
    counted = ctx.stats(reset=True)
    ctx.count_flops(False)
    times_a = np.asarray(times)
    p95 = float(np.percentile(times_a, 95))
    launches = max(st.anchor_launches, 1)
    anchor_ms = st.anchor_ms / launches
    flops = counted.anchor_flops / max(counted.anchor_launches, 1)
    flops32 = counted.anchor_flops32 / max(counted.anchor_launches, 1)
    bytes_alg = st.anchor_bytes / launches
    tflops = (flops + flops32) / (anchor_ms * 1e-3) / 1e12
    gbs = bytes_alg / (anchor_ms * 1e-3) / 1e9
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
            "status": "exact" if info.status == 0 and tower["info"].status == 0 else "conditional",
            "tower_nodes": int(tower["info"].node_count),
            "descent_steps": int(tower["info"].descent_steps),
            "p50_ms": round(float(np.percentile(times_a, 50)), 3),
            "p99_ms": round(float(np.percentile(times_a, 99)), 3),
            "max_ms": round(float(times_a.max()), 3),
            "samples_ms": [round(t, 3) for t in times],
        },
        # The catalogue anchor kernel is bound by vector work (FP32 tests with
        # FP64 for what they leave open, divergent lanes), not by HBM: its
        # roofline is the VALU's FP32 (unpacked) / FP64 peak, with the HBM
        # figures beside it.
        "roofline": {
            "kernel": "k_anchor_lds (critical-sphere catalogue, one workgroup per anchor)",
            "bound": "valu",
            "achieved": round(tflops, 4),
            "peak": VALU_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": round(tflops * 1e12 / VALU_PEAK, 5),
            "traffic": traffic_from_profile(),
            "avg_launch_ms": round(anchor_ms, 3),
            "fp64_ops_per_launch": int(flops),
            "fp32_ops_per_launch": int(flops32),
            "hbm": {"algorithmic_bytes_per_launch": int(bytes_alg), "achieved": round(gbs, 3),
                    "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(gbs * 1e9 / HBM_PEAK, 6)},
        },
        "cpu_baseline": None,
    }
    ctx.close()
    if not args.no_cpu_baseline:
        c1 = c1_cloud()
        base, ref = cpu_baseline_c1(c1)
        base["gpu_same_cloud"] = c1_gpu(c1, ref)
        base["c2_extrapolated"] = c2_cpu_extrapolated(pts, radii, args.kmax)
        line["cpu_baseline"] = base
        print(f"cpu baseline: {base['value']} ms (256 points), gpu {base['gpu_same_cloud']}", file=sys.stderr,
              flush=True)
    if not args.no_served:
        line["config"]["served"] = served_stream()
    if not args.no_batched:
        line["config"]["batched"] = batched()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
