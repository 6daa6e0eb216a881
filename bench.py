#!/usr/bin/env python3
"""Benchmark of the exact grid min cut on one MI355X.

Workload (one "step"): the plain GPU push-relabel solve of the 256x256x256 6-connected synthetic
instance (blob volume + Gaussian noise, int32 t-links and n-links, seed 0), from capacities
already resident in HBM to the exact max-flow value and the canonical cut S_min resident in HBM.
Prints one JSON line with the throughput (Mvoxels/s), the roofline of the dominant kernel
(HIP events on the solver's stream, algorithmic bytes from the solver's voxel-visit counters)
and the CPU baseline (Boykov-Kolmogorov on one host core, maxflow() only).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--seed 0] [--no-cpu]
"""
import argparse
import hashlib
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(nx, ny, nz, conn, cs, ct, cn, runs):
    """Boykov-Kolmogorov (the reference's CPU comparator role) on the same instance, one core."""
    import oracle
    ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, conn, cn)
    n = nx * ny * nz
    secs = []
    flow = None
    for i in range(runs + 1):
        f, _, _, t = oracle.bk(n, ea, eb, fwd, bwd, cs.reshape(-1), ct.reshape(-1))
        flow = f
        if i > 0:  # first run is the warm-up
            secs.append(t)
        log(f"cpu baseline run {i}: BK maxflow {t:.3f} s")
    med = statistics.median(secs)
    return dict(value=round(n / med / 1e6, 3), unit="Mvoxels/s", cores=1, kind="port",
                sample=f"full {nx}x{ny}x{nz} {conn}-connected seed-0 instance; Boykov-Kolmogorov maxflow() only "
                       f"(graph construction excluded), median of {runs} runs after 1 warm-up",
                seconds=round(med, 4), flow=flow, host_cpu=cpu_model(), host_nproc=os.cpu_count())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--conn", type=int, default=6)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); the grid "
              "solve is not sharded across GPUs in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cugip_amd

    if cugip_amd.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)

    n = args.size
    nx = ny = nz = n
    vox = nx * ny * nz
    log(f"generating {nx}x{ny}x{nz} {args.conn}-connected instance seed {args.seed}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=args.seed, conn=args.conn)

    g = cugip_amd.GridGraph()
    t0 = time.perf_counter()
    g.set_grid(cs, ct, cn, conn=args.conn)
    h2d_ms = (time.perf_counter() - t0) * 1e3

    for i in range(args.warmup):
        g.max_flow()
        log(f"warmup {i}: {g.stats.total_ms:.2f} ms")
    # cgp_max_flow is synchronous: it returns after the flow value and S_min are final on the device
    t0 = time.perf_counter()
    flow = None
    for _ in range(args.steps):
        flow = g.max_flow()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = vox / (ms_per_step / 1e3) / 1e6
    stats = g.stats.as_dict()
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} Mvoxels/s, flow {flow}")

    # verification against the committed golden fixture (outside the timed region)
    verified = None
    smin = g.source_set()
    try:
        fx = json.load(open(os.path.join(ROOT, "tests", "golden", "grid_fixtures.json")))
        for rec in fx["instances"]:
            if rec["dims"] == [nx, ny, nz] and rec["conn"] == args.conn and rec["seed"] == args.seed \
                    and rec["instance"] == 0:
                digest = hashlib.sha256(np.packbits(smin)).hexdigest()
                verified = bool(rec["flow"] == flow and rec["smin_sha256"] == digest)
    except OSError:
        pass
    log(f"verified against golden: {verified}")

    # profiled solve: per-kernel device time with HIP events on the solver's stream
    g.set_profiling(True)
    g.max_flow()
    prof = g.profile()
    g.set_profiling(False)
    for p in prof:
        log(f"  {p['name']:>20}: {p['launches']:5d} launches {p['device_ms']:8.3f} ms "
            f"{p['algorithmic_bytes'] / 1e9:8.3f} GB alg")
    dom = max(prof, key=lambda p: p["device_ms"])
    achieved = dom["algorithmic_bytes"] / (dom["device_ms"] / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        try:
            tdata = json.load(open(tpath))
            traffic = tdata.get(dom["name"], {}).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic, kernel=dom["name"],
                    kernel_launches=dom["launches"],
                    kernel_avg_us=round(dom["device_ms"] * 1e3 / max(1, dom["launches"]), 2),
                    alg_bytes_per_launch=round(dom["algorithmic_bytes"] / max(1, dom["launches"])),
                    bytes_per_voxel_visit=dom["bytes_per_voxel"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(nx, ny, nz, args.conn, cs, ct, cn, args.cpu_runs)
        if cpu["flow"] != flow:
            log(f"CPU baseline flow {cpu['flow']} differs from GPU flow {flow}")
            cpu["flow_matches_gpu"] = False
        else:
            cpu["flow_matches_gpu"] = True
    g.close()

    out = {
        "metric": "exact min-cut Mvoxels/sec at 256^3 6-conn; learned+batched vs plain GPU speedup",
        "value": round(value, 2),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"plain GPU push-relabel exact min cut (flow value + S_min), {nx}x{ny}x{nz} "
                        f"{args.conn}-connected blob+noise instance, seed {args.seed}",
            "voxels": vox,
            "connectivity": args.conn,
            "tile": [stats["tile_x"], stats["tile_y"], stats["tile_z"]],
            "rounds": stats["rounds"],
            "flow": flow,
            "verified_bit_exact": verified,
            "h2d_setup_ms": round(h2d_ms, 2),
            "learned_batched_speedup": None,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
