#!/usr/bin/env python3
"""Benchmark of the exact grid min cut on one MI355X.

Workload (one "step"): the plain GPU push-relabel solve of the 256x256x256 6-connected synthetic
instance (blob volume + Gaussian noise, int32 t-links and n-links, seed 0), from capacities
already resident in HBM to the exact max-flow value and the canonical cut S_min resident in HBM.
Prints one JSON line with the throughput (Mvoxels/s), the roofline of the dominant kernel
(HIP events on the solver's stream, algorithmic bytes from the solver's voxel-visit counters)
and the CPU baseline (Boykov-Kolmogorov on one host core, maxflow() only).

Also reported: config.batched_vs_plain (8 grids of the same size solved as one batch against the
plain solver run on each in turn, bit-exact per grid), config.learned_vs_plain (the learned cut
proposal's inference + certified solve against the plain solve on the same grids) and
config.learned_batched_speedup (inference for every grid + one batched solve, against the plain
solves).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--seed 0] [--no-cpu] [--batch 8]
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
PROFILED_SOLVES = 5


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
    log(f"cpu baseline: edge list of the {nx}x{ny}x{nz} {conn}-connected grid")
    ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, conn, cn)
    log(f"cpu baseline: {len(ea)} edges; Boykov-Kolmogorov runs")
    n = nx * ny * nz
    secs = []
    flow = None
    # runs == 1: a single timed run without warm-up (for instances whose one run takes minutes)
    warm = 1 if runs > 1 else 0
    for i in range(runs + warm):
        f, _, _, t = oracle.bk(n, ea, eb, fwd, bwd, cs.reshape(-1), ct.reshape(-1))
        flow = f
        if i >= warm:
            secs.append(t)
        log(f"cpu baseline run {i}: BK maxflow {t:.3f} s")
    med = statistics.median(secs)
    how = f"median of {runs} runs after 1 warm-up" if warm else "one run"
    return dict(value=round(n / med / 1e6, 3), unit="Mvoxels/s", cores=1, kind="port",
                sample=f"full {nx}x{ny}x{nz} {conn}-connected seed-0 instance; Boykov-Kolmogorov maxflow() only "
                       f"(graph construction excluded), {how}",
                seconds=round(med, 4), flow=flow, host_cpu=cpu_model(), host_nproc=os.cpu_count())


def batched_vs_plain(cugip_amd, n, conn, seed, count, repeats=3):
    """`count` grids of n^3 (instances 0..count-1 of `seed`): one batched solve (cgp_solve_batch) against
    the plain solver run on each grid in turn; both from HBM-resident capacities, per-grid flows and
    S_min compared bit for bit."""
    import numpy as np
    grids = []
    for i in range(count):
        _, cs, ct, cn = cugip_amd.make_instance(n, n, n, seed, i, conn)
        grids.append((cs, ct, cn))
    vox = count * n ** 3
    with cugip_amd.GridBatch() as b:
        b.set_batch(grids, conn)
        b.solve()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            flows = b.solve()
            ts.append(time.perf_counter() - t0)
        batch_s = min(ts)
        smin = [np.packbits(b.batch_source_set(i)) for i in range(count)]
    plain_s = 0.0
    same = True
    for i, (cs, ct, cn) in enumerate(grids):
        with cugip_amd.GridGraph() as g:
            g.set_grid(cs, ct, cn, conn=conn)
            g.max_flow()
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                f = g.max_flow()
                ts.append(time.perf_counter() - t0)
            plain_s += min(ts)
            same &= f == flows[i] and np.array_equal(np.packbits(g.source_set()), smin[i])
    return dict(workload=f"{count} grids of {n}^3 {conn}-connected (seed {seed}, instances 0-{count - 1})",
                batched_mvox_s=round(vox / batch_s / 1e6, 1), plain_mvox_s=round(vox / plain_s / 1e6, 1),
                speedup=round(plain_s / batch_s, 3), batched_ms=round(batch_s * 1e3, 3),
                plain_ms=round(plain_s * 1e3, 3), bit_exact=bool(same))


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_batch_baseline(grids, threads):
    """Boykov-Kolmogorov on every grid of the batch workload, one grid per thread at a time on
    `threads` host threads (the native solver releases the GIL). value = voxels / the longest
    per-thread sum of maxflow() times (graph construction excluded, as for the one-core leg);
    wall_s is the pool's wall time including construction."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    busy = {}
    flows = [None] * len(grids)

    def one(i):
        cs, ct, cn = grids[i]
        nz, ny, nx = cs.shape
        ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, 6, cn)
        f, _, _, t = oracle.bk(nx * ny * nz, ea, eb, fwd, bwd, cs.reshape(-1), ct.reshape(-1))
        flows[i] = f
        tid = threading.get_ident()
        busy[tid] = busy.get(tid, 0.0) + t

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(one, range(len(grids))))
    wall = time.perf_counter() - t0
    vox = sum(g[0].size for g in grids)
    longest = max(busy.values())
    return dict(value=round(vox / longest / 1e6, 2), unit="Mvoxels/s", cores=threads, kind="port",
                sample=f"all {len(grids)} grids of the batch workload, Boykov-Kolmogorov maxflow() only, one grid "
                       f"per thread on {threads} threads",
                seconds=round(longest, 3), wall_s=round(wall, 3), flows=flows)


def batch_workload(cugip_amd, count, seed, repeats=3, cpu=True, plain=True):
    """The batch workload: `count` grids with edge lengths log-uniform in [32, 128] per axis and grid,
    solved as one batch (capacities resident in HBM) against the plain single-grid solver run on
    each grid in turn; every grid's flow and S_min compared bit for bit.
    With cpu, the all-cores Boykov-Kolmogorov baseline over the same grids."""
    import numpy as np
    grids = cugip_amd.make_batch(count, seed)
    vox = sum(g[0].size for g in grids)
    with cugip_amd.GridBatch() as b:
        b.set_batch(grids, 6)
        b.solve()  # warm-up (tile set-up)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            flows = b.solve()
            ts.append(time.perf_counter() - t0)
        stats = b.stats.as_dict()
        batch_s = statistics.median(ts)
        smin = [np.packbits(b.batch_source_set(i)) for i in range(count)]
        # the same batch on the device-driven engine (one captured round graph, replayed)
        b.set_policy(engine=cugip_amd.ENGINE_GRAPH)
        b.solve()  # captures the round graph
        tg = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            graph_flows = b.solve()
            tg.append(time.perf_counter() - t0)
        graph_stats = b.stats.as_dict()
        graph_s = statistics.median(tg)
        graph_same = list(graph_flows) == list(flows) and all(
            np.array_equal(np.packbits(b.batch_source_set(i)), smin[i]) for i in range(count))
    plain_s = 0.0
    same = True
    with cugip_amd.GridGraph() as g:
        for i, (cs, ct, cn) in enumerate(grids if plain else []):
            g.set_grid(cs, ct, cn)
            g.max_flow()
            tg = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                f = g.max_flow()
                tg.append(time.perf_counter() - t0)
            plain_s += statistics.median(tg)
            same &= f == flows[i] and np.array_equal(np.packbits(g.source_set()), smin[i])
    out = dict(workload=f"{count} grids, edge lengths log-uniform in [32, 128] per axis (seed {seed}), "
                        f"6-connected, {vox} voxels",
               batched_ms=round(batch_s * 1e3, 3), batched_mvox_s=round(vox / batch_s / 1e6, 1),
               rounds=stats["rounds"], tile=[stats["tile_x"], stats["tile_y"], stats["tile_z"]],
               graph_engine=dict(ms=round(graph_s * 1e3, 3), mvox_s=round(vox / graph_s / 1e6, 1),
                                 replays=graph_stats["graph_replays"], same_flows_and_smin=bool(graph_same)))
    if plain:
        out.update(plain_ms=round(plain_s * 1e3, 3), plain_mvox_s=round(vox / plain_s / 1e6, 1),
                   speedup=round(plain_s / batch_s, 3), bit_exact=bool(same))
    if cpu:
        c = cpu_batch_baseline(grids, host_cores())
        c["flows_match_gpu"] = c.pop("flows") == flows
        out["cpu_all_cores"] = c
    return out


def random_csr_graph(n, seed):
    """A sparse irregular graph for the generic (CSR) path, shaped like a region adjacency graph:
    the vertices sit on a k x k x k lattice (k = round(n^(1/3)), n = k^3 used), each joined to its
    +x, +y and +z lattice neighbours and to one random vertex within 2 steps per axis, with
    capacities 1..99 each way; a tenth of the vertices get a source t-link and a tenth a sink t-link
    (0..399). Edge records are normalised to first < second (EdgeRecord, edge_record.hpp:10-29),
    fwd being first -> second."""
    import numpy as np
    rng = np.random.default_rng(seed)
    k = max(2, int(round(n ** (1.0 / 3.0))))
    n = k ** 3
    z, y, x = (c.reshape(-1) for c in np.indices((k, k, k)))
    v = np.arange(n, dtype=np.int64)
    ends = []
    for dx, dy, dz in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        ok = (x + dx < k) & (y + dy < k) & (z + dz < k)
        ends.append((v[ok], v[ok] + dx + dy * k + dz * k * k))
    off = rng.integers(-2, 3, (n, 3))
    xx, yy, zz = np.clip(x + off[:, 0], 0, k - 1), np.clip(y + off[:, 1], 0, k - 1), np.clip(z + off[:, 2], 0, k - 1)
    u = (zz * k + yy) * k + xx
    ok = u != v
    ends.append((v[ok], u[ok]))
    a = np.concatenate([e[0] for e in ends])
    b = np.concatenate([e[1] for e in ends])
    fwd = rng.integers(1, 100, a.size).astype(np.int32)
    bwd = rng.integers(1, 100, a.size).astype(np.int32)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    f = np.where(a < b, fwd, bwd)
    bw = np.where(a < b, bwd, fwd)
    cs = (rng.integers(0, 400, n) * (rng.random(n) < 0.1)).astype(np.int32)
    ct = (rng.integers(0, 400, n) * (rng.random(n) < 0.1)).astype(np.int32)
    return n, lo.astype(np.int32), hi.astype(np.int32), f.astype(np.int32), bw.astype(np.int32), cs, ct


def generic_graph(cugip_amd, n, seed, repeats=3, cpu=True):
    """The generic (CSR) path on a random sparse graph: median solve time from HBM-resident arcs
    (set_nweights excluded), checked against the CPU Boykov-Kolmogorov flow."""
    import numpy as np
    n, a, b, fwd, bwd, cs, ct = random_csr_graph(n, seed)
    with cugip_amd.Graph() as g:
        g.set_vertex_count(n)
        g.set_nweights(np.stack([a, b], 1), fwd, bwd)
        g.set_tweights(cs, ct)
        g.max_flow()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            flow = g.max_flow()
            ts.append(time.perf_counter() - t0)
        st = g.stats.as_dict()
    sec = statistics.median(ts)
    out = dict(workload=f"random region-adjacency-like graph, {n} vertices on a lattice, {a.size} edges (lattice "
                        f"neighbours plus one random edge within 2 steps per vertex), seed {seed}",
               ms=round(sec * 1e3, 3), mvertices_s=round(n / sec / 1e6, 2), medges_s=round(a.size / sec / 1e6, 2),
               flow=flow, rounds=st["rounds"], bfs_levels=st["bfs_passes"])
    if cpu:
        import oracle
        f, _, _, t = oracle.bk(n, a, b, fwd, bwd, cs, ct)
        out["cpu_bk"] = dict(seconds=round(t, 3), mvertices_s=round(n / t / 1e6, 3), cores=1, kind="port")
        out["flow_matches_cpu"] = bool(f == flow)
    return out


def learned_vs_plain(cugip_amd, n, seed, count, weights, repeats=3):
    """The learned cut proposal on the batched-comparison grids: per grid, inference + certificate
    (capacity of the predicted cut) + the certified solve, against the plain solve; flows and S_min
    must be identical. exact_certificate_ms: the certified solve given the max flow itself (no
    inference), the bound on what any proposal can save. Also the learned+batched pipeline:
    inference for every grid plus one batched solve (the certificate applies to single grids, so
    the batch solve runs to completion)."""
    import numpy as np
    params = np.load(weights)
    plain = learned = infer_total = ceiling = 0.0
    early = 0
    exact = True
    gaps = []
    with cugip_amd.Cnn() as net:
        net.set_params(params)
        for i in range(count):
            vol = cugip_amd.generate_volume(n, n, n, seed, i)
            with cugip_amd.GridGraph() as g:
                g.set_grid_volume(vol)
                flow = g.max_flow()
                smin = np.packbits(g.source_set())
                ts = []
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    g.max_flow()
                    ts.append(time.perf_counter() - t0)
                plain += min(ts)
                net.certificate(g)  # warm-up
                best = None
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    cert = net.certificate(g)
                    t1 = time.perf_counter()
                    f, stopped = g.max_flow_certified(cert)
                    t2 = time.perf_counter()
                    if best is None or t2 - t0 < best[0]:
                        best = (t2 - t0, t1 - t0)
                learned += best[0]
                infer_total += best[1]
                early += int(stopped)
                gaps.append((cert - flow) / flow)
                exact &= f == flow and np.array_equal(np.packbits(g.source_set()), smin)
                # the most any proposal could save: the max flow itself as the certificate, no inference
                ts = []
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    f, stopped = g.max_flow_certified(flow)
                    ts.append(time.perf_counter() - t0)
                ceiling += min(ts)
                exact &= f == flow and bool(stopped)
    return dict(workload=f"{count} grids of {n}^3 6-connected (seed {seed}), weights {os.path.basename(weights)} "
                         "(trained on 2000 grids of 128^3)",
                speedup=round(plain / learned, 3), plain_ms=round(plain * 1e3, 3), learned_ms=round(learned * 1e3, 3),
                inference_ms=round(infer_total * 1e3, 3), stopped_early=early,
                exact_certificate_ms=round(ceiling * 1e3, 3), exact_certificate_speedup=round(plain / ceiling, 3),
                median_relative_certificate_gap=float(np.median(gaps)), bit_exact=bool(exact))


def calibration_copies(cugip_amd, cs, ct, cn):
    """Device-to-device copies of known size (64, 64 and 192 MiB for 256^3) through cgp_set_grid_device:
    under rocprofv3 --pmc their FETCH_SIZE / WRITE_SIZE give the counters' scale."""
    from tests.hip_rt import DeviceArray
    dev = [DeviceArray(a) for a in (cs, ct, cn)]
    with cugip_amd.GridGraph() as g:
        nz, ny, nx = cs.shape
        g.set_grid_device((nz, ny, nx), 6, dev[0].address, dev[1].address, dev[2].address)
    for d in dev:
        d.free()
    return [a.nbytes for a in (cs, ct, cn)]


def traffic_from_counters(fetch_csv, write_csv, out_path, copy_bytes):
    """HBM bytes per launch of each grid kernel of the last solve, from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE, in KiB per dispatch) of `bench.py --calibration-copies`, scaled by what the
    counters read for the calibration copies (the first three dispatches)."""
    import csv
    import collections

    def per_dispatch(path):
        d = collections.OrderedDict()
        for r in csv.DictReader(open(path)):
            k = int(r["Dispatch_Id"])
            d.setdefault(k, [r["Kernel_Name"], 0.0])
            d[k][1] += float(r["Counter_Value"])
        return d

    fa, wa = per_dispatch(fetch_csv), per_dispatch(write_csv)
    ids = list(fa)
    known_kib = sum(copy_bytes) / 1024.0
    fcal = known_kib / sum(fa[i][1] for i in ids[:3])
    wcal = known_kib / sum(wa[i][1] for i in ids[:3])
    last_init = [i for i in ids if "k_grid_init" in fa[i][0]][-1]
    names = {"k_grid_discharge": "grid_discharge", "k_grid_discharge_wave": "grid_discharge", "k_grid_bfs": "grid_bfs",
             "k_grid_bfs_wave": "grid_bfs", "k_grid_init": "grid_init", "k_grid_source_closure": "grid_source_closure",
             "k_grid_closure_wave": "grid_source_closure"}
    out = dict(workload="last solve of `bench.py --calibration-copies --steps 1 --warmup 1 --no-cpu --batch 0 --generic 0 --batch-workload 0`",
               calibration=dict(fetch_factor=round(fcal, 4), write_factor=round(wcal, 4),
                                how="bytes of the three device-to-device copies / their FETCH_SIZE and WRITE_SIZE"))
    for i in ids:
        if i < last_init:
            continue
        base = fa[i][0].split("(")[0].replace("void ", "").replace("cgp::", "").split("<")[0]
        if base not in names:
            continue
        rec = out.setdefault(names[base], dict(launches=0, fetch_bytes=0.0, write_bytes=0.0))
        rec["launches"] += 1
        rec["fetch_bytes"] += fa[i][1] * fcal * 1024
        rec["write_bytes"] += wa.get(i, ["", 0.0])[1] * wcal * 1024
    for k in names.values():
        if k in out:
            r = out[k]
            r["bytes_per_launch"] = round((r["fetch_bytes"] + r["write_bytes"]) / r["launches"])
            r["fetch_bytes"], r["write_bytes"] = round(r["fetch_bytes"]), round(r["write_bytes"])
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--shape", type=int, nargs=3, metavar=("NX", "NY", "NZ"),
                    help="grid dimensions (default: --size in every axis)")
    ap.add_argument("--volume", action="store_true",
                    help="capacities built on the device from a generated uint8 volume (the image front end), "
                         "as for the 512x512x384 26-connected configuration; no batch legs (the CPU baseline "
                         "computes the same capacities on the host; at 512x512x384 26-connected it needs about "
                         "90 GB of host memory: use --cpu-runs 1)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--conn", type=int, default=6)
    ap.add_argument("--engine", choices=["auto", "tiles", "graph"], default="auto",
                    help="grid engine: host-driven loop (tiles), device-driven graph replay (graph), or the "
                         "library's choice (auto)")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=8, help="grids of the batched-vs-plain comparison (0: skip)")
    ap.add_argument("--generic", type=int, default=1000000,
                    help="vertices of the generic-graph (CSR) leg, about 4 edges each (0: skip)")
    ap.add_argument("--batch-workload", type=int, default=512,
                    help="grids of the batch workload leg, edge lengths log-uniform in [32, 128] (0: skip)")
    ap.add_argument("--batch-plain", type=int, default=1,
                    help="0: skip the batch leg's one-by-one plain solves (each grid captures its own round graph)")
    ap.add_argument("--batch-workload-only", action="store_true",
                    help="run only the batch workload leg and print its record (for kernel traces of that leg)")
    ap.add_argument("--calibration-copies", action="store_true",
                    help="start with device-to-device copies of known size (for rocprofv3 --pmc)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="only turn two --pmc counter_collection.csv files into profiles/hbm_traffic.json")
    args = ap.parse_args()

    if args.traffic_from:
        n = args.size
        copy = [4 * n ** 3, 4 * n ** 3, 12 * n ** 3]
        res = traffic_from_counters(args.traffic_from[0], args.traffic_from[1],
                                    os.path.join(ROOT, "profiles", "hbm_traffic.json"), copy)
        print(json.dumps(res))
        return

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

    if args.batch_workload_only:
        rec = batch_workload(cugip_amd, args.batch_workload, args.seed, cpu=not args.no_cpu,
                             plain=bool(args.batch_plain))
        print(json.dumps(rec))
        return

    n = args.size
    nx, ny, nz = args.shape if args.shape else (n, n, n)
    vox = nx * ny * nz
    log(f"generating {nx}x{ny}x{nz} {args.conn}-connected instance seed {args.seed}")
    if args.volume:
        vol = cugip_amd.generate_volume(nx, ny, nz, seed=args.seed)
        cs = ct = cn = None
    else:
        vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=args.seed, conn=args.conn)
    if args.calibration_copies and cs is not None:
        log(f"calibration copies: {calibration_copies(cugip_amd, cs, ct, cn)} bytes")

    engine = {"auto": cugip_amd.ENGINE_AUTO, "tiles": cugip_amd.ENGINE_TILES,
              "graph": cugip_amd.ENGINE_GRAPH}[args.engine]
    log("instance ready; capacities to the device")
    g = cugip_amd.GridGraph()
    t0 = time.perf_counter()
    if args.volume:
        g.set_grid_volume(vol, conn=args.conn)
    else:
        g.set_grid(cs, ct, cn, conn=args.conn)
    h2d_ms = (time.perf_counter() - t0) * 1e3
    g.set_policy(engine=engine)
    log(f"grid set ({h2d_ms:.1f} ms); warm-up solves")

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

    # profiled solves: per-kernel device time from a HIP event pair around every launch on the
    # solver's stream, read after each solve (no host wait between launches; the host-driven loop,
    # which runs the same kernels); summed over several solves
    g.set_policy()
    g.set_profiling(True)
    acc = {}
    for _ in range(PROFILED_SOLVES):
        g.max_flow()
        for p in g.profile():
            a = acc.setdefault(p["name"], dict(name=p["name"], launches=0, device_ms=0.0, algorithmic_bytes=0.0,
                                               bytes_per_voxel=p["bytes_per_voxel"]))
            a["launches"] += p["launches"]
            a["device_ms"] += p["device_ms"]
            a["algorithmic_bytes"] += p["algorithmic_bytes"]
    g.set_profiling(False)
    prof = list(acc.values())
    for p in prof:
        log(f"  {p['name']:>20}: {p['launches']:5d} launches {p['device_ms']:8.3f} ms "
            f"{p['algorithmic_bytes'] / 1e9:8.3f} GB alg ({PROFILED_SOLVES} profiled solves)")
    dom = max(prof, key=lambda p: p["device_ms"])
    achieved = dom["algorithmic_bytes"] / (dom["device_ms"] / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    # the counter runs behind that file measured the default workload only
    default_workload = (not args.shape and not args.volume and n == ap.get_default("size")
                        and args.conn == ap.get_default("conn") and args.seed == ap.get_default("seed"))
    if default_workload and os.path.exists(tpath):
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

    batch = None
    learned = None
    learned_batched = None
    if args.batch > 0 and not args.volume and nx == ny == nz == n:
        batch = batched_vs_plain(cugip_amd, n, args.conn, args.seed, args.batch)
        log(f"batched vs plain: {batch}")
        weights = os.path.join(ROOT, "cugip_amd", "cnn_weights.npy")
        if args.conn == 6 and os.path.exists(weights):
            learned = learned_vs_plain(cugip_amd, n, args.seed, args.batch, weights)
            log(f"learned vs plain: {learned}")
            # inference for every grid plus one batched solve, against the plain solves
            learned_batched = round(batch["plain_ms"] / (batch["batched_ms"] + learned["inference_ms"]), 3)

    batch512 = None
    if args.batch_workload > 0 and not args.volume and args.conn == 6:
        log(f"batch workload: {args.batch_workload} grids")
        batch512 = batch_workload(cugip_amd, args.batch_workload, args.seed, cpu=not args.no_cpu,
                                  plain=args.batch_plain != 0)
        log(f"batch workload: {batch512}")

    generic = None
    if args.generic > 0 and not args.volume:
        log(f"generic graph: {args.generic} vertices")
        generic = generic_graph(cugip_amd, args.generic, args.seed, cpu=not args.no_cpu)
        log(f"generic graph: {generic}")

    cpu = None
    if not args.no_cpu:
        if args.volume:
            # the CPU solver needs the capacities on the host: the same tables, computed here
            cs, ct, cn = cugip_amd.caps_from_volume(vol, args.conn)
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
            "batched_vs_plain": batch,
            "learned_vs_plain": learned,
            "learned_batched_speedup": learned_batched,
            "batch512": batch512,
            "generic_graph": generic,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
