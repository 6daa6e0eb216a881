"""Training and evaluation of the learned cut proposal (cgp_cnn_*) on synthetic grid instances.

train:     python -m cugip_amd.train_cnn train --instances 2000 --size 128 --out cugip_amd/cnn_weights.npy
evaluate:  python -m cugip_amd.train_cnn evaluate --weights cugip_amd/cnn_weights.npy
infer:     python -m cugip_amd.train_cnn infer --size 256   (the workload of the inference profile)

Training instances are seeds seed_base + i of the blob generator, solved by the plain solver (their
S_min is the label); the capacities are built on the device from each volume. Evaluation uses a
disjoint seed range (64 grids of 192^3 and 8 of 256^3 by default) and compares, per grid, the plain
solve against inference + the certified solve: the flow and S_min must be identical, and the
report gives the certificate gaps, how often the certificate stopped the solve, and the times.
"""
import argparse
import json
import math
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import cugip_amd


def _volumes(size, seeds, threads=8):
    """Generator volumes for `seeds`, made on host threads (the library call releases the GIL)."""
    with ThreadPoolExecutor(threads) as pool:
        for vol in pool.map(lambda s: cugip_amd.generate_volume(size, size, size, s), seeds):
            yield vol


def train(instances, size, epochs, lr, seed_base, init_seed, out, log=print):
    net = cugip_amd.Cnn()
    net.set_params(cugip_amd.Cnn.init_params(init_seed))
    g = cugip_amd.GridGraph()
    total = instances * epochs
    step = 0
    window = []
    t0 = time.time()
    for epoch in range(epochs):
        for vol in _volumes(size, [seed_base + i for i in range(instances)]):
            g.set_grid_volume(vol)
            g.max_flow()
            rate = lr * 0.5 * (1 + math.cos(math.pi * step / total))  # cosine decay
            window.append(net.train_step(g, None, rate))
            step += 1
            if step % 100 == 0 or step == total:
                log(f"step {step}/{total} loss {np.mean(window):.5f} lr {rate:.2e} {time.time() - t0:.1f}s")
                window = []
    params = net.get_params()
    if out:
        np.save(out, params)
    g.close()
    net.close()
    return params


def evaluate(params, sets, seed_base, log=print):
    """sets: list of (count, size). Returns a report dict."""
    net = cugip_amd.Cnn()
    net.set_params(params)
    rows = []
    seed = seed_base
    for count, size in sets:
        for vol in _volumes(size, [seed + i for i in range(count)]):
            with cugip_amd.GridGraph() as g:
                g.set_grid_volume(vol)
                flow = g.max_flow()  # warm-up and reference
                smin = g.source_set()
                t = time.perf_counter()
                g.max_flow()
                plain = time.perf_counter() - t
                t = time.perf_counter()
                cert = net.certificate(g)  # the volume is already on the device
                infer = time.perf_counter() - t
                t = time.perf_counter()
                fc, early = g.max_flow_certified(cert)
                certified = time.perf_counter() - t
                exact = fc == flow and np.array_equal(g.source_set(), smin)
                rows.append(dict(size=size, flow=flow, certificate=cert, gap=cert - flow, early=early, exact=exact,
                                 plain_ms=plain * 1e3, infer_ms=infer * 1e3, certified_ms=certified * 1e3))
        seed += count
    plain = sum(r["plain_ms"] for r in rows)
    learned = sum(r["infer_ms"] + r["certified_ms"] for r in rows)
    rep = dict(grids=len(rows), exact=all(r["exact"] for r in rows), stopped_early=sum(r["early"] for r in rows),
               plain_ms=round(plain, 2), learned_ms=round(learned, 2), speedup=round(plain / learned, 4),
               inference_ms=round(sum(r["infer_ms"] for r in rows), 2),
               median_relative_gap=float(np.median([r["gap"] / r["flow"] for r in rows])),
               sets=[dict(count=c, size=s) for c, s in sets])
    log(json.dumps(rep))
    net.close()
    return rep, rows


def infer(params, size, seed, repeats, log=print):
    """Certificate computations (inference + cut capacity) on one grid built on the device from its
    volume, then one plain solve: the workload of the inference profile."""
    vol = cugip_amd.generate_volume(size, size, size, seed)
    with cugip_amd.GridGraph() as g, cugip_amd.Cnn() as net:
        g.set_grid_volume(vol)
        net.set_params(params)
        net.certificate(g)
        ts = []
        cert = 0
        for _ in range(repeats):
            t = time.perf_counter()
            cert = net.certificate(g)
            ts.append(time.perf_counter() - t)
        flow = g.max_flow()
        log(f"inference + certificate {size}^3: {1e3 * min(ts):.2f} ms; flow {flow}, certificate {cert}, "
            f"gap {cert - flow}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--instances", type=int, default=2000)
    t.add_argument("--size", type=int, default=128)
    t.add_argument("--epochs", type=int, default=1)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed-base", type=int, default=100000)
    t.add_argument("--init-seed", type=int, default=0)
    t.add_argument("--out", default="cugip_amd/cnn_weights.npy")
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default="cugip_amd/cnn_weights.npy")
    e.add_argument("--sets", default="64x192,8x256", help="count x size, comma separated")
    e.add_argument("--seed-base", type=int, default=900000)
    e.add_argument("--out", default=None)
    f = sub.add_parser("infer")
    f.add_argument("--weights", default="cugip_amd/cnn_weights.npy")
    f.add_argument("--size", type=int, default=256)
    f.add_argument("--seed", type=int, default=424242)
    f.add_argument("--repeats", type=int, default=5)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, flush=True)
    if args.cmd == "train":
        train(args.instances, args.size, args.epochs, args.lr, args.seed_base, args.init_seed, args.out, log)
    elif args.cmd == "infer":
        infer(np.load(args.weights), args.size, args.seed, args.repeats, log)
    else:
        sets = [tuple(int(v) for v in s.split("x")) for s in args.sets.split(",")]
        rep, rows = evaluate(np.load(args.weights), sets, args.seed_base, log)
        if args.out:
            with open(args.out, "w") as f:
                json.dump(dict(report=rep, grids=rows), f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
