"""Autotuner of the grid schedule: tile shape, discharge phases per global relabel and in-tile
iterations per discharge.

For each workload class it times every candidate policy on sample grids (median of a few solves
after a warm-up), requires every candidate to return the same per-grid flows (a schedule may
change the work done, never the result), and keeps the fastest. `--write-header` turns the
winners into cugip_amd/csrc/tuned_policy.h, which the library compiles in as its defaults.

    python -m cugip_amd.autotune --out tuned.json --write-header
"""
import argparse
import itertools
import json
import os
import statistics
import sys
import time

import cugip_amd

SHAPES = {6: [(32, 8, 4), (32, 8, 8), (64, 4, 4), (32, 4, 4), (64, 4, 8), (16, 8, 4), (16, 4, 4), (32, 2, 4)],
          26: [(32, 4, 4), (16, 8, 4), (16, 4, 4), (32, 4, 2)]}
PHASES = {6: [2, 4, 6, 8, 12, 16], 26: [8, 16, 24]}
ITERATIONS = [1, 2, 4, 8, 16, 32]

# name: (conn, batch, max_voxels of the table row, builder of the sample grids). Single-grid
# workloads are timed as the sum over their sample grids (one solver each), so that the winner is
# not fitted to one instance.
WORKLOADS = {
    "grid6_small": (6, False, 1 << 21, lambda: [cugip_amd.make_instance(96, 96, 96, s)[1:] for s in (11, 12, 13)]),
    "grid6_large": (6, False, None, lambda: [cugip_amd.make_instance(256, 256, 256, s)[1:] for s in (1, 2, 3)]),
    "batch6": (6, True, None, lambda: [cugip_amd.make_instance(nx, ny, nz, 1, i, 6)[1:]
                                       for i, (nz, ny, nx) in enumerate(cugip_amd.batch_shapes(128, seed=1))]),
    "grid26_small": (26, False, 1 << 22, lambda: [cugip_amd.make_instance(160, 160, 128, s, 0, 26)[1:]
                                                  for s in (3, 4)]),
    "grid26": (26, False, None, lambda: [cugip_amd.make_instance(256, 256, 192, s, 0, 26)[1:] for s in (3, 4)]),
    "batch26": (26, True, None, lambda: [cugip_amd.make_instance(nx, ny, nz, 2, i, 26)[1:]
                                         for i, (nz, ny, nx) in enumerate(cugip_amd.batch_shapes(24, 2, 24, 64))]),
}


class _Graphs:
    """One GridGraph per sample grid, with the policy calls of a single handle."""

    def __init__(self, grids, conn):
        self.graphs = []
        for cs, ct, cn in grids:
            g = cugip_amd.GridGraph()
            g.set_grid(cs, ct, cn, conn=conn)
            self.graphs.append(g)

    def set_policy(self, **kw):
        for g in self.graphs:
            g.set_policy(**kw)

    def solve(self):
        return [g.max_flow() for g in self.graphs]

    def close(self):
        for g in self.graphs:
            g.close()


def _solver(grids, conn, batch):
    if batch:
        h = cugip_amd.GridBatch()
        h.set_batch(grids, conn)
        return h, lambda: h.solve()
    h = _Graphs(grids, conn)
    return h, h.solve


def tune_workload(name, repeats=3, log=print, shapes=None):
    """Times every (tile shape, phases, iterations) candidate on the workload's sample grids;
    shapes (a list of (tx, ty, tz)) restricts the tile shapes to those of the connectivity's list."""
    conn, batch, max_voxels, build = WORKLOADS[name]
    candidates = [s for s in SHAPES[conn] if shapes is None or tuple(s) in shapes]
    grids = build()
    voxels = sum(int(g[0].size) for g in grids)
    handle, solve = _solver(grids, conn, batch)
    results = []
    reference = None
    try:
        for shape, phases, iters in itertools.product(candidates, PHASES[conn], ITERATIONS):
            handle.set_policy(tile_x=shape[0], tile_y=shape[1], tile_z=shape[2], discharges_per_relabel=phases,
                              max_tile_iterations=iters)
            flows = solve()
            if reference is None:
                reference = flows
            elif flows != reference:
                raise RuntimeError("%s: policy %s %d %d changed the flows" % (name, shape, phases, iters))
            times = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                solve()
                times.append((time.perf_counter() - t0) * 1e3)
            ms = statistics.median(times)
            results.append(dict(tile=list(shape), phases=phases, iterations=iters, ms=ms))
            log("%s tile %dx%dx%d phases %d iters %d: %.2f ms" % (name, shape[0], shape[1], shape[2], phases,
                                                                   iters, ms))
    finally:
        handle.close()
    best = min(results, key=lambda r: r["ms"])
    return dict(workload=name, conn=conn, batch=batch, max_voxels=max_voxels, grids=len(grids), voxels=voxels,
                best=best, candidates=results)


def _row_class(line):
    """(conn, batch, bound) of a table row of the header."""
    f = [x.strip() for x in line.strip().lstrip("{").split("}")[0].split(",")]
    return int(f[0]), int(f[1]), f[2]


def header_text(tuned):
    """The header with one row per tuned workload; rows of the classes (connectivity, batch flag,
    size bound) not tuned in this run are kept as they are."""
    src = open(os.path.join(os.path.dirname(__file__), "csrc", "tuned_policy.h")).read()
    start = src.index("static const TunedPolicy kTunedPolicies[] = {")
    end = src.index("};", start)
    classes = {(e["conn"], 1 if e["batch"] else 0, "INT64_MAX" if e["max_voxels"] is None else "%dLL" % e["max_voxels"])
               for e in tuned}
    kept = [line for line in src[start:end].splitlines()
            if line.strip().startswith("{") and _row_class(line) not in classes]
    rows = []
    # bounded rows first, so that they match before the unbounded row of the same class
    for entry in sorted(tuned, key=lambda e: (e["conn"], e["batch"], e["max_voxels"] is None,
                                              e["max_voxels"] or 0)):
        b = entry["best"]
        mv = "INT64_MAX" if entry["max_voxels"] is None else "%dLL" % entry["max_voxels"]
        rows.append("\t{%d, %d, %s, %d, %d, %d, %d, %d},  // %s: %.2f ms" % (
            entry["conn"], 1 if entry["batch"] else 0, mv, b["tile"][0], b["tile"][1], b["tile"][2], b["phases"],
            b["iterations"], entry["workload"], b["ms"]))
    # keep the table ordered by connectivity, batch flag and bound, whichever run wrote a row
    def key(line):
        conn, batch, bound = _row_class(line)
        return conn, batch, bound == "INT64_MAX", 0 if bound == "INT64_MAX" else int(bound.rstrip("L"))
    rows = sorted(rows + kept, key=key)
    head = src[:start] + "static const TunedPolicy kTunedPolicies[] = {\n" \
        "\t// conn batch max_voxels  tile        phases iters\n"
    return head + "\n".join(rows) + "\n" + src[end:]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workloads", default=",".join(WORKLOADS))
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default=None, help="write all timings as JSON")
    ap.add_argument("--write-header", action="store_true", help="rewrite csrc/tuned_policy.h")
    ap.add_argument("--from-json", default=None, help="take the timings of an earlier --out file")
    ap.add_argument("--shapes", default=None, help="only these tile shapes, e.g. 16x4x4,32x2x4")
    args = ap.parse_args(argv)
    shapes = None
    if args.shapes:
        shapes = {tuple(int(v) for v in s.split("x")) for s in args.shapes.split(",")}
    tuned = []
    if args.from_json:
        with open(args.from_json) as f:
            tuned = json.load(f)
    for name in ([] if args.from_json else args.workloads.split(",")):
        tuned.append(tune_workload(name, args.repeats, log=lambda *a: print(*a, flush=True), shapes=shapes))
        b = tuned[-1]["best"]
        print("best %s: %s" % (name, json.dumps(b)), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(tuned, f, indent=1)
    if args.write_header:
        text = header_text(tuned)
        with open(os.path.join(os.path.dirname(__file__), "csrc", "tuned_policy.h"), "w") as f:
            f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
