#!/usr/bin/env python3
"""Writes the golden fixtures of the grid graph cut (tests/golden/grid_fixtures.json and
tests/golden/masks_64_seed0.npz).

For every synthetic instance (the integer generator of cugip_amd: blob volume, Boykov-Jolly
t-links, contrast-sensitive n-links) it records the max-flow value and SHA-256 digests of the
packed S_min and T_min masks. A value is written only when the independent CPU solvers of the
oracle agree on it (Boykov-Kolmogorov and Dinic everywhere; the restatement of the reference's
relabel/push loop where it finishes in seconds), and, where the total fits int32, when
scipy.sparse.csgraph.maximum_flow (a third-party implementation) returns the same flow value.

Run from the repository root: python tests/golden/make_golden.py (everything), or
python tests/golden/make_golden.py --large-26c (only the 256x256x192 26-connected record).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import cugip_amd  # noqa: E402  (host-side generator only)
import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def digest(mask):
    return hashlib.sha256(np.packbits(np.asarray(mask, np.uint8).reshape(-1))).hexdigest()


def scipy_flow(nx, ny, nz, conn, cs, ct, cn):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import maximum_flow

    ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, conn, cn)
    n = nx * ny * nz
    s, t = n, n + 1
    v = np.arange(n, dtype=np.int64)
    rows = np.concatenate([ea, eb, np.full(n, s), v]).astype(np.int64)
    cols = np.concatenate([eb, ea, v, np.full(n, t)]).astype(np.int64)
    data = np.concatenate([fwd, bwd, cs.reshape(-1), ct.reshape(-1)]).astype(np.int64)
    keep = data > 0
    mat = csr_matrix((data[keep].astype(np.int32), (rows[keep], cols[keep])), shape=(n + 2, n + 2))
    mat.sum_duplicates()
    return int(maximum_flow(mat, s, t, method="dinic").flow_value)


def solve_instance(nx, ny, nz, conn, seed, instance, with_refpr=True):
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=seed, instance=instance, conn=conn)
    n = nx * ny * nz
    ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, conn, cn)
    t0 = time.time()
    f_bk, s_bk, t_bk, bk_secs = oracle.bk(n, ea, eb, fwd, bwd, cs.reshape(-1), ct.reshape(-1))
    f_di, s_di, t_di = oracle.dinic(n, ea, eb, fwd, bwd, cs.reshape(-1), ct.reshape(-1))
    checked = ["bk", "dinic"]
    assert f_bk == f_di and np.array_equal(s_bk, s_di) and np.array_equal(t_bk, t_di), (nx, ny, nz, conn, seed)
    if with_refpr:
        r = oracle.refpr(n, ea, eb, fwd, bwd, cs.reshape(-1), ct.reshape(-1))
        assert r["flow"] == f_bk and np.array_equal(r["smin"], s_bk) and np.array_equal(r["tmin"], t_bk)
        checked.append("refpr")
    if f_bk < 2 ** 31 - 1 and n <= 300000:
        assert scipy_flow(nx, ny, nz, conn, cs, ct, cn) == f_bk
        checked.append("scipy")
    rec = dict(dims=[nx, ny, nz], conn=conn, seed=seed, instance=instance, flow=int(f_bk),
               smin_count=int(s_bk.sum()), tmin_count=int(t_bk.sum()), smin_sha256=digest(s_bk),
               tmin_sha256=digest(t_bk), volume_sha256=hashlib.sha256(vol.tobytes()).hexdigest(),
               checked_by=checked, bk_maxflow_seconds=round(bk_secs, 4))
    print(f"{nx}x{ny}x{nz} c{conn} seed {seed} inst {instance}: flow {f_bk} |S| {rec['smin_count']} "
          f"checked {checked} ({time.time() - t0:.1f}s)", flush=True)
    return rec, s_bk, t_bk


LARGE_26C = ((256, 256, 192), 26, 0, 0)


def append(dims, conn, seed, inst):
    """Solves one more instance (BK and Dinic) and adds its record to grid_fixtures.json."""
    path = os.path.join(HERE, "grid_fixtures.json")
    data = json.load(open(path))
    data["instances"] = [r for r in data["instances"]
                         if not (r["dims"] == list(dims) and r["conn"] == conn and r["seed"] == seed
                                 and r["instance"] == inst)]
    rec, _, _ = solve_instance(*dims, conn, seed, inst, with_refpr=False)
    data["instances"].append(rec)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--large-26c":
        # the largest 26-connected instance the two CPU solvers finish here (about 12.6 M voxels)
        append(*LARGE_26C)
        return
    records = []
    masks = {}
    for seed in range(16):
        rec, s, t = solve_instance(64, 64, 64, 6, seed, 0)
        records.append(rec)
        if seed == 0:
            masks["smin_64_seed0"] = np.packbits(s)
            masks["tmin_64_seed0"] = np.packbits(t)
    for dims, conn, seed, inst in [((48, 40, 32), 26, 0, 0), ((48, 40, 32), 26, 1, 0), ((37, 101, 64), 6, 0, 3),
                                   ((128, 32, 90), 6, 5, 1), ((96, 80, 64), 26, 0, 0), ((128, 128, 128), 6, 1000, 0)]:
        rec, _, _ = solve_instance(*dims, conn, seed, inst, with_refpr=np.prod(dims) <= 600000)
        records.append(rec)
    rec, _, _ = solve_instance(256, 256, 256, 6, 0, 0, with_refpr=False)
    records.append(rec)
    rec, _, _ = solve_instance(*LARGE_26C[0], *LARGE_26C[1:], with_refpr=False)
    records.append(rec)
    out = dict(
        description="Synthetic grid graph-cut instances: exact max flow and SHA-256 of np.packbits of the "
                    "S_min / T_min byte masks (voxel order z, y, x with x fastest).",
        instances=records,
    )
    with open(os.path.join(HERE, "grid_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "masks_64_seed0.npz"), **masks)


if __name__ == "__main__":
    main()
