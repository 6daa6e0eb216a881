"""The CPU oracle pinned before it is trusted: the reference's own known-answer graphs, brute
force over every cut of small graphs, agreement of three independent max-flow solvers, scipy's
max flow, and the committed golden fixtures. CPU only."""
import hashlib

import numpy as np
import pytest

import cugip_amd
import oracle
from tests.golden import kats

SOLVERS = ("refpr", "dinic", "bk")


def run(name, g):
    n, ea, eb, fwd, bwd, cs, ct = g
    if name == "refpr":
        r = oracle.refpr(n, ea, eb, fwd, bwd, cs, ct)
        return r["flow"], r["smin"], r["tmin"]
    if name == "dinic":
        return oracle.dinic(n, ea, eb, fwd, bwd, cs, ct)
    if name == "bk":
        f, s, t, _ = oracle.bk(n, ea, eb, fwd, bwd, cs, ct)
        return f, s, t
    if name == "brute":
        return oracle.bruteforce(n, ea, eb, fwd, bwd, cs, ct)
    raise ValueError(name)


def cut_capacity(g, source_side):
    """Capacity of the s-t cut whose source side is {s} + source_side."""
    n, ea, eb, fwd, bwd, cs, ct = g
    S = np.asarray(source_side, bool)
    c = int(np.asarray(cs, np.int64)[~S].sum() + np.asarray(ct, np.int64)[S].sum())
    ea, eb = np.asarray(ea), np.asarray(eb)
    c += int(np.asarray(fwd, np.int64)[S[ea] & ~S[eb]].sum() + np.asarray(bwd, np.int64)[S[eb] & ~S[ea]].sum())
    return c


# ---- the reference's known answers -----------------------------------------------------------

@pytest.mark.parametrize("solver", SOLVERS + ("brute",))
def test_kat_grid_4x4(solver):
    g, want = kats.grid_4x4()
    flow, smin, tmin = run(solver, g)
    assert flow == want
    assert smin.tolist() == [1] * 4 + [1] * 4 + [0] * 8       # rows 0 and 1 stay with s
    assert tmin.tolist() == [0] * 8 + [1] * 8                 # rows 2 and 3 reach t


@pytest.mark.parametrize("flip", [False, True])
@pytest.mark.parametrize("solver", SOLVERS)
def test_kat_layered(solver, flip):
    g, want = kats.layered(flip)
    assert run(solver, g)[0] == want


@pytest.mark.parametrize("solver", SOLVERS)
def test_kat_grid_1000(solver):
    g, want = kats.grid_1000()
    flow, smin, tmin = run(solver, g)
    assert flow == want
    # the unique minimum cut runs through the weight-1 row y = 50 -> 51
    assert smin.reshape(1000, 1000)[:51].all() and not smin.reshape(1000, 1000)[51:].any()


def test_kat_bfs_seeds_and_levels():
    g, seeds = kats.two_layer()
    n, ea, eb, fwd, bwd, cs, ct = g
    assert oracle.bfs_seed_count(ct) == seeds
    labels = oracle.initial_labels(n, ea, eb, fwd, bwd, ct)
    # level 1: the 16 vertices with a sink link; their 4 neighbours each in the other layer: level 2
    assert labels.tolist() == [1] * 16 + [2] * 16


# ---- brute force over every cut ----------------------------------------------------------------

def random_graph(rng, n, density, cap_max, asym=True):
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < density]
    if not pairs:
        pairs = [(0, n - 1)]
    ea = np.array([p[0] for p in pairs], np.int32)
    eb = np.array([p[1] for p in pairs], np.int32)
    fwd = rng.integers(0, cap_max + 1, len(pairs)).astype(np.int32)
    bwd = rng.integers(0, cap_max + 1, len(pairs)).astype(np.int32) if asym else fwd.copy()
    cs = (rng.integers(0, cap_max + 1, n) * (rng.random(n) < 0.5)).astype(np.int32)
    ct = (rng.integers(0, cap_max + 1, n) * (rng.random(n) < 0.5)).astype(np.int32)
    return (n, ea, eb, fwd, bwd, cs, ct)


@pytest.mark.parametrize("seed", range(120))
def test_random_graphs_vs_bruteforce(seed):
    rng = np.random.default_rng(seed)
    g = random_graph(rng, int(rng.integers(2, 15)), float(rng.uniform(0.1, 0.7)), int(rng.integers(1, 12)),
                     asym=bool(seed % 2))
    want = run("brute", g)
    for name in SOLVERS:
        got = run(name, g)
        assert got[0] == want[0], name
        assert np.array_equal(got[1], want[1]), (name, "S_min")
        assert np.array_equal(got[2], want[2]), (name, "T_min")
    assert cut_capacity(g, want[1]) == want[0]
    assert cut_capacity(g, 1 - want[2]) == want[0]


@pytest.mark.parametrize("conn", [6, 26])
@pytest.mark.parametrize("seed", range(6))
def test_small_grids_vs_bruteforce(conn, seed):
    rng = np.random.default_rng(100 + seed)
    nx, ny, nz = 3, 3, 2
    n = nx * ny * nz
    cn = rng.integers(0, 6, (conn // 2, nz, ny, nx)).astype(np.int32)
    cnb = rng.integers(0, 6, (conn // 2, nz, ny, nx)).astype(np.int32) if seed % 2 else None
    cs = rng.integers(0, 9, n).astype(np.int32)
    ct = rng.integers(0, 9, n).astype(np.int32)
    ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, conn, cn, cnb)
    g = (n, ea, eb, fwd, bwd, cs, ct)
    want = run("brute", g)
    for name in SOLVERS:
        got = run(name, g)
        assert got[0] == want[0] and np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2]), name


def test_grid_edges_neighbourhoods():
    # every stored edge joins 26-neighbours exactly once; 6-connected edges are axis steps
    for conn, count in ((6, 3), (26, 13)):
        nx, ny, nz = 4, 5, 3
        cn = np.ones((count, nz, ny, nx), np.int32)
        ea, eb, _, _ = oracle.grid_edges(nx, ny, nz, conn, cn)
        pa = np.stack(np.unravel_index(ea, (nz, ny, nx)), 1)
        pb = np.stack(np.unravel_index(eb, (nz, ny, nx)), 1)
        d = np.abs(pa - pb)
        assert d.max() == 1 and (d.sum(1) > 0).all()
        if conn == 6:
            assert (d.sum(1) == 1).all()
        keys = set(zip(np.minimum(ea, eb).tolist(), np.maximum(ea, eb).tolist()))
        assert len(keys) == len(ea)
        full = sum(1 for z in range(nz) for y in range(ny) for x in range(nx)
                   for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                   if (dx, dy, dz) != (0, 0, 0) and (conn == 26 or abs(dx) + abs(dy) + abs(dz) == 1)
                   and 0 <= x + dx < nx and 0 <= y + dy < ny and 0 <= z + dz < nz) // 2
        assert len(ea) == full


# ---- an independent implementation: scipy ------------------------------------------------------

@pytest.mark.parametrize("conn", [6, 26])
def test_oracle_matches_scipy(conn):
    from tests.golden.make_golden import scipy_flow
    nx, ny, nz = 24, 20, 16
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=4, conn=conn)
    flow = oracle.grid_solve("dinic", nx, ny, nz, conn, cs, ct, cn)[0]
    assert scipy_flow(nx, ny, nz, conn, cs, ct, cn) == flow


# ---- golden fixtures ---------------------------------------------------------------------------

def digest(mask):
    return hashlib.sha256(np.packbits(np.asarray(mask, np.uint8).reshape(-1))).hexdigest()


def golden_cases(golden, limit_voxels):
    return [r for r in golden["instances"] if np.prod(r["dims"]) <= limit_voxels]


def test_golden_fixtures_reproduced_by_oracle(golden):
    cases = golden_cases(golden, 262144)
    assert len(cases) >= 18
    for rec in cases:
        nx, ny, nz = rec["dims"]
        vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, rec["seed"], rec["instance"], rec["conn"])
        flow, smin, tmin = oracle.grid_solve("bk", nx, ny, nz, rec["conn"], cs, ct, cn)
        assert flow == rec["flow"]
        assert digest(smin) == rec["smin_sha256"] and digest(tmin) == rec["tmin_sha256"]


def test_golden_masks_file(golden):
    import os
    data = np.load(os.path.join(os.path.dirname(__file__), "golden", "masks_64_seed0.npz"))
    rec = golden["instances"][0]
    assert rec["dims"] == [64, 64, 64] and rec["seed"] == 0
    smin = np.unpackbits(data["smin_64_seed0"])[:64 ** 3]
    assert digest(smin) == rec["smin_sha256"] and int(smin.sum()) == rec["smin_count"]


def test_planted_error_is_detected():
    """A one-unit change on an edge of the minimum cut must change the golden comparison."""
    nx, ny, nz = 32, 24, 16
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=9)
    flow, smin, _ = oracle.grid_solve("dinic", nx, ny, nz, 6, cs, ct, cn)
    s = smin.reshape(nz, ny, nx).astype(bool)
    # find a cut n-link v -> v + (-1,0,0) with v in S and its neighbour outside, capacity > 0
    cand = np.argwhere(s[:, :, 1:] & ~s[:, :, :-1] & (cn[0][:, :, 1:] > 0))
    assert len(cand)
    z, y, x = cand[0]
    cn2 = cn.copy()
    cn2[0, z, y, x + 1] += 1
    flow2 = oracle.grid_solve("dinic", nx, ny, nz, 6, cs, ct, cn2)[0]
    assert flow2 == flow + 1


def test_preflow_closure_equals_source_reachability():
    """S_min from a maximum preflow (closure of the excess vertices) equals BFS from s on a maximum
    flow: the restatement and Dinic compute S_min by these two different routes."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        g = random_graph(rng, int(rng.integers(20, 60)), 0.15, 20)
        r = oracle.refpr(*g)
        f, s, t = oracle.dinic(*g)
        assert r["flow"] == f and np.array_equal(r["smin"], s) and np.array_equal(r["tmin"], t)


@pytest.mark.parametrize("conn,seed", [(6, 0), (6, 1), (26, 2), (26, 3)])
def test_residual_labels_vs_explicit_bfs(conn, seed):
    """oracle.residual_labels (the C relabel on the residual planes' edge list) equals a BFS written
    out voxel by voxel over the neighbour offsets, on random residual planes with zeros (entries
    whose neighbour lies outside the volume carry values that must be ignored)."""
    from collections import deque
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(2, 7, 3))
    r = (rng.integers(0, 5, (conn, nz, ny, nx)) * (rng.random((conn, nz, ny, nx)) < 0.45)).astype(np.int32)
    rt = (rng.integers(0, 3, (nz, ny, nx)) * (rng.random((nz, ny, nx)) < 0.2)).astype(np.int32)
    offs = oracle.half_offsets(conn)
    offs = offs + [(-dx, -dy, -dz) for dx, dy, dz in offs]
    inf = 0x3FFFFFFF
    lab = np.full((nz, ny, nx), inf, np.int64)
    q = deque()
    for z, y, x in zip(*np.nonzero(rt > 0)):
        lab[z, y, x] = 1
        q.append((z, y, x))
    while q:
        z, y, x = q.popleft()
        for (vz, vy, vx) in [(z - dz, y - dy, x - dx) for dx, dy, dz in offs]:
            if not (0 <= vx < nx and 0 <= vy < ny and 0 <= vz < nz) or lab[vz, vy, vx] != inf:
                continue
            # v reaches u = (z, y, x) when the residual v -> u > 0: u = v + o_d for the d with that offset
            d = offs.index((x - vx, y - vy, z - vz))
            if r[d, vz, vy, vx] > 0:
                lab[vz, vy, vx] = lab[z, y, x] + 1
                q.append((vz, vy, vx))
    got = oracle.residual_labels(nx, ny, nz, conn, r, rt)
    assert np.array_equal(got, lab.reshape(-1).astype(np.int32))
    # a planted mistake (the mirror planes read at v instead of v + o_j) is caught
    if conn == 6:
        wrong = r.copy()
        wrong[3:] = np.roll(wrong[3:], 1, axis=3)
        assert not np.array_equal(oracle.residual_labels(nx, ny, nz, conn, wrong, rt), got) or r[3:].sum() == 0
