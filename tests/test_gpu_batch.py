"""Batched grid path on the GPU (cgp_set_grid_batch / cgp_solve_batch): every grid of a batch of
uneven grids gets the flow, S_min and T_min of the CPU oracle, and of the single-grid solver."""
import os
import subprocess

import numpy as np
import pytest

import cugip_amd
import oracle
from tests.gpu_util import grid_cut_capacity, say

pytestmark = pytest.mark.gpu


def uneven_instances(count, seed, conn, lo=6, hi=48):
    shapes = cugip_amd.batch_shapes(count, seed, lo, hi)
    out = []
    for i, (nz, ny, nx) in enumerate(shapes):
        vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed, i, conn)
        out.append((cs, ct, cn))
    return out


def solve_batch(grids, conn, **policy):
    with cugip_amd.GridBatch() as b:
        b.set_batch(grids, conn)
        if policy:
            b.set_policy(**policy)
        flows = b.solve()
        smin = [b.batch_source_set(i).reshape(-1) for i in range(len(grids))]
        tmin = [b.batch_fill_segments(i, 1, 0).reshape(-1) for i in range(len(grids))]
        return flows, smin, tmin, b.stats.as_dict()


@pytest.mark.parametrize("conn,seed,count,policy", [
    (6, 0, 12, {}),
    (6, 1, 9, dict(tile_x=64, tile_y=4, tile_z=4)),
    (6, 2, 7, dict(discharges_per_relabel=1, max_tile_iterations=2)),
    (26, 3, 6, {}),
    (26, 4, 5, dict(tile_x=16, tile_y=8, tile_z=4)),
])
def test_batch_parity_oracle(conn, seed, count, policy):
    say(f"batch parity conn {conn} seed {seed} count {count} policy {policy}")
    grids = uneven_instances(count, seed, conn)
    flows, smin, tmin, stats = solve_batch(grids, conn, **policy)
    assert stats["flow"] == sum(flows)
    for i, (cs, ct, cn) in enumerate(grids):
        nz, ny, nx = cs.shape
        ref_flow, ref_smin, ref_tmin = oracle.grid_solve("bk", nx, ny, nz, conn, cs, ct, cn)
        assert flows[i] == ref_flow, i
        assert np.array_equal(smin[i], ref_smin), (i, int((smin[i] != ref_smin).sum()))
        assert np.array_equal(tmin[i], ref_tmin), (i, int((tmin[i] != ref_tmin).sum()))


def test_batch_mixed_back_capacities():
    """Random capacities with zeros; some grids carry independent back capacities, some are symmetric."""
    say("batch mixed back capacities")
    rng = np.random.default_rng(5)
    grids = []
    for i in range(6):
        nx, ny, nz = (int(v) for v in rng.integers(2, 30, 3))
        cs = (rng.integers(0, 200, (nz, ny, nx)) * (rng.random((nz, ny, nx)) < 0.3)).astype(np.int32)
        ct = (rng.integers(0, 200, (nz, ny, nx)) * (rng.random((nz, ny, nx)) < 0.3)).astype(np.int32)
        cn = (rng.integers(0, 60, (3, nz, ny, nx)) * (rng.random((3, nz, ny, nx)) < 0.8)).astype(np.int32)
        cnb = (rng.integers(0, 60, (3, nz, ny, nx)) * (rng.random((3, nz, ny, nx)) < 0.8)).astype(np.int32)
        grids.append((cs, ct, cn, cnb if i % 2 == 0 else None))
    flows, smin, tmin, _ = solve_batch(grids, 6)
    for i, (cs, ct, cn, cnb) in enumerate(grids):
        nz, ny, nx = cs.shape
        ref_flow, ref_smin, ref_tmin = oracle.grid_solve("dinic", nx, ny, nz, 6, cs, ct, cn, cnb)
        assert flows[i] == ref_flow, i
        assert np.array_equal(smin[i], ref_smin), i
        assert np.array_equal(tmin[i], ref_tmin), i


def test_batch_matches_single_solves_on_batch_workload():
    """A slice of the batch workload (edge lengths log-uniform in [32, 128] per axis and grid): every
    grid's flow and S_min equal the single-grid solver's, and the S_min cut has the flow's capacity."""
    say("batch workload slice, 24 grids")
    shapes = cugip_amd.batch_shapes(24, seed=0)
    grids = []
    for i, (nz, ny, nx) in enumerate(shapes):
        vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, 0, i, 6)
        grids.append((cs, ct, cn))
    flows, smin, _, stats = solve_batch(grids, 6)
    for i, (cs, ct, cn) in enumerate(grids):
        with cugip_amd.GridGraph() as g:
            g.set_grid(cs, ct, cn)
            assert g.max_flow() == flows[i], i
            assert np.array_equal(g.source_set().reshape(-1), smin[i]), i
        assert grid_cut_capacity(smin[i], cs, ct, cn) == flows[i], i


def test_batch_workload_512():
    """The full batch workload: 512 grids with edge lengths log-uniform in [32, 128] per axis and
    grid (165.6 M voxels), solved as one batch by the host-driven and by the device-driven engine
    (the same flows). Every grid's flow and S_min equal the single-grid solver's bit for bit; the 16
    smallest grids also match the CPU oracle (Boykov-Kolmogorov) in flow, S_min and T_min."""
    import time
    say("batch workload, 512 grids: building")
    grids = cugip_amd.make_batch(512, seed=0)
    vox = sum(cs.size for cs, _, _ in grids)
    with cugip_amd.GridBatch() as b:
        b.set_batch(grids, 6)
        b.set_policy(engine=cugip_amd.ENGINE_GRAPH)
        t0 = time.perf_counter()
        graph_flows = b.solve()
        stats = b.stats.as_dict()
        say(f"  graph batch solve {1e3 * (time.perf_counter() - t0):.1f} ms for {vox} voxels, "
            f"{stats['graph_replays']} replays, {stats['rounds']} rounds")
        assert stats["graph_replays"] > 0
        b.set_policy()
        t0 = time.perf_counter()
        flows = b.solve()
        say(f"  tiles batch solve {1e3 * (time.perf_counter() - t0):.1f} ms")
        assert b.stats.graph_replays == 0
        assert list(flows) == list(graph_flows)
        smin = [b.batch_source_set(i).reshape(-1) for i in range(len(grids))]
        order = sorted(range(len(grids)), key=lambda i: grids[i][0].size)
        small = order[:16]
        tmin = {i: b.batch_fill_segments(i, 1, 0).reshape(-1) for i in small}
    say("  single-grid solves")
    with cugip_amd.GridGraph() as g:
        for i, (cs, ct, cn) in enumerate(grids):
            g.set_grid(cs, ct, cn)
            assert g.max_flow() == flows[i], i
            assert np.array_equal(g.source_set().reshape(-1), smin[i]), i
    say("  oracle on the 16 smallest grids")
    for i in small:
        cs, ct, cn = grids[i]
        nz, ny, nx = cs.shape
        ref_flow, ref_smin, ref_tmin = oracle.grid_solve("bk", nx, ny, nz, 6, cs, ct, cn)
        assert flows[i] == ref_flow, i
        assert np.array_equal(smin[i], ref_smin), i
        assert np.array_equal(tmin[i], ref_tmin), i


def test_batch_of_one_and_errors():
    say("batch of one, errors")
    vol, cs, ct, cn = cugip_amd.make_instance(33, 20, 17, 9, 0, 6)
    flows, smin, tmin, _ = solve_batch([(cs, ct, cn)], 6)
    ref_flow, ref_smin, _ = oracle.grid_solve("bk", 33, 20, 17, 6, cs, ct, cn)
    assert flows == [ref_flow] and np.array_equal(smin[0], ref_smin)
    grids = uneven_instances(3, 1, 6)
    with cugip_amd.GridBatch() as b:
        b.set_batch(grids, 6)
        b.solve()
        with pytest.raises(cugip_amd.CgpError) as e:
            b.batch_source_set(3)
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        buf = np.zeros(1, np.uint8)  # the library checks the index before it writes
        lib = cugip_amd.library()
        for bad in (3, -1):
            assert lib.cgp_batch_source_set(b._h, bad, buf.ctypes.data_as(cugip_amd._u8p)) == \
                cugip_amd.CGP_ERR_INVALID
        with pytest.raises(cugip_amd.CgpError) as e:
            b.source_set()  # a batch must name its grid
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        b.set_policy(engine=7)  # no such engine
        with pytest.raises(cugip_amd.CgpError) as e:
            b.solve()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        b.set_policy()
        assert len(b.solve()) == 3
    with cugip_amd.GridBatch() as b:
        with pytest.raises(cugip_amd.CgpError) as e:
            b.set_batch([], 6)
        assert e.value.code == cugip_amd.CGP_ERR_INVALID


def test_cpp_dropin_batch():
    say("cpp batch program")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "graph_cut_batch")
    out = subprocess.run(["timeout", "-k", "5", "120", exe, "9", "6", "5"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.strip().splitlines()
    assert len(lines) == 9 and all(line.endswith(" ok") for line in lines), out.stdout
