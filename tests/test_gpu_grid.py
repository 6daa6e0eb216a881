"""Grid path on the GPU, through the C ABI: bit-exact parity with the CPU oracle (max-flow value,
S_min and T_min), the committed golden fixtures, full-size properties, edge cases and errors."""
import numpy as np
import pytest

import cugip_amd
import oracle
from tests.gpu_util import digest, grid_cut_capacity, say

pytestmark = pytest.mark.gpu


def gpu_solve(cs, ct, cn, cnb=None, conn=6, **policy):
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn, cnb, conn=conn)
        if policy:
            g.set_policy(**policy)
        flow = g.max_flow()
        return flow, g.source_set(), g.fill_segments(1, 0), g.stats.as_dict()


def check_against_oracle(cs, ct, cn, cnb=None, conn=6, method="bk", **policy):
    nz, ny, nx = cs.shape
    flow, smin, tmin, stats = gpu_solve(cs, ct, cn, cnb, conn, **policy)
    ref_flow, ref_smin, ref_tmin = oracle.grid_solve(method, nx, ny, nz, conn, cs, ct, cn, cnb)
    assert flow == ref_flow
    assert np.array_equal(smin, ref_smin), int((smin != ref_smin).sum())
    assert np.array_equal(tmin, ref_tmin), int((tmin != ref_tmin).sum())
    return stats


CASES = [
    # dims, conn, seed, instance, policy
    ((37, 29, 23), 6, 3, 0, {}),
    ((64, 64, 64), 6, 0, 0, {}),
    ((64, 64, 64), 6, 1, 0, dict(tile_x=32, tile_y=8, tile_z=8)),
    ((64, 64, 64), 6, 2, 0, dict(tile_x=64, tile_y=4, tile_z=4)),
    ((64, 64, 64), 6, 3, 0, dict(tile_x=32, tile_y=4, tile_z=4)),
    ((64, 64, 64), 6, 4, 0, dict(tile_x=64, tile_y=4, tile_z=8)),
    ((100, 7, 33), 6, 5, 2, dict(discharges_per_relabel=1, max_tile_iterations=1)),
    ((45, 61, 12), 6, 6, 1, dict(discharges_per_relabel=7, max_tile_iterations=64)),
    ((70, 52, 31), 6, 7, 0, dict(minimal_residual=3.5)),
    ((24, 20, 16), 26, 1, 0, {}),
    ((40, 33, 17), 26, 2, 0, dict(tile_x=16, tile_y=8, tile_z=4)),
    ((35, 18, 21), 26, 3, 4, dict(tile_x=16, tile_y=4, tile_z=4)),
    ((41, 23, 19), 26, 5, 1, dict(tile_x=32, tile_y=4, tile_z=2)),
    ((48, 40, 32), 26, 0, 0, dict(discharges_per_relabel=3)),
]


@pytest.mark.parametrize("dims,conn,seed,instance,policy", CASES)
def test_parity_synthetic(dims, conn, seed, instance, policy):
    nx, ny, nz = dims
    say(f"grid parity {dims} conn {conn} seed {seed} instance {instance} policy {policy}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed, instance, conn)
    check_against_oracle(cs, ct, cn, None, conn, **policy)


@pytest.mark.parametrize("engine", [cugip_amd.ENGINE_GRAPH, cugip_amd.ENGINE_TILES])
@pytest.mark.parametrize("dims,seed,policy", [
    ((64, 64, 64), 0, {}),
    ((96, 40, 28), 1, dict(max_tile_iterations=1)),
    ((33, 17, 9), 2, dict(discharges_per_relabel=1)),
    ((128, 64, 48), 3, dict(max_tile_iterations=64, discharges_per_relabel=6)),
    ((61, 45, 37), 4, dict(tile_x=32, tile_y=8, tile_z=8)),
    ((20, 300, 3), 5, dict(tile_x=64, tile_y=4, tile_z=4)),
])
def test_parity_engines(engine, dims, seed, policy):
    """The host-driven loop and the device-driven graph replay give the oracle's flow, S_min and
    T_min (T_min after a graph solve comes from one exact relabel on request)."""
    nx, ny, nz = dims
    say(f"grid engine {engine} {dims} seed {seed} policy {policy}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed, 0, 6)
    stats = check_against_oracle(cs, ct, cn, None, 6, engine=engine, **policy)
    assert (stats["graph_replays"] > 0) == (engine == cugip_amd.ENGINE_GRAPH)


def test_graph_engine_refuses_host_decisions():
    """The graph engine takes no per-stage host decision: minimal_residual and an unknown engine
    are refused, and so are 26-connected grids; AUTO runs the host loop."""
    say("graph engine argument checks")
    vol, cs, ct, cn = cugip_amd.make_instance(16, 16, 16, 1, 0, 6)
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn)
        g.set_policy(engine=cugip_amd.ENGINE_GRAPH, minimal_residual=2.0)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        g.set_policy(engine=7)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        g.set_policy(minimal_residual=2.0)
        flow = g.max_flow()
        assert g.stats.graph_replays == 0
        g.set_policy()
        assert g.max_flow() == flow
        assert g.stats.graph_replays == 0
        g.set_policy(engine=cugip_amd.ENGINE_GRAPH)
        assert g.max_flow() == flow
        assert g.stats.graph_replays > 0
    vol, cs, ct, cn = cugip_amd.make_instance(16, 16, 16, 1, 0, 26)
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn, conn=26)
        g.set_policy(engine=cugip_amd.ENGINE_GRAPH)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        g.set_policy()
        g.max_flow()
        assert g.stats.graph_replays == 0


@pytest.mark.parametrize("conn", [6, 26])
@pytest.mark.parametrize("seed", range(4))
def test_parity_random_asymmetric(conn, seed):
    """Random integer capacities (with zeros) and independent back capacities."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(3, 40, 3))
    say(f"grid random asym {nx}x{ny}x{nz} conn {conn} seed {seed}")
    half = conn // 2
    cs = (rng.integers(0, 200, (nz, ny, nx)) * (rng.random((nz, ny, nx)) < 0.3)).astype(np.int32)
    ct = (rng.integers(0, 200, (nz, ny, nx)) * (rng.random((nz, ny, nx)) < 0.3)).astype(np.int32)
    cn = (rng.integers(0, 60, (half, nz, ny, nx)) * (rng.random((half, nz, ny, nx)) < 0.8)).astype(np.int32)
    cnb = (rng.integers(0, 60, (half, nz, ny, nx)) * (rng.random((half, nz, ny, nx)) < 0.8)).astype(np.int32)
    check_against_oracle(cs, ct, cn, cnb, conn, method="dinic")


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 1, 50), (50, 1, 1), (1, 37, 1), (7, 1, 9), (33, 9, 5), (65, 9, 5)])
def test_degenerate_shapes(dims):
    nx, ny, nz = dims
    say(f"grid degenerate {dims}")
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz)
    cs = rng.integers(0, 50, (nz, ny, nx)).astype(np.int32)
    ct = rng.integers(0, 50, (nz, ny, nx)).astype(np.int32)
    cn = rng.integers(0, 30, (3, nz, ny, nx)).astype(np.int32)
    check_against_oracle(cs, ct, cn, method="dinic")
    check_against_oracle(cs, ct, cn, method="dinic", engine=cugip_amd.ENGINE_GRAPH)
    cn26 = rng.integers(0, 30, (13, nz, ny, nx)).astype(np.int32)
    check_against_oracle(cs, ct, cn26, conn=26, method="dinic")


def test_closed_forms():
    say("grid closed forms")
    rng = np.random.default_rng(3)
    shape = (9, 10, 11)
    cs = rng.integers(0, 100, shape).astype(np.int32)
    ct = rng.integers(0, 100, shape).astype(np.int32)
    zero = np.zeros((3,) + shape, np.int32)
    # no n-links: every voxel is cut on its own
    flow, smin, tmin, _ = gpu_solve(cs, ct, zero)
    assert flow == int(np.minimum(cs, ct).astype(np.int64).sum())
    assert np.array_equal(smin, (cs > ct).reshape(-1)) and np.array_equal(tmin, (ct > cs).reshape(-1))
    # no t-links: no flow at all
    cn = rng.integers(0, 100, (3,) + shape).astype(np.int32)
    flow, smin, tmin, _ = gpu_solve(np.zeros(shape, np.int32), np.zeros(shape, np.int32), cn)
    assert flow == 0 and not smin.any() and not tmin.any()


def test_golden_fixtures(golden):
    for rec in golden["instances"]:
        nx, ny, nz = rec["dims"]
        say(f"golden {rec['dims']} conn {rec['conn']} seed {rec['seed']} instance {rec['instance']}")
        vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, rec["seed"], rec["instance"], rec["conn"])
        flow, smin, tmin, _ = gpu_solve(cs, ct, cn, conn=rec["conn"])
        assert flow == rec["flow"]
        assert digest(smin) == rec["smin_sha256"] and digest(tmin) == rec["tmin_sha256"]


@pytest.mark.parametrize("conn,policy", [
    (26, dict(tile_x=16, tile_y=4, tile_z=4)),
    (26, dict(tile_x=32, tile_y=4, tile_z=2)),
    (26, dict(tile_x=16, tile_y=4, tile_z=4, discharges_per_relabel=5)),
    (6, dict(engine=cugip_amd.ENGINE_GRAPH)),
    (6, dict(tile_x=32, tile_y=2, tile_z=4, discharges_per_relabel=3, max_tile_iterations=4)),
])
def test_golden_multi_workgroup_lists(golden, conn, policy):
    """The large golden instances (256^3 6C, 256x256x192 26C) with tile shapes of more than 32768
    tiles, so every list goes through the multi-workgroup compactions (class regions for the
    discharge, alternating counts, one-launch flag counts), also with an odd number of discharge
    phases per round (the cursor sets then alternate across rounds) and on the device-driven
    engine: flow, S_min and T_min equal the fixtures bit for bit."""
    rec = [r for r in golden["instances"] if r["conn"] == conn and r["dims"][0] >= 256][0]
    nx, ny, nz = rec["dims"]
    say(f"golden multi-workgroup {rec['dims']} conn {conn} policy {policy}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, rec["seed"], rec["instance"], conn)
    flow, smin, tmin, stats = gpu_solve(cs, ct, cn, conn=conn, **policy)
    tiles = -(-nx // stats["tile_x"]) * -(-ny // stats["tile_y"]) * -(-nz // stats["tile_z"])
    say(f"  {tiles} tiles, {stats['rounds']} rounds, {stats['discharge_phases']} phases")
    assert tiles > 32768
    assert flow == rec["flow"]
    assert digest(smin) == rec["smin_sha256"] and digest(tmin) == rec["tmin_sha256"]


def test_golden_repeated_solves_across_tile_shapes(golden):
    """Repeated solves of the 256^3 golden instance on one handle while the tile shape changes
    (16x4x4, 32x2x4, 16x4x4): the relabel learns from each solve how many colour cycles come
    before its first host check (two-row tiles need more) and starts over for a new shape; every
    solve must still give the fixture's flow, S_min and T_min bit for bit."""
    rec = [r for r in golden["instances"] if r["conn"] == 6 and r["dims"][0] >= 256][0]
    nx, ny, nz = rec["dims"]
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, rec["seed"], rec["instance"], 6)
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn)
        for shape in [(16, 4, 4), (32, 2, 4), (16, 4, 4)]:
            g.set_policy(tile_x=shape[0], tile_y=shape[1], tile_z=shape[2])
            for rep in range(3):
                say(f"golden 256^3 tile {shape} solve {rep}")
                flow = g.max_flow()
                stats = g.stats.as_dict()
                assert (stats["tile_x"], stats["tile_y"], stats["tile_z"]) == shape
                assert flow == rec["flow"]
                assert digest(g.source_set()) == rec["smin_sha256"]
                assert digest(g.fill_segments(1, 0)) == rec["tmin_sha256"]


def test_full_size_properties():
    """256^3: the S_min and T_min cuts both have capacity equal to the flow, S_min and T_min are
    disjoint, and repeated solves on one handle agree bit for bit."""
    nx = ny = nz = 256
    say("grid 256^3 properties")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=2)
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn)
        flow = g.max_flow()
        smin = g.source_set()
        tmin = g.fill_segments(1, 0)
        flow2 = g.max_flow()
        assert flow2 == flow and np.array_equal(g.source_set(), smin)
    assert not (smin.astype(bool) & tmin.astype(bool)).any()
    assert grid_cut_capacity(smin, cs, ct, cn) == flow
    assert grid_cut_capacity(1 - tmin, cs, ct, cn) == flow


def test_device_pointer_entry():
    from tests.hip_rt import DeviceArray
    nx, ny, nz = 48, 40, 24
    say("grid device pointers")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed=12)
    dev = [DeviceArray(a) for a in (cs, ct, cn)]
    with cugip_amd.GridGraph() as g:
        g.set_grid_device((nz, ny, nx), 6, dev[0].address, dev[1].address, dev[2].address)
        flow = g.max_flow()
        smin = g.source_set()
    for d in dev:
        d.free()
    ref_flow, ref_smin, _ = oracle.grid_solve("bk", nx, ny, nz, 6, cs, ct, cn)
    assert flow == ref_flow and np.array_equal(smin, ref_smin)


def test_errors():
    say("grid errors")
    shape = (4, 5, 6)
    ones = np.ones(shape, np.int32)
    cn = np.ones((3,) + shape, np.int32)
    with cugip_amd.GridGraph() as g:
        with pytest.raises(cugip_amd.CgpError) as e:
            g.source_set()
        assert e.value.code == cugip_amd.CGP_ERR_STATE
        g.set_grid(-ones, ones, cn)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        big = np.full(shape, 2 ** 31 - 1, np.int32)
        g.set_grid(big, ones, cn)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_OVERFLOW
        g.set_grid(ones, ones, cn)
        g.set_policy(tile_x=7, tile_y=3, tile_z=1)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        g.set_policy()  # defaults
        assert g.max_flow() == int(np.prod(shape))
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.GridGraph().set_grid(ones, ones, cn, conn=8)


@pytest.mark.parametrize("conn", [6, 26])
def test_capacity_errors_found_by_the_solve(conn):
    """The init kernel's checks (negative capacity, int32 headroom) reach the caller as errors, also
    though the solver reads them only with the relabel's first count; the handle then solves a
    valid grid as usual."""
    say(f"grid capacity errors conn {conn}")
    vol, cs, ct, cn = cugip_amd.make_instance(40, 24, 16, 2, 0, conn)
    flow = oracle.grid_solve("bk", 40, 24, 16, conn, cs, ct, cn)[0]
    with cugip_amd.GridGraph() as g:
        bad = cs.copy()
        bad.reshape(-1)[123] = -1
        g.set_grid(bad, ct, cn, conn=conn)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        with pytest.raises(cugip_amd.CgpError):
            g.source_set()  # no result after a failed solve
        big, zero = cs.copy(), ct.copy()
        big.reshape(-1)[77] = 2 ** 31 - 1  # plus the incoming n-links: beyond int32
        zero.reshape(-1)[77] = 0
        g.set_grid(big, zero, cn, conn=conn)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_OVERFLOW
        g.set_grid(cs, ct, cn, conn=conn)
        assert g.max_flow() == flow


def test_round_cap_reports_not_converged():
    say("grid round cap")
    vol, cs, ct, cn = cugip_amd.make_instance(64, 64, 32, seed=1)
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn)
        g.set_policy(max_rounds=1, discharges_per_relabel=1, max_tile_iterations=1)
        with pytest.raises(cugip_amd.CgpError) as e:
            g.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_NOT_CONVERGED
        with pytest.raises(cugip_amd.CgpError):
            g.source_set()  # no result after a failed solve
        g.set_policy()
        flow = g.max_flow()
    assert flow == oracle.grid_solve("bk", 64, 64, 32, 6, cs, ct, cn)[0]


def test_cpp_dropin_grid():
    import os
    import subprocess
    say("cpp grid program")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "graph_cut_grid")
    out = subprocess.run(["timeout", "-k", "5", "120", exe, "53", "41", "29", "6", "17"], capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    vol, cs, ct, cn = cugip_amd.make_instance(53, 41, 29, seed=17)
    flow, smin, tmin = oracle.grid_solve("bk", 53, 41, 29, 6, cs, ct, cn)
    assert out.stdout.split() == ["flow", str(flow), "source_side", str(int(smin.sum())), "sink_side",
                                  str(int(tmin.sum()))]


@pytest.mark.parametrize("conn", [6, 26])
def test_mfi_instance(tmp_path, conn):
    """An MFI file (asymmetric capacities) solved through cgp_set_grid_mfi: the flow equals the
    header's max flow, and S_min / T_min equal the oracle's on the converted grid."""
    say(f"grid mfi conn {conn}")
    rng = np.random.default_rng(40 + conn)
    shape = (23, 31, 37)
    nz, ny, nx = shape
    cs = (rng.integers(0, 400, shape) * (rng.random(shape) < 0.25)).astype(np.int32)
    ct = (rng.integers(0, 400, shape) * (rng.random(shape) < 0.25)).astype(np.int32)
    cn = rng.integers(0, 90, (conn // 2,) + shape).astype(np.int32)
    cnb = rng.integers(0, 90, (conn // 2,) + shape).astype(np.int32)
    ref_flow, ref_smin, ref_tmin = oracle.grid_solve("dinic", nx, ny, nz, conn, cs, ct, cn, cnb)
    path = str(tmp_path / "i.mfi")
    cugip_amd.write_mfi(path, cs, ct, cn, cnb, conn=conn, maxflow=ref_flow)
    with cugip_amd.GridGraph() as g:
        info = g.set_grid_mfi(path)
        assert info["maxflow"] == ref_flow and info["grid_conn"] == conn
        assert g.max_flow() == info["maxflow"]
        assert np.array_equal(g.source_set(), ref_smin)
        assert np.array_equal(g.fill_segments(1, 0), ref_tmin)


def frontend_caps(vol, markers, w, conn):
    """numpy restatement of the device front end (t-link tables, marker then border forcing,
    n-link tables by |I_v - I_u| and neighbour class)."""
    from tests.gpu_util import HALF6, HALF26, _shift
    vol = vol.astype(np.int64)
    nz, ny, nx = vol.shape
    cs = np.array(w.tlink_source, np.int64)[vol]
    ct = np.array(w.tlink_sink, np.int64)[vol]
    if markers is not None:
        src = markers == w.marker_source
        snk = (markers == w.marker_sink) & ~src
        cs[src], ct[src] = w.marker_capacity, 0
        cs[snk], ct[snk] = 0, w.marker_capacity
    b = w.border
    if b > 0:
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        m = (x < b) | (y < b) | (z < b) | (x >= nx - b) | (y >= ny - b) | (z >= nz - b)
        cs[m], ct[m] = w.marker_capacity, 0
    table = np.array([list(w.nlink[c]) for c in range(3)], np.int64)
    offs = HALF6 if conn == 6 else HALF26
    cn = np.zeros((len(offs),) + vol.shape, np.int64)
    for j, (dx, dy, dz) in enumerate(offs):
        nb = _shift(vol, dx, dy, dz, -1)
        cls = (dx != 0) + (dy != 0) + (dz != 0) - 1
        inside = nb >= 0
        cn[j][inside] = table[cls][np.abs(vol - nb)[inside]]
    return cs.astype(np.int32), ct.astype(np.int32), cn.astype(np.int32)


@pytest.mark.parametrize("conn", [6, 26])
def test_device_front_end_default_tables(conn):
    """Capacities computed on the device from the volume equal the host generator's, bit for bit,
    and the solve matches the one from host capacities."""
    say(f"grid device front end conn {conn}")
    vol, cs, ct, cn = cugip_amd.make_instance(61, 47, 35, 21, 0, conn)
    with cugip_amd.GridGraph() as g:
        g.set_grid_volume(vol, conn=conn)
        dcs, dct, dcn, dcnb = g.capacities()
        assert np.array_equal(dcs, cs) and np.array_equal(dct, ct)
        assert np.array_equal(dcn, cn) and np.array_equal(dcnb, cn)
        flow = g.max_flow()
        smin = g.source_set()
    ref_flow, ref_smin, _ = oracle.grid_solve("bk", 61, 47, 35, conn, cs, ct, cn)
    assert flow == ref_flow and np.array_equal(smin, ref_smin)


def test_device_front_end_markers_and_border():
    """The reference's scheme: zero data t-links, markers 128 (source) / 255 (sink), border 2
    forced to the source, contrast n-links; checked against the numpy restatement and the oracle."""
    say("grid device front end markers")
    rng = np.random.default_rng(3)
    vol = cugip_amd.generate_volume(40, 36, 28, 5)
    markers = np.zeros(vol.shape, np.uint8)
    markers[rng.random(vol.shape) < 0.01] = 128
    markers[(rng.random(vol.shape) < 0.01) & (vol > 150)] = 255
    w = cugip_amd.default_volume_weights()
    for i in range(256):
        w.tlink_source[i] = 0
        w.tlink_sink[i] = 0
    w.marker_source, w.marker_sink, w.marker_capacity, w.border = 128, 255, 1000, 2
    cs, ct, cn = frontend_caps(vol, markers, w, 6)
    with cugip_amd.GridGraph() as g:
        g.set_grid_volume(vol, markers, w, conn=6)
        dcs, dct, dcn, _ = g.capacities()
        assert np.array_equal(dcs, cs) and np.array_equal(dct, ct) and np.array_equal(dcn, cn)
        flow = g.max_flow()
        smin = g.source_set()
    nz, ny, nx = vol.shape
    ref_flow, ref_smin, _ = oracle.grid_solve("bk", nx, ny, nz, 6, cs, ct, cn)
    assert flow == ref_flow and np.array_equal(smin, ref_smin)


def test_grid_state_and_graphml(tmp_path):
    """State volumes after a solve: every voxel with excess is in S_min, labels equal cgp_labels,
    and every edge leaving S_min is saturated (its direction bit is set). The GraphML dump, parsed
    back, has the same max flow as the grid."""
    import xml.etree.ElementTree as ET
    from tests.gpu_util import HALF6
    say("grid state and graphml")
    nx, ny, nz = 19, 13, 11
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, 33)
    seen = []
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn)

        def hook(event, iteration, done, flow, levels):
            if event == cugip_amd.TRACE_AFTER_PUSH and not seen:
                seen.append(g.grid_state()[0].copy())

        g.set_trace(hook)
        flow = g.max_flow()
        g.set_trace(None)
        smin = g.source_set().astype(bool)
        excess, labels, sat = g.grid_state()
        assert np.array_equal(labels, g.labels())
        path = str(tmp_path / "g.graphml")
        g.save_graphml(path)
    assert seen and seen[0].shape == (nx * ny * nz,)
    assert not (excess > 0)[~smin].any()
    S = smin.reshape(nz, ny, nx)
    satv = sat.reshape(nz, ny, nx)
    for j, (dx, dy, dz) in enumerate(HALF6):
        z, y, x = np.nonzero(S)
        ux, uy, uz = x + dx, y + dy, z + dz
        ok = (ux >= 0) & (uy >= 0) & (uz >= 0)
        leaving = ok.copy()
        leaving[ok] = ~S[uz[ok], uy[ok], ux[ok]] & (cn[j][z[ok], y[ok], x[ok]] > 0)
        assert ((satv[z[leaving], y[leaving], x[leaving]] >> j) & 1).all()
    ns = {"g": "http://graphml.graphdrawing.org/xmlns"}
    root = ET.parse(path).getroot()
    ea, eb, cap, src, snk = [], [], [], np.zeros(nx * ny * nz, np.int64), np.zeros(nx * ny * nz, np.int64)
    for e in root.iter("{%s}edge" % ns["g"]):
        a, b = e.get("source"), e.get("target")
        c = int(e.find("g:data", ns).text)
        if a == "source":
            src[int(b[1:])] = c
        elif b == "sink":
            snk[int(a[1:])] = c
        else:
            ea.append(int(a[1:]))
            eb.append(int(b[1:]))
            cap.append(c)
    assert len(root.findall(".//g:node", ns)) == nx * ny * nz + 2
    ref = oracle.dinic(nx * ny * nz, np.array(ea), np.array(eb), np.array(cap), np.zeros(len(cap)), src, snk)[0]
    assert ref == flow


def test_cut_capacity_entry():
    """cgp_cut_capacity against the numpy cut capacity on random masks (asymmetric capacities,
    both connectivities), and on S_min / V minus T_min, whose capacity is the flow."""
    say("grid cut capacity entry")
    rng = np.random.default_rng(77)
    for conn in (6, 26):
        shape = (13, 17, 19)
        half = conn // 2
        cs = rng.integers(0, 300, shape).astype(np.int32)
        ct = rng.integers(0, 300, shape).astype(np.int32)
        cn = rng.integers(0, 90, (half,) + shape).astype(np.int32)
        cnb = rng.integers(0, 90, (half,) + shape).astype(np.int32)
        with cugip_amd.GridGraph() as g:
            g.set_grid(cs, ct, cn, cnb, conn=conn)
            for _ in range(3):
                mask = (rng.random(shape) < 0.5).astype(np.uint8)
                assert g.cut_capacity(mask) == grid_cut_capacity(mask, cs, ct, cn, cnb, conn)
            flow = g.max_flow()
            assert g.cut_capacity(g.source_set()) == flow
            assert g.cut_capacity(1 - g.fill_segments(1, 0)) == flow


def test_6c_above_the_one_workgroup_compaction():
    """A 6-connected grid with more tiles than one workgroup compacts (61440 tiles of the 16x8x4
    shape): the relabel and closure lists take the multi-workgroup compaction, the discharge lists
    the ordered pieces path. Capacities on the device from the volume; S_min and V minus T_min are
    both minimum cuts, nested, and the flow and S_min match a solve with 16x4x4 tiles (122880 tiles,
    four pieces a thread in the ordered compaction)."""
    nx, ny, nz = 384, 320, 256
    say("grid 6C 384x320x256")
    vol = cugip_amd.generate_volume(nx, ny, nz, seed=3)
    with cugip_amd.GridGraph() as g:
        g.set_grid_volume(vol, conn=6)
        g.set_policy(tile_x=16, tile_y=8, tile_z=4)
        flow = g.max_flow()
        smin = g.source_set()
        tmin = g.fill_segments(1, 0)
        assert g.cut_capacity(smin) == flow
        assert g.cut_capacity(1 - tmin) == flow
        assert not (smin.astype(bool) & tmin.astype(bool)).any()
        g.set_policy(tile_x=16, tile_y=4, tile_z=4)
        assert g.max_flow() == flow
        assert digest(g.source_set()) == digest(smin)


def test_full_size_26c_512x512x384():
    """The 26-connected 512x512x384 workload (about 100 M voxels, capacities computed on the device
    from the volume): the S_min and the V minus T_min cuts both have the flow's capacity (so both are
    minimum cuts and the flow is maximum), they are nested, and a second solve gives the same bits."""
    import time
    nx, ny, nz = 512, 512, 384
    say("grid 26C 512x512x384")
    vol = cugip_amd.generate_volume(nx, ny, nz, seed=0)
    with cugip_amd.GridGraph() as g:
        g.set_grid_volume(vol, conn=26)
        t0 = time.perf_counter()
        flow = g.max_flow()
        say(f"  solve {1e3 * (time.perf_counter() - t0):.1f} ms, flow {flow}, stats {g.stats.as_dict()}")
        smin = g.source_set()
        tmin = g.fill_segments(1, 0)
        assert g.cut_capacity(smin) == flow
        assert g.cut_capacity(1 - tmin) == flow
        assert not (smin.astype(bool) & tmin.astype(bool)).any()
        assert g.max_flow() == flow
        assert digest(g.source_set()) == digest(smin)


@pytest.mark.parametrize("conn,dims,seed,engine", [
    (6, (48, 40, 36), 0, cugip_amd.ENGINE_GRAPH),
    (6, (48, 40, 36), 1, cugip_amd.ENGINE_TILES),
    (6, (64, 64, 64), 2, cugip_amd.ENGINE_AUTO),
    (26, (32, 28, 20), 3, cugip_amd.ENGINE_AUTO),
    (26, (30, 22, 17), 4, cugip_amd.ENGINE_TILES),
])
def test_labels_are_sink_distances(conn, dims, seed, engine):
    """cgp_labels after a solve are the exact BFS distances to t in the final residual graph: the
    oracle's relabel run on the residual planes the device holds (cgp_grid_residuals). The
    residuals also keep every edge's capacity pair: r(v->u) + r(u->v) = c(v->u) + c(u->v)."""
    nx, ny, nz = dims
    say(f"labels vs residual BFS {dims} conn {conn} seed {seed} engine {engine}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed, 0, conn)
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn, conn=conn)
        g.set_policy(engine=engine)
        g.max_flow()
        labels = g.labels().reshape(-1)
        r, rt = g.residuals()
    ref = oracle.residual_labels(nx, ny, nz, conn, r, rt)
    assert np.array_equal(labels, ref), int((labels != ref).sum())
    assert (ref < cugip_amd.LABEL_UNREACHABLE).any() and (ref == cugip_amd.LABEL_UNREACHABLE).any()
    H = conn // 2
    for j, (dx, dy, dz) in enumerate(oracle.half_offsets(conn)):
        zd = slice(max(0, -dz), nz - max(0, dz))
        yd = slice(max(0, -dy), ny - max(0, dy))
        xd = slice(max(0, -dx), nx - max(0, dx))
        zs, ys, xs = (slice(a.start + o, a.stop + o) for a, o in ((zd, dz), (yd, dy), (xd, dx)))
        pair = r[j][zd, yd, xd].astype(np.int64) + r[j + H][zs, ys, xs]
        assert np.array_equal(pair, 2 * cn[j][zd, yd, xd].astype(np.int64)), j


@pytest.mark.parametrize("conn,dims,seed", [(6, (40, 36, 28), 5), (26, (24, 20, 16), 6)])
def test_first_relabel_labels(conn, dims, seed):
    """The labels of the first global relabel (read from the trace hook after it, before any push)
    equal the oracle's restatement of the reference's first BFS (initBFSKernel + level sweeps,
    graph_cut_relabeling.hpp:47-112) on the initial residual graph, the sink t-links taken after the
    per-voxel cancellation min(cs, ct) that the init kernel sends straight to t."""
    nx, ny, nz = dims
    say(f"first relabel labels {dims} conn {conn} seed {seed}")
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed, 0, conn)
    captured = {}
    with cugip_amd.GridGraph() as g:
        g.set_grid(cs, ct, cn, conn=conn)

        def hook(event, iteration, done, flow, starts):
            if event == cugip_amd.TRACE_AFTER_RELABEL and iteration == 0 and "labels" not in captured:
                captured["labels"] = g.grid_state()[1].copy()

        g.set_trace(hook)
        g.max_flow()
        g.set_trace(None)
    ct2 = (ct.astype(np.int64) - np.minimum(cs, ct)).astype(np.int32)
    ea, eb, fwd, bwd = oracle.grid_edges(nx, ny, nz, conn, cn)
    ref = oracle.initial_labels(nx * ny * nz, ea, eb, fwd, bwd, ct2.reshape(-1))
    assert np.array_equal(captured["labels"], ref), int((captured["labels"] != ref).sum())
