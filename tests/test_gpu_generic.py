"""Generic-graph path (CSR, the reference's Graph<TFlow> data layout) on the GPU: the reference's
known-answer graphs through the Python binding and through the C++ drop-in header, random graphs
against the oracle, edge cases, errors and trace hooks."""
import os
import subprocess

import numpy as np
import pytest

import cugip_amd
import oracle
from tests.golden import kats
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

CPP = os.path.join(os.path.dirname(__file__), "cpp")


def solve(g, **policy):
    n, ea, eb, fwd, bwd, cs, ct = g
    with cugip_amd.Graph() as gr:
        gr.set_vertex_count(n)
        gr.set_nweights(np.stack([ea, eb], 1), fwd, bwd)
        gr.set_tweights(cs, ct)
        if policy:
            gr.set_policy(**policy)
        flow = gr.max_flow()
        return flow, gr.source_set(), gr.fill_segments(1, 0)


def test_kats_python():
    for name, (g, want) in [("4x4", kats.grid_4x4()), ("layered", kats.layered(False)),
                            ("layered flipped", kats.layered(True)), ("1000^2", kats.grid_1000())]:
        say(f"generic KAT {name}")
        flow, smin, tmin = solve(g)
        assert flow == want, name
        ref = oracle.dinic(*g)
        assert np.array_equal(smin, ref[1]) and np.array_equal(tmin, ref[2]), name


def test_bfs_seed_count_kat():
    """unit_tests/graph_cut_test.cu:102-169: the first relabel of the two-layer graph seeds 16 vertices."""
    say("generic BFS seeds")
    (n, ea, eb, fwd, bwd, cs, ct), seeds = kats.two_layer()
    starts = []
    with cugip_amd.Graph() as gr:
        gr.set_vertex_count(n)
        gr.set_nweights(np.stack([ea, eb], 1), fwd, bwd)
        gr.set_tweights(cs, ct)
        gr.set_trace(lambda ev, it, done, flow, ls: starts.append(ls) if ev == cugip_amd.TRACE_AFTER_RELABEL else None)
        gr.max_flow()
    assert starts and starts[0][:2] == [0, seeds]


def test_cpp_kats():
    say("cpp KAT program")
    out = subprocess.run(["timeout", "-k", "5", "300", os.path.join(CPP, "graph_cut_kats")], capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "MISMATCH" not in out.stdout and out.stdout.count(" ok") == 5


def test_cpp_trace_and_policy():
    say("cpp trace program")
    out = subprocess.run(["timeout", "-k", "5", "120", os.path.join(CPP, "graph_cut_trace"), "11"],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "source_side 33" in out.stdout


def random_graph(rng, n, m, cap):
    ea = rng.integers(0, n, m).astype(np.int32)
    eb = rng.integers(0, n, m).astype(np.int32)
    fwd = rng.integers(0, cap, m).astype(np.int32)
    bwd = rng.integers(0, cap, m).astype(np.int32)
    cs = (rng.integers(0, 3 * cap, n) * (rng.random(n) < 0.2)).astype(np.int32)
    ct = (rng.integers(0, 3 * cap, n) * (rng.random(n) < 0.2)).astype(np.int32)
    return (n, ea, eb, fwd, bwd, cs, ct)


@pytest.mark.parametrize("seed", range(8))
def test_random_graphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 5000))
    m = int(rng.integers(0, 4 * n))
    g = random_graph(rng, n, m, int(rng.integers(1, 1000)))
    say(f"generic random n {n} m {m} seed {seed}")
    # EdgeRecord normalises every pair to first = min; swap the capacities accordingly
    n_, ea, eb, fwd, bwd, cs, ct = g
    swap = ea > eb
    lo, hi = np.where(swap, eb, ea), np.where(swap, ea, eb)
    f2, b2 = np.where(swap, bwd, fwd), np.where(swap, fwd, bwd)
    g2 = (n_, lo, hi, f2, b2, cs, ct)
    flow, smin, tmin = solve(g2, **({"minimal_residual": 5.0} if seed % 3 == 2 else {}))
    ref = oracle.dinic(*g2)
    assert flow == ref[0] and np.array_equal(smin, ref[1]) and np.array_equal(tmin, ref[2])


def test_edge_cases():
    say("generic edge cases")
    # no edges at all
    g = (5, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
         np.array([3, 0, 7, 1, 0], np.int32), np.array([1, 5, 7, 0, 0], np.int32))
    flow, smin, tmin = solve(g)
    assert flow == 1 + 0 + 7 + 0 + 0
    # loops, duplicates and zero capacities
    g = (4, np.array([0, 0, 1, 1, 2], np.int32), np.array([0, 1, 2, 2, 3], np.int32),
         np.array([9, 4, 2, 3, 0], np.int32), np.array([9, 0, 1, 0, 0], np.int32),
         np.array([10, 0, 0, 0], np.int32), np.array([0, 0, 10, 10], np.int32))
    flow, smin, tmin = solve(g)
    ref = oracle.dinic(*g)
    assert flow == ref[0] == 4 and np.array_equal(smin, ref[1]) and np.array_equal(tmin, ref[2])


def test_errors():
    say("generic errors")
    with cugip_amd.Graph() as gr:
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.set_tweights(np.zeros(0), np.zeros(0))
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_STATE
        gr.set_vertex_count(3)
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.set_nweights(np.array([[0, 3]]), [1], [1])
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.set_nweights(np.array([[0, 1]]), [-1], [1])
        assert e.value.code == cugip_amd.CGP_ERR_INVALID
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.set_nweights(np.array([[0, 1]]), [2 ** 31 - 1], [1])
        assert e.value.code == cugip_amd.CGP_ERR_OVERFLOW
        gr.set_nweights(np.array([[0, 1], [1, 2]]), [5, 5], [5, 5])
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.max_flow()  # t-links missing
        assert e.value.code == cugip_amd.CGP_ERR_STATE
        gr.set_tweights([2 ** 31 - 1, 0, 0], [0, 0, 1])
        with pytest.raises(cugip_amd.CgpError) as e:
            gr.max_flow()
        assert e.value.code == cugip_amd.CGP_ERR_OVERFLOW
        gr.set_tweights([7, 0, 0], [0, 0, 9])
        assert gr.max_flow() == 5


def test_trace_hooks():
    say("generic trace hooks")
    g, want = kats.grid_4x4()
    events = []
    n, ea, eb, fwd, bwd, cs, ct = g
    with cugip_amd.Graph() as gr:
        gr.set_vertex_count(n)
        gr.set_nweights(np.stack([ea, eb], 1), fwd, bwd)
        gr.set_tweights(cs, ct)
        gr.set_trace(lambda ev, it, done, flow, ls: events.append((ev, it, done, flow)))
        assert gr.max_flow() == want
    kinds = [e[0] for e in events]
    assert kinds[0] == cugip_amd.TRACE_STARTED and kinds[-1] == cugip_amd.TRACE_FINISHED
    assert kinds.count(cugip_amd.TRACE_BEGIN_ITERATION) == kinds.count(cugip_amd.TRACE_AFTER_RELABEL)
    assert events[-1][3] == want and events[-1][2] == 1


@pytest.mark.parametrize("n,seed", [(300, 2), (5000, 3)])
def test_region_graph_file(tmp_path, n, seed):
    """A region adjacency graph file and marker file in the reference's format, read by
    cgp_region_graph_read and solved by the generic GPU path: flow, S_min and the object mask
    (fill_segments, the regions that still reach t) equal Dinic's on the independently parsed
    graph."""
    from tests.region_util import parse_region_graph, write_random_region_graph
    say(f"region graph file n={n} seed {seed}")
    g, mk = tmp_path / "g.txt", tmp_path / "m.txt"
    write_random_region_graph(g, mk, n, seed)
    nv, edges, w, cs, ct = cugip_amd.read_region_graph(str(g), str(mk), 1000.0, 10000)
    with cugip_amd.Graph() as gr:
        gr.set_vertex_count(nv)
        gr.set_nweights(edges, w, w)
        gr.set_tweights(cs, ct)
        flow = gr.max_flow()
        smin, tmin = gr.source_set(), gr.fill_segments(1, 0)
    rn, ea, eb, rw, rcs, rct = parse_region_graph(g, mk, 1000.0, 10000)
    ref_flow, ref_smin, ref_tmin = oracle.dinic(rn, ea, eb, rw, rw, rcs, rct)
    assert flow == ref_flow
    assert np.array_equal(smin, ref_smin) and np.array_equal(tmin, ref_tmin)
    assert 0 < tmin.sum() < nv


def test_large_lattice_graph_multi_workgroup_levels():
    """A 64^3-vertex lattice graph with one random extra edge per vertex (parallel edges and both
    orientations among them, asymmetric capacities): BFS and push levels far above one workgroup's
    size, so the multi-workgroup level kernels and their wave-aggregated queue appends and slotted
    flow sums run; flow, S_min and T_min equal Boykov-Kolmogorov's."""
    rng = np.random.default_rng(11)
    k = 64
    n = k ** 3
    z, y, x = (c.reshape(-1) for c in np.indices((k, k, k)))
    v = np.arange(n)
    ea, eb = [], []
    for d, ok in ((1, x + 1 < k), (k, y + 1 < k), (k * k, z + 1 < k)):
        ea.append(v[ok])
        eb.append(v[ok] + d)
    off = rng.integers(-2, 3, (n, 3))
    u = (np.clip(z + off[:, 2], 0, k - 1) * k + np.clip(y + off[:, 1], 0, k - 1)) * k + np.clip(x + off[:, 0], 0, k - 1)
    ea.append(u[u != v])  # (reversed orientation on purpose)
    eb.append(v[u != v])
    ea, eb = np.concatenate(ea), np.concatenate(eb)
    fwd = rng.integers(1, 100, ea.size).astype(np.int32)
    bwd = rng.integers(1, 100, ea.size).astype(np.int32)
    cs = (rng.integers(0, 400, n) * (rng.random(n) < 0.1)).astype(np.int32)
    ct = (rng.integers(0, 400, n) * (rng.random(n) < 0.1)).astype(np.int32)
    swap = ea > eb
    g = (n, np.where(swap, eb, ea).astype(np.int32), np.where(swap, ea, eb).astype(np.int32),
         np.where(swap, bwd, fwd).astype(np.int32), np.where(swap, fwd, bwd).astype(np.int32), cs, ct)
    say(f"generic lattice graph n {n} m {ea.size}")
    flow, smin, tmin = solve(g)
    ref_flow, ref_smin, ref_tmin, _ = oracle.bk(*g)
    assert flow == ref_flow
    assert np.array_equal(smin, ref_smin) and np.array_equal(tmin, ref_tmin)
