"""MFI instance files on the CPU: the library's writer and reader against the numpy restatement of
the format and against the reference's own reader (benchmarks/mfi.cpp, built into oracle/_ref by
oracle/Makefile when the reference tree is present), and the arc directions of the MFI -> grid
conversion, pinned by max flows computed straight from the arcs."""
import os
import subprocess

import numpy as np
import pytest

import cugip_amd
import oracle
from oracle import mfi as omfi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "mfi_dump")


def random_grid(rng, shape, conn, sym):
    nz, ny, nx = shape
    cs = (rng.integers(0, 300, shape) * (rng.random(shape) < 0.3)).astype(np.int32)
    ct = (rng.integers(0, 300, shape) * (rng.random(shape) < 0.3)).astype(np.int32)
    cn = rng.integers(0, 80, (conn // 2,) + shape).astype(np.int32)
    cnb = None if sym else rng.integers(0, 80, (conn // 2,) + shape).astype(np.int32)
    return cs, ct, cn, cnb


@pytest.mark.parametrize("conn,sym", [(6, True), (6, False), (26, False)])
def test_write_read_round_trip(tmp_path, conn, sym):
    rng = np.random.default_rng(conn + sym)
    shape = (5, 7, 9)
    cs, ct, cn, cnb = random_grid(rng, shape, conn, sym)
    flow = oracle.grid_solve("dinic", 9, 7, 5, conn, cs, ct, cn, cnb)[0]
    path = str(tmp_path / "g.mfi")
    cugip_amd.write_mfi(path, cs, ct, cn, cnb, conn=conn, maxflow=flow, metadata="synthetic")
    m = omfi.read(path)
    assert (m["dimension"], m["connectivity"], m["width"], m["height"], m["depth"]) == (3, conn, 9, 7, 5)
    assert m["maxflow"] == flow and m["symmetric_arcs"] == (1 if sym else 0) and m["metadata"] == b"synthetic"
    assert np.array_equal(m["source"], cs) and np.array_equal(m["sink"], ct)
    # flow straight from the file's arcs: pins every arc's direction and plane
    assert omfi.solve_arcs(m) == flow
    info, cs2, ct2, cn2, cnb2, gconn = cugip_amd.mfi_to_grid(path)
    assert gconn == conn and info["maxflow"] == flow
    assert np.array_equal(cs2, cs) and np.array_equal(ct2, ct)
    # back planes: the symmetric file stores the forward planes twice; edges leaving the volume are 0
    _, _, fwd_ref, bwd_ref = oracle.grid_edges(9, 7, 5, conn, cn, cnb)
    _, _, fwd, bwd = oracle.grid_edges(9, 7, 5, conn, cn2, cnb2)
    assert np.array_equal(fwd, fwd_ref) and np.array_equal(bwd, bwd_ref)


@pytest.mark.parametrize("dimension,connectivity,tt,nt", [
    (2, 4, 0, 0), (2, 8, 1, 0), (3, 6, 1, 1), (3, 18, 2, 1), (3, 26, 0, 2)])
def test_conversion_matches_arc_flow(tmp_path, dimension, connectivity, tt, nt):
    """Files of every supported kind and capacity type, written by the numpy restatement: the grid
    the library converts them to has the max flow of their arcs taken one by one."""
    rng = np.random.default_rng(dimension * 100 + connectivity)
    shape = (1, 11, 13) if dimension == 2 else (4, 6, 7)
    hi_t = 250 if tt == 0 else 3000
    hi_n = 250 if nt == 0 else 3000
    source = (rng.integers(0, hi_t, shape) * (rng.random(shape) < 0.3))
    sink = (rng.integers(0, hi_t, shape) * (rng.random(shape) < 0.3))
    arcs = rng.integers(0, hi_n, (connectivity,) + shape) * (rng.random((connectivity,) + shape) < 0.7)
    path = str(tmp_path / "a.mfi")
    omfi.write(path, dimension, connectivity, source, sink, arcs, maxflow=0, terminal_type=tt, neighbor_type=nt)
    m = omfi.read(path)
    want = omfi.solve_arcs(m)
    info, cs, ct, cn, cnb, conn = cugip_amd.mfi_to_grid(path)
    assert conn == (6 if connectivity in (4, 6) else 26)
    assert (info["width"], info["height"], info["depth"]) == (shape[2], shape[1], shape[0])
    nz, ny, nx = cs.shape
    assert oracle.grid_solve("dinic", nx, ny, nz, conn, cs, ct, cn, cnb)[0] == want


def test_reference_reader_agrees(tmp_path):
    if not os.access(REF_DUMP, os.X_OK):
        pytest.skip("oracle/_ref/mfi_dump not built (reference tree absent)")
    rng = np.random.default_rng(7)
    for conn, sym in ((6, True), (26, False)):
        cs, ct, cn, cnb = random_grid(rng, (3, 4, 5), conn, sym)
        path = str(tmp_path / ("r%d.mfi" % conn))
        cugip_amd.write_mfi(path, cs, ct, cn, cnb, conn=conn, maxflow=123456789012)
        out = subprocess.run([REF_DUMP, path], capture_output=True, text=True, check=True).stdout.split("\n")
        m = omfi.read(path)
        assert out[0].split() == [str(v) for v in (3, conn, 5, 4, 3, 123456789012, 1 if sym else 0, 0, 2, 2)]
        planes = [m["source"], m["sink"]] + list(m["arcs"]) + [m["active"].astype(np.int64)]
        for line, plane in zip(out[1:], planes):
            flat = plane.reshape(-1).astype(np.uint64)
            s = int(flat.sum())
            w = int((flat * np.arange(1, flat.size + 1, dtype=np.uint64)).sum())
            assert line.split() == [str(s), str(w)]


def test_errors(tmp_path):
    bad = tmp_path / "bad.mfi"
    bad.write_bytes(b"XYZ" + b"\0" * 100)
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.mfi_info(str(bad))
    assert e.value.code == cugip_amd.CGP_ERR_INVALID and "not an MFI file" in str(e.value)
    with pytest.raises(cugip_amd.CgpError):
        cugip_amd.mfi_info(str(tmp_path / "missing.mfi"))
    shape = (2, 3, 4)
    path = str(tmp_path / "t.mfi")
    omfi.write(path, 3, 6, np.ones(shape), np.ones(shape), np.ones((6,) + shape))
    data = open(path, "rb").read()
    (tmp_path / "trunc.mfi").write_bytes(data[:-10])
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.mfi_to_grid(str(tmp_path / "trunc.mfi"))
    assert "truncated" in str(e.value)
    omfi.write(path, 3, 10, np.ones(shape), np.ones(shape), np.ones((10,) + shape))
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.mfi_info(path)
    assert "unsupported" in str(e.value)
    big = np.full(shape, 2 ** 32 - 1, np.uint64)
    omfi.write(path, 3, 6, big, np.ones(shape), np.ones((6,) + shape))
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.mfi_to_grid(path)
    assert e.value.code == cugip_amd.CGP_ERR_OVERFLOW
