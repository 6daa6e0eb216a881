"""CPU oracle of the push-relabel graph cut -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may import this
module, and only as the checker (or as the timed CPU baseline). The product package
cugip_amd never imports it.

ctypes binding of oracle/liboracle.so (built from oracle/oracle.c by oracle/Makefile):
  refpr       restatement of the reference's relabel/push loop
              (include/cugip/advanced_operations/detail/graph_cut_implementation.hpp:122-198)
  dinic       independent textbook max flow
  bk          Boykov-Kolmogorov, the timed CPU baseline
  bruteforce  enumeration of every source side, n <= 24
All take a generic graph: n vertices, edge arrays ea/eb/fwd/bwd (fwd = ea -> eb), t-links
cs/ct, int32; flows are exact int64. grid_edges() turns the SoA n-link planes of a 3D grid
into such an edge list.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class RefprStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("pushes", ctypes.c_int64), ("max_level", ctypes.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        graph_args = [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.orc_refpr.restype = ctypes.c_int64
        L.orc_refpr.argtypes = graph_args + [_u8p, _u8p, _i32p, ctypes.POINTER(RefprStats)]
        L.orc_dinic.restype = ctypes.c_int64
        L.orc_dinic.argtypes = graph_args + [_u8p, _u8p]
        L.orc_bk.restype = ctypes.c_int64
        L.orc_bk.argtypes = graph_args + [_u8p, _u8p, ctypes.POINTER(ctypes.c_double)]
        L.orc_bruteforce.restype = ctypes.c_int64
        L.orc_bruteforce.argtypes = graph_args + [_u8p, _u8p]
        L.orc_bfs_seed_count.restype = ctypes.c_int64
        L.orc_bfs_seed_count.argtypes = [ctypes.c_int32, _i32p]
        L.orc_initial_labels.restype = ctypes.c_int
        L.orc_initial_labels.argtypes = [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.orc_grid_edges.restype = ctypes.c_int64
        L.orc_grid_edges.argtypes = [ctypes.c_int32] * 4 + [_i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p32(a):
    return a.ctypes.data_as(_i32p)


def _p8(a):
    return a.ctypes.data_as(_u8p)


def _graph(n, ea, eb, fwd, bwd, cs, ct):
    arrs = [_i32(x) for x in (ea, eb, fwd, bwd, cs, ct)]
    m = len(arrs[0])
    assert all(len(a) == m for a in arrs[:4]) and len(arrs[4]) == n and len(arrs[5]) == n
    return arrs, m


def _check(flow, name):
    if flow < 0:
        raise ValueError(f"{name}: invalid graph (code {flow})")
    return int(flow)


def refpr(n, ea, eb, fwd, bwd, cs, ct):
    """Reference-algorithm restatement: returns dict(flow, smin, tmin, labels, iterations, pushes, max_level)."""
    arrs, m = _graph(n, ea, eb, fwd, bwd, cs, ct)
    smin = np.zeros(n, np.uint8)
    tmin = np.zeros(n, np.uint8)
    labels = np.zeros(n, np.int32)
    st = RefprStats()
    flow = lib().orc_refpr(n, m, *[_p32(a) for a in arrs], _p8(smin), _p8(tmin), _p32(labels), ctypes.byref(st))
    return dict(flow=_check(flow, "refpr"), smin=smin, tmin=tmin, labels=labels, iterations=st.iterations,
                pushes=st.pushes, max_level=st.max_level)


def dinic(n, ea, eb, fwd, bwd, cs, ct):
    arrs, m = _graph(n, ea, eb, fwd, bwd, cs, ct)
    smin = np.zeros(n, np.uint8)
    tmin = np.zeros(n, np.uint8)
    flow = lib().orc_dinic(n, m, *[_p32(a) for a in arrs], _p8(smin), _p8(tmin))
    return _check(flow, "dinic"), smin, tmin


def bk(n, ea, eb, fwd, bwd, cs, ct):
    """Boykov-Kolmogorov: returns (flow, smin, tmin, seconds of the maxflow loop only)."""
    arrs, m = _graph(n, ea, eb, fwd, bwd, cs, ct)
    smin = np.zeros(n, np.uint8)
    tmin = np.zeros(n, np.uint8)
    secs = ctypes.c_double(0.0)
    flow = lib().orc_bk(n, m, *[_p32(a) for a in arrs], _p8(smin), _p8(tmin), ctypes.byref(secs))
    return _check(flow, "bk"), smin, tmin, secs.value


def bruteforce(n, ea, eb, fwd, bwd, cs, ct):
    arrs, m = _graph(n, ea, eb, fwd, bwd, cs, ct)
    if n > 24:
        raise ValueError("bruteforce needs n <= 24")
    smin = np.zeros(n, np.uint8)
    tmin = np.zeros(n, np.uint8)
    flow = lib().orc_bruteforce(n, m, *[_p32(a) for a in arrs], _p8(smin), _p8(tmin))
    return _check(flow, "bruteforce"), smin, tmin


def bfs_seed_count(ct):
    ct = _i32(ct)
    return int(lib().orc_bfs_seed_count(len(ct), _p32(ct)))


def initial_labels(n, ea, eb, fwd, bwd, ct):
    ea, eb, fwd, bwd, ct = (_i32(x) for x in (ea, eb, fwd, bwd, ct))
    out = np.zeros(n, np.int32)
    rc = lib().orc_initial_labels(n, len(ea), _p32(ea), _p32(eb), _p32(fwd), _p32(bwd), _p32(ct), _p32(out))
    if rc:
        raise RuntimeError("initial_labels failed")
    return out


def grid_edges(nx, ny, nz, conn, cap_n, cap_nb=None):
    """Edge list of a grid: cap_n (and cap_nb) are (conn/2, nz, ny, nx) int32 planes."""
    cap_n = _i32(cap_n).reshape(-1)
    nb = _i32(cap_nb).reshape(-1) if cap_nb is not None else None
    cap = (conn // 2) * nx * ny * nz
    assert cap_n.size == cap
    ea = np.empty(cap, np.int32)
    eb = np.empty(cap, np.int32)
    fwd = np.empty(cap, np.int32)
    bwd = np.empty(cap, np.int32)
    m = lib().orc_grid_edges(nx, ny, nz, conn, _p32(cap_n), _p32(nb) if nb is not None else None,
                             _p32(ea), _p32(eb), _p32(fwd), _p32(bwd))
    if m < 0:
        raise ValueError("grid_edges: bad connectivity")
    if m == cap:
        return ea, eb, fwd, bwd
    return ea[:m].copy(), eb[:m].copy(), fwd[:m].copy(), bwd[:m].copy()


def half_offsets(conn):
    """(dx, dy, dz) of the negative half neighbourhood in the grid layout's plane order
    (VonNeumannNeighborhood<3>::offset(1..3), MooreNeighborhood<3>::offset(1..13),
    include/cugip/neighborhood.hpp:51-91, 160-262)."""
    if conn == 6:
        return [(-1, 0, 0), (0, -1, 0), (0, 0, -1)]
    return [(j % 3 - 1, (j // 3) % 3 - 1, j // 9 - 1) for j in range(13)]


def residual_labels(nx, ny, nz, conn, residuals, sink_residual):
    """Labels of an exact global relabel of a grid's residual graph: BFS distance to t of every
    voxel (1 where the residual v -> t is positive, then 1 + the lowest label of a neighbour u with
    residual v -> u > 0), 0x3fffffff where t is unreachable. residuals: conn planes, plane d the
    residual v -> v + o_d (d < conn/2 the half offsets, d >= conn/2 their mirrors). The BFS is the
    oracle's restatement of the reference's relabel (orc_initial_labels) on the edge list of the
    residual planes."""
    H = conn // 2
    r = _i32(residuals).reshape(conn, nz, ny, nx)
    back = np.zeros((H, nz, ny, nx), np.int32)
    for j, (dx, dy, dz) in enumerate(half_offsets(conn)):
        # back[j] at v = residual (v + o_j) -> v, which plane j + H holds at voxel v + o_j
        zd = slice(max(0, -dz), nz - max(0, dz))
        yd = slice(max(0, -dy), ny - max(0, dy))
        xd = slice(max(0, -dx), nx - max(0, dx))
        zs = slice(zd.start + dz, zd.stop + dz)
        ys = slice(yd.start + dy, yd.stop + dy)
        xs = slice(xd.start + dx, xd.stop + dx)
        back[j][zd, yd, xd] = r[j + H][zs, ys, xs]
    ea, eb, fwd, bwd = grid_edges(nx, ny, nz, conn, r[:H], back)
    return initial_labels(nx * ny * nz, ea, eb, fwd, bwd, _i32(sink_residual).reshape(-1))


def grid_solve(method, nx, ny, nz, conn, cs, ct, cap_n, cap_nb=None):
    """Solve a grid instance with one of refpr/dinic/bk; returns (flow, smin, tmin) with the
    masks as flat uint8 arrays in voxel order."""
    ea, eb, fwd, bwd = grid_edges(nx, ny, nz, conn, cap_n, cap_nb)
    n = nx * ny * nz
    cs = _i32(cs).reshape(-1)
    ct = _i32(ct).reshape(-1)
    if method == "refpr":
        r = refpr(n, ea, eb, fwd, bwd, cs, ct)
        return r["flow"], r["smin"], r["tmin"]
    if method == "dinic":
        return dinic(n, ea, eb, fwd, bwd, cs, ct)
    if method == "bk":
        f, s, t, _ = bk(n, ea, eb, fwd, bwd, cs, ct)
        return f, s, t
    raise ValueError(method)
