"""Python host binding of the MI355X push-relabel graph cut.

A thin ctypes layer over the C ABI of libcugip_hip.so (include/cugip_hip.h), mirroring the
reference's cugip::Graph<TFlow> interface (include/cugip/advanced_operations/graph_cut.hpp:47-146):
same method names, same argument meaning, errors raised as exceptions. The C++ header
include/cugip/graph_cut.hpp is the drop-in for C++ callers; this module serves Python
callers, the tests and bench.py.

The library must have been built (python -c "import __graft_entry__ as g; g.build()"); there is
no fallback: without it, importing this module raises ImportError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcugip_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build the HIP library first (__graft_entry__.build())")

_lib = ctypes.CDLL(LIB_PATH)

CGP_OK = 0
CGP_ERR_INVALID = 1
CGP_ERR_NOMEM = 2
CGP_ERR_NOT_CONVERGED = 3
CGP_ERR_OVERFLOW = 4
CGP_ERR_DEVICE = 5
CGP_ERR_STATE = 6
LABEL_UNREACHABLE = 0x3FFFFFFF

TRACE_STARTED, TRACE_BEGIN_ITERATION, TRACE_AFTER_RELABEL, TRACE_AFTER_PUSH, TRACE_FINISHED = range(5)
ENGINE_AUTO, ENGINE_TILES, ENGINE_GRAPH = 0, 1, 2


class Edge(ctypes.Structure):
    """cgp_edge / cugip::EdgeRecord layout: second at offset 0, first at offset 4."""
    _fields_ = [("second", ctypes.c_uint32), ("first", ctypes.c_uint32)]


class Policy(ctypes.Structure):
    """cgp_policy; zero fields take the library default (cugip::GraphCutPolicy counterpart)."""
    _fields_ = [
        ("tile_x", ctypes.c_int32), ("tile_y", ctypes.c_int32), ("tile_z", ctypes.c_int32),
        ("discharges_per_relabel", ctypes.c_int32),
        ("max_tile_iterations", ctypes.c_int32),
        ("max_rounds", ctypes.c_int32),
        ("minimal_residual", ctypes.c_float),
        ("engine", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("flow", ctypes.c_int64),
        ("rounds", ctypes.c_int32), ("bfs_passes", ctypes.c_int32),
        ("discharge_phases", ctypes.c_int32), ("cut_passes", ctypes.c_int32),
        ("bfs_tile_visits", ctypes.c_int64), ("discharge_tile_visits", ctypes.c_int64),
        ("cut_tile_visits", ctypes.c_int64),
        ("init_ms", ctypes.c_double), ("relabel_ms", ctypes.c_double), ("discharge_ms", ctypes.c_double),
        ("cut_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("voxels", ctypes.c_int64),
        ("tile_x", ctypes.c_int32), ("tile_y", ctypes.c_int32), ("tile_z", ctypes.c_int32), ("conn", ctypes.c_int32),
        ("graph_replays", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class GridInstance(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("reserved0", ctypes.c_int32),
                ("cap_source", ctypes.c_void_p), ("cap_sink", ctypes.c_void_p),
                ("cap_n", ctypes.c_void_p), ("cap_n_back", ctypes.c_void_p)]


class MfiInfo(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_int32), ("connectivity", ctypes.c_int32),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("depth", ctypes.c_int32),
                ("symmetric_arcs", ctypes.c_int32), ("maxflow", ctypes.c_int64),
                ("active_count", ctypes.c_int32), ("type_terminal_cap", ctypes.c_int32),
                ("type_neighbor_cap", ctypes.c_int32), ("grid_conn", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4), ("metadata", ctypes.c_char * 4096)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "metadata")}
        d["metadata"] = self.metadata
        return d


class VolumeWeights(ctypes.Structure):
    _fields_ = [("tlink_source", ctypes.c_int32 * 256), ("tlink_sink", ctypes.c_int32 * 256),
                ("nlink", (ctypes.c_int32 * 256) * 3), ("marker_source", ctypes.c_int32),
                ("marker_sink", ctypes.c_int32), ("marker_capacity", ctypes.c_int32), ("border", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]


class KernelProfile(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("voxel_visits", ctypes.c_int64),
        ("device_ms", ctypes.c_double), ("algorithmic_bytes", ctypes.c_double), ("bytes_per_voxel", ctypes.c_double),
    ]


TraceFn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int32)

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

_lib.cgp_create.argtypes = [ctypes.POINTER(_vp)]
_lib.cgp_destroy.argtypes = [_vp]
_lib.cgp_last_error.argtypes = [_vp]
_lib.cgp_last_error.restype = ctypes.c_char_p
_lib.cgp_set_vertex_count.argtypes = [_vp, ctypes.c_int32]
_lib.cgp_set_nweights.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(Edge), _i32p, _i32p]
_lib.cgp_set_tweights.argtypes = [_vp, _i32p, _i32p]
_lib.cgp_set_grid.argtypes = [_vp] + [ctypes.c_int32] * 4 + [_i32p] * 4
_lib.cgp_set_grid_device.argtypes = [_vp] + [ctypes.c_int32] * 4 + [_vp] * 4
_lib.cgp_set_policy.argtypes = [_vp, ctypes.POINTER(Policy)]
_lib.cgp_get_policy.argtypes = [_vp, ctypes.POINTER(Policy)]
_lib.cgp_set_trace.argtypes = [_vp, TraceFn, _vp]
_lib.cgp_max_flow.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Stats)]
_lib.cgp_fill_segments.argtypes = [_vp, _u8p, ctypes.c_uint8, ctypes.c_uint8]
_lib.cgp_source_set.argtypes = [_vp, _u8p]
_lib.cgp_labels.argtypes = [_vp, _i32p]
_lib.cgp_generate_volume.argtypes = [ctypes.c_int32] * 3 + [ctypes.c_uint64, ctypes.c_uint64, _u8p]
_lib.cgp_caps_from_volume.argtypes = [ctypes.c_int32] * 4 + [_u8p, _i32p, _i32p, _i32p]
_lib.cgp_device_count.argtypes = [_i32p]
_lib.cgp_set_profiling.argtypes = [_vp, ctypes.c_int32]
_lib.cgp_get_profile.argtypes = [_vp, ctypes.POINTER(KernelProfile), ctypes.c_int32, _i32p]
_lib.cgp_abi_version.restype = ctypes.c_int32
_cp = ctypes.c_char_p
_lib.cgp_mfi_read_info.argtypes = [_cp, ctypes.POINTER(MfiInfo)]
_lib.cgp_mfi_to_grid.argtypes = [_cp, _i32p, _i32p, _i32p, _i32p]
_lib.cgp_set_grid_mfi.argtypes = [_vp, _cp, ctypes.POINTER(MfiInfo)]
_lib.cgp_write_mfi.argtypes = [_cp] + [ctypes.c_int32] * 4 + [_i32p] * 4 + [ctypes.c_int64, _cp]
_lib.cgp_default_volume_weights.argtypes = [ctypes.POINTER(VolumeWeights)]
_lib.cgp_set_grid_volume.argtypes = [_vp] + [ctypes.c_int32] * 4 + [_vp, _vp, ctypes.POINTER(VolumeWeights),
                                                                    ctypes.c_int32]
_lib.cgp_grid_capacities.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p, _i32p, _i32p]
_lib.cgp_grid_state.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p, ctypes.POINTER(ctypes.c_uint32)]
_lib.cgp_grid_residuals.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p]
_lib.cgp_region_graph_info.argtypes = [ctypes.c_char_p, _i32p, _i32p]
_lib.cgp_region_graph_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int32,
                                       ctypes.POINTER(Edge), _i32p, _i32p, _i32p]
_lib.cgp_grid_max_flow_noalloc.argtypes = [ctypes.c_int32] * 4 + [_vp] * 5 + [ctypes.POINTER(ctypes.c_int64), _vp]
_lib.cgp_save_graphml.argtypes = [_vp, ctypes.c_char_p]
_lib.cgp_cut_capacity.argtypes = [_vp, ctypes.c_int32, _u8p, ctypes.POINTER(ctypes.c_int64)]
_lib.cgp_cnn_create.argtypes = [ctypes.POINTER(_vp)]
_lib.cgp_cnn_destroy.argtypes = [_vp]
_lib.cgp_cnn_last_error.argtypes = [_vp]
_lib.cgp_cnn_last_error.restype = ctypes.c_char_p
_fp = ctypes.POINTER(ctypes.c_float)
_lib.cgp_cnn_set_params.argtypes = [_vp, _fp, ctypes.c_int32]
_lib.cgp_cnn_get_params.argtypes = [_vp, _fp, ctypes.c_int32]
_lib.cgp_cnn_predict.argtypes = [_vp, _vp, _u8p, _u8p, _fp, ctypes.POINTER(ctypes.c_int64)]
_lib.cgp_cnn_gradients.argtypes = [_vp, _vp, _u8p, ctypes.POINTER(ctypes.c_double), _fp]
_lib.cgp_cnn_train_step.argtypes = [_vp, _vp, _u8p, ctypes.c_float, ctypes.POINTER(ctypes.c_double)]
_lib.cgp_max_flow_certified.argtypes = [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Stats),
                                        ctypes.POINTER(ctypes.c_int32)]
_lib.cgp_set_grid_batch.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(GridInstance),
                                    ctypes.c_int32]
_lib.cgp_batch_size.argtypes = [_vp, _i32p]
_lib.cgp_solve_batch.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Stats)]
_lib.cgp_batch_source_set.argtypes = [_vp, ctypes.c_int32, _u8p]
_lib.cgp_batch_fill_segments.argtypes = [_vp, ctypes.c_int32, _u8p, ctypes.c_uint8, ctypes.c_uint8]


class CgpError(RuntimeError):
    """Error status of the C ABI (the reference raises cugip::ExceptionBase, exception.hpp:20-46)."""

    def __init__(self, code, message):
        super().__init__(f"[code {code}] {message}")
        self.code = code


def library():
    return _lib


def abi_version():
    return int(_lib.cgp_abi_version())


def device_count():
    n = ctypes.c_int32(0)
    rc = _lib.cgp_device_count(ctypes.byref(n))
    return int(n.value) if rc == CGP_OK else 0


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p32(a):
    return a.ctypes.data_as(_i32p)


class _Handle:
    def __init__(self):
        h = _vp()
        rc = _lib.cgp_create(ctypes.byref(h))
        if rc != CGP_OK:
            raise CgpError(rc, _lib.cgp_last_error(None).decode())
        self._h = h
        self._trace_cb = None
        self.stats = None
        self._n = 0

    def close(self):
        if getattr(self, "_h", None):
            _lib.cgp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != CGP_OK:
            raise CgpError(rc, _lib.cgp_last_error(self._h).decode())

    def set_policy(self, policy=None, **fields):
        if policy is None:
            policy = Policy()
        for k, v in fields.items():
            setattr(policy, k, v)
        self._check(_lib.cgp_set_policy(self._h, ctypes.byref(policy)))

    def get_policy(self):
        p = Policy()
        self._check(_lib.cgp_get_policy(self._h, ctypes.byref(p)))
        return p

    def set_trace(self, fn):
        """fn(event, iteration, done, flow, level_starts) or None."""
        if fn is None:
            self._trace_cb = None
            self._check(_lib.cgp_set_trace(self._h, TraceFn(), None))
            return

        def _cb(_user, event, iteration, done, flow, levels, count):
            starts = [levels[i] for i in range(count)] if levels and count > 0 else []
            fn(event, iteration, done, flow, starts)

        self._trace_cb = TraceFn(_cb)
        self._check(_lib.cgp_set_trace(self._h, self._trace_cb, None))

    def max_flow(self, policy=None):
        """Exact max-flow value (Python int). Stats of the solve land in self.stats."""
        if policy is not None:
            self.set_policy(policy)
        flow = ctypes.c_int64(0)
        st = Stats()
        self._check(_lib.cgp_max_flow(self._h, ctypes.byref(flow), ctypes.byref(st)))
        self.stats = st
        return int(flow.value)

    def fill_segments(self, mask=255, background=0):
        """Reference segmentation (T_min): mask where the vertex still reaches t."""
        out = np.empty(self._n, np.uint8)
        self._check(_lib.cgp_fill_segments(self._h, out.ctypes.data_as(_u8p), mask, background))
        return out

    def source_set(self):
        """Canonical minimum cut S_min as 0/1 bytes."""
        out = np.empty(self._n, np.uint8)
        self._check(_lib.cgp_source_set(self._h, out.ctypes.data_as(_u8p)))
        return out

    def set_profiling(self, enable=True):
        """Per-kernel event timing of the next solves (serialises them)."""
        self._check(_lib.cgp_set_profiling(self._h, 1 if enable else 0))

    def profile(self):
        """Per-kernel records of the last solve: list of dicts (name, launches, voxel_visits, device_ms,
        algorithmic_bytes, bytes_per_voxel)."""
        arr = (KernelProfile * 16)()
        n = ctypes.c_int32(0)
        self._check(_lib.cgp_get_profile(self._h, arr, 16, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            r = arr[i]
            out.append(dict(name=r.name.decode(), launches=r.launches, voxel_visits=r.voxel_visits,
                            device_ms=r.device_ms, algorithmic_bytes=r.algorithmic_bytes,
                            bytes_per_voxel=r.bytes_per_voxel))
        return out

    def labels(self):
        out = np.empty(self._n, np.int32)
        self._check(_lib.cgp_labels(self._h, _p32(out)))
        return out

    def cut_capacity(self, source_side, index=0):
        """Exact capacity of the cut with the given source-side mask (computed on the device)."""
        side = np.ascontiguousarray(source_side, dtype=np.uint8).reshape(-1)
        want = int(np.prod(self.shapes[index])) if hasattr(self, "shapes") else self._n
        if side.size != want:
            raise CgpError(CGP_ERR_INVALID, "mask has %d entries, the grid %d voxels" % (side.size, want))
        cap = ctypes.c_int64(0)
        self._check(_lib.cgp_cut_capacity(self._h, index, side.ctypes.data_as(_u8p), ctypes.byref(cap)))
        return int(cap.value)

    def max_flow_certified(self, certificate):
        """Solve that stops once the flow into t reaches `certificate` (a cut capacity); returns
        (flow, stopped_early)."""
        flow = ctypes.c_int64(0)
        early = ctypes.c_int32(0)
        st = Stats()
        self._check(_lib.cgp_max_flow_certified(self._h, int(certificate), ctypes.byref(flow), ctypes.byref(st),
                                                ctypes.byref(early)))
        self.stats = st
        return int(flow.value), bool(early.value)

    def save_graphml(self, path):
        """GraphML dump of the graph (Graph::save_to_graphml)."""
        self._check(_lib.cgp_save_graphml(self._h, os.fsencode(path)))

    def grid_state(self, index=0):
        """(excess, labels, saturated direction bits) of grid `index` once a solve has started,
        e.g. from a trace hook; flat arrays in voxel order."""
        n = int(np.prod(self.shapes[index])) if hasattr(self, "shapes") else self._n
        e = np.empty(n, np.int32)
        lab = np.empty(n, np.int32)
        sat = np.empty(n, np.uint32)
        self._check(_lib.cgp_grid_state(self._h, index, _p32(e), _p32(lab),
                                        sat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return e, lab, sat


    def residuals(self, index=0):
        """(residuals, sink_residual) of grid `index` once a solve has started: the conn planes of
        the residual v -> v + o_d, shaped (conn, nz, ny, nx), and the residual v -> t, (nz, ny, nx)."""
        shape = tuple(self.shapes[index]) if hasattr(self, "shapes") else tuple(self.shape)
        r = np.empty((self.conn,) + shape, np.int32)
        rt = np.empty(shape, np.int32)
        self._check(_lib.cgp_grid_residuals(self._h, index, _p32(r), _p32(rt)))
        return r, rt


class Graph(_Handle):
    """Generic graph: cugip::Graph<TFlow> (set_vertex_count / set_nweights / set_tweights / max_flow /
    fill_segments, graph_cut.hpp:54-107)."""

    def set_vertex_count(self, n):
        self._check(_lib.cgp_set_vertex_count(self._h, int(n)))
        self._n = int(n)

    def set_nweights(self, edges, fwd, bwd):
        """edges: (m, 2) vertex pairs (normalised to first = min, second = max as EdgeRecord does);
        fwd: capacity first -> second of the normalised record, bwd: second -> first."""
        edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        m = edges.shape[0]
        rec = (Edge * max(m, 1))()
        if m:
            lo = np.minimum(edges[:, 0], edges[:, 1]).astype(np.uint32)
            hi = np.maximum(edges[:, 0], edges[:, 1]).astype(np.uint32)
            arr = np.frombuffer(rec, dtype=np.uint32, count=2 * m).reshape(m, 2)
            arr[:, 0] = hi
            arr[:, 1] = lo
        fwd = _i32(fwd).reshape(-1)
        bwd = _i32(bwd).reshape(-1)
        if fwd.size != m or bwd.size != m:
            raise CgpError(CGP_ERR_INVALID, "fwd and bwd must hold one capacity per edge")
        self._check(_lib.cgp_set_nweights(self._h, m, rec, _p32(fwd), _p32(bwd)))

    def set_tweights(self, source, sink):
        source = _i32(source)
        sink = _i32(sink)
        if source.size != self._n or sink.size != self._n:
            raise CgpError(CGP_ERR_INVALID, "t-link arrays must have one entry per vertex")
        self._check(_lib.cgp_set_tweights(self._h, _p32(source), _p32(sink)))


class GridGraph(_Handle):
    """3D grid graph with implicit 6- or 26-connected edges (cgp_set_grid)."""

    def set_grid(self, cap_source, cap_sink, cap_n, cap_n_back=None, conn=6):
        cs = _i32(cap_source)
        if cs.ndim != 3:
            raise CgpError(CGP_ERR_INVALID, "cap_source must be a (nz, ny, nx) volume")
        nz, ny, nx = cs.shape
        ct = _i32(cap_sink)
        cn = _i32(cap_n)
        if ct.shape != cs.shape or cn.shape != (conn // 2, nz, ny, nx):
            raise CgpError(CGP_ERR_INVALID, "capacity shapes do not match the volume")
        cb = _i32(cap_n_back) if cap_n_back is not None else None
        if cb is not None and cb.shape != cn.shape:
            raise CgpError(CGP_ERR_INVALID, "cap_n_back shape does not match cap_n")
        self._check(_lib.cgp_set_grid(self._h, nx, ny, nz, conn, _p32(cs), _p32(ct), _p32(cn),
                                      _p32(cb) if cb is not None else None))
        self._n = nx * ny * nz
        self.shape = (nz, ny, nx)
        self.conn = conn

    def set_grid_volume(self, volume, markers=None, weights=None, conn=6):
        """Capacities computed on the device from a uint8 (nz, ny, nx) volume (and optional marker
        volume) with the lookup tables of `weights` (VolumeWeights; None = default_volume_weights())."""
        vol = np.ascontiguousarray(volume, dtype=np.uint8)
        if vol.ndim != 3:
            raise CgpError(CGP_ERR_INVALID, "volume must be (nz, ny, nx)")
        mk = None
        if markers is not None:
            mk = np.ascontiguousarray(markers, dtype=np.uint8)
            if mk.shape != vol.shape:
                raise CgpError(CGP_ERR_INVALID, "markers must match the volume")
        nz, ny, nx = vol.shape
        self._check(_lib.cgp_set_grid_volume(self._h, nx, ny, nz, conn, vol.ctypes.data,
                                             mk.ctypes.data if mk is not None else None,
                                             ctypes.byref(weights) if weights is not None else None, 0))
        self._n = nx * ny * nz
        self.shape = (nz, ny, nx)
        self.conn = conn

    def capacities(self):
        """Input capacities held by the handle: (cap_source, cap_sink, cap_n, cap_n_back)."""
        return _capacities(self, 0, self.shape, self.conn)

    def set_grid_mfi(self, path):
        """Loads an MFI instance file; returns its header as a dict."""
        info = MfiInfo()
        self._check(_lib.cgp_set_grid_mfi(self._h, os.fsencode(path), ctypes.byref(info)))
        self._n = info.width * info.height * info.depth
        self.shape = (info.depth, info.height, info.width)
        self.conn = info.grid_conn
        return info.as_dict()

    def set_grid_device(self, shape, conn, cs_ptr, ct_ptr, cn_ptr, cnb_ptr=None):
        """Capacities already resident in device memory (raw device pointers, e.g. tensor.data_ptr())."""
        nz, ny, nx = shape
        self._check(_lib.cgp_set_grid_device(self._h, nx, ny, nz, conn, _vp(cs_ptr), _vp(ct_ptr), _vp(cn_ptr),
                                             _vp(cnb_ptr) if cnb_ptr else None))
        self._n = nx * ny * nz
        self.shape = (nz, ny, nx)
        self.conn = conn


def _capacities(handle, index, shape, conn):
    cs = np.empty(shape, np.int32)
    ct = np.empty(shape, np.int32)
    cn = np.empty((conn // 2,) + tuple(shape), np.int32)
    cnb = np.empty((conn // 2,) + tuple(shape), np.int32)
    handle._check(_lib.cgp_grid_capacities(handle._h, index, _p32(cs), _p32(ct), _p32(cn), _p32(cnb)))
    return cs, ct, cn, cnb


def default_volume_weights():
    """The lookup tables of the synthetic instances as front-end weights (VolumeWeights)."""
    w = VolumeWeights()
    _global_check(_lib.cgp_default_volume_weights(ctypes.byref(w)))
    return w


class GridBatch(_Handle):
    """Independent 3D grids of one connectivity solved together (cgp_set_grid_batch / cgp_solve_batch)."""

    def set_batch(self, grids, conn=6):
        """grids: sequence of (cap_source, cap_sink, cap_n[, cap_n_back]) with (nz, ny, nx) volumes."""
        keep = []
        arr = (GridInstance * max(len(grids), 1))()
        shapes = []
        for i, g in enumerate(grids):
            cs, ct, cn = (_i32(a) for a in g[:3])
            cb = _i32(g[3]) if len(g) > 3 and g[3] is not None else None
            if cs.ndim != 3 or ct.shape != cs.shape or cn.shape != (conn // 2,) + cs.shape or \
                    (cb is not None and cb.shape != cn.shape):
                raise CgpError(CGP_ERR_INVALID, "grid %d: capacity shapes do not match the volume" % i)
            nz, ny, nx = cs.shape
            keep += [cs, ct, cn, cb]
            arr[i] = GridInstance(nx, ny, nz, 0, cs.ctypes.data, ct.ctypes.data, cn.ctypes.data,
                                  cb.ctypes.data if cb is not None else None)
            shapes.append((nz, ny, nx))
        self._check(_lib.cgp_set_grid_batch(self._h, len(grids), conn, arr, 0))
        self.shapes = shapes
        self.conn = conn
        self._n = sum(a * b * c for a, b, c in shapes)

    def set_batch_device(self, shapes, conn, pointers):
        """pointers: per grid (cs_ptr, ct_ptr, cn_ptr[, cnb_ptr]) raw device pointers."""
        arr = (GridInstance * max(len(shapes), 1))()
        for i, ((nz, ny, nx), ptrs) in enumerate(zip(shapes, pointers)):
            cnb = ptrs[3] if len(ptrs) > 3 else None
            arr[i] = GridInstance(nx, ny, nz, 0, ptrs[0], ptrs[1], ptrs[2], cnb)
        self._check(_lib.cgp_set_grid_batch(self._h, len(shapes), conn, arr, 1))
        self.shapes = [tuple(s) for s in shapes]
        self.conn = conn
        self._n = sum(a * b * c for a, b, c in self.shapes)

    def solve(self, policy=None):
        """Per-grid exact max-flow values (list of int). Stats of the solve land in self.stats."""
        if policy is not None:
            self.set_policy(policy)
        flows = (ctypes.c_int64 * max(len(self.shapes), 1))()
        st = Stats()
        self._check(_lib.cgp_solve_batch(self._h, flows, ctypes.byref(st)))
        self.stats = st
        return [int(flows[i]) for i in range(len(self.shapes))]

    def _shape_of(self, index):
        if not 0 <= index < len(self.shapes):
            raise CgpError(CGP_ERR_INVALID, "grid index %d out of range (batch of %d)" % (index, len(self.shapes)))
        return self.shapes[index]

    def batch_source_set(self, index):
        """Canonical minimum cut S_min of grid `index`, 0/1 bytes shaped like its volume."""
        out = np.empty(self._shape_of(index), np.uint8)
        self._check(_lib.cgp_batch_source_set(self._h, index, out.ctypes.data_as(_u8p)))
        return out

    def capacities(self, index):
        """Input capacities of grid `index`: (cap_source, cap_sink, cap_n, cap_n_back)."""
        return _capacities(self, index, self._shape_of(index), self.conn)

    def batch_fill_segments(self, index, mask=255, background=0):
        """T_min of grid `index` (fill_segments semantics), shaped like its volume."""
        out = np.empty(self._shape_of(index), np.uint8)
        self._check(_lib.cgp_batch_fill_segments(self._h, index, out.ctypes.data_as(_u8p), mask, background))
        return out


def _global_check(rc):
    if rc != CGP_OK:
        raise CgpError(rc, (_lib.cgp_last_error(None) or b"").decode())


def mfi_info(path):
    """Header of an MFI instance file as a dict."""
    info = MfiInfo()
    _global_check(_lib.cgp_mfi_read_info(os.fsencode(path), ctypes.byref(info)))
    return info.as_dict()


def mfi_to_grid(path):
    """An MFI file in the grid layout: (info, cap_source, cap_sink, cap_n, cap_n_back, conn)."""
    info = mfi_info(path)
    shape = (info["depth"], info["height"], info["width"])
    conn = info["grid_conn"]
    cs = np.empty(shape, np.int32)
    ct = np.empty(shape, np.int32)
    cn = np.empty((conn // 2,) + shape, np.int32)
    cnb = np.empty((conn // 2,) + shape, np.int32)
    _global_check(_lib.cgp_mfi_to_grid(os.fsencode(path), _p32(cs), _p32(ct), _p32(cn), _p32(cnb)))
    return info, cs, ct, cn, cnb, conn


def write_mfi(path, cap_source, cap_sink, cap_n, cap_n_back=None, conn=6, maxflow=0, metadata=""):
    """Writes a 3D grid (layout of set_grid) as an MFI file with uint32 capacities."""
    cs, ct, cn = _i32(cap_source), _i32(cap_sink), _i32(cap_n)
    nz, ny, nx = cs.shape
    cb = _i32(cap_n_back) if cap_n_back is not None else None
    if ct.shape != cs.shape or cn.shape != (conn // 2, nz, ny, nx) or (cb is not None and cb.shape != cn.shape):
        raise CgpError(CGP_ERR_INVALID, "capacity shapes do not match the volume")
    _global_check(_lib.cgp_write_mfi(os.fsencode(path), nx, ny, nz, conn, _p32(cs), _p32(ct), _p32(cn),
                                     _p32(cb) if cb is not None else None, int(maxflow), metadata.encode()))


CNN_PARAMS = 22545


def _vol_ptr(vol):
    return vol.ctypes.data_as(_u8p) if vol is not None else None


class Cnn:
    """Learned cut proposal (cgp_cnn_*): a 3D CNN run on the matrix cores. Parameter layout: see
    cugip_hip.h. Prediction, gradients and training act on grid 0 of a GridGraph plus its volume."""

    def __init__(self):
        h = _vp()
        rc = _lib.cgp_cnn_create(ctypes.byref(h))
        if rc != CGP_OK:
            raise CgpError(rc, (_lib.cgp_last_error(None) or b"").decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.cgp_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != CGP_OK:
            raise CgpError(rc, _lib.cgp_cnn_last_error(self._h).decode())

    @staticmethod
    def init_params(seed=0):
        """He-initialised parameters (numpy float32), the padding input channel zero."""
        rng = np.random.default_rng(seed)
        p = np.zeros(CNN_PARAMS, np.float32)
        w1 = rng.normal(0, np.sqrt(2.0 / (27 * 3)), (27, 4, 16)).astype(np.float32)
        w1[:, 3, :] = 0
        p[:27 * 4 * 16] = w1.reshape(-1)
        off = 27 * 4 * 16 + 16
        for _ in range(3):
            p[off:off + 27 * 256] = rng.normal(0, np.sqrt(2.0 / (27 * 16)), 27 * 256)
            off += 27 * 256 + 16
        p[off:off + 16] = rng.normal(0, np.sqrt(1.0 / 16), 16)
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(_lib.cgp_cnn_set_params(self._h, p.ctypes.data_as(_fp), p.size))

    def get_params(self):
        p = np.empty(CNN_PARAMS, np.float32)
        self._check(_lib.cgp_cnn_get_params(self._h, p.ctypes.data_as(_fp), p.size))
        return p

    @staticmethod
    def _volume(graph, volume):
        if volume is None:
            return None  # the volume the grid was built from (set_grid_volume), on the device
        vol = np.ascontiguousarray(volume, dtype=np.uint8).reshape(-1)
        if vol.size != graph._n:
            raise CgpError(CGP_ERR_INVALID, "volume does not match the graph's grid")
        return vol

    def predict(self, graph, volume):
        """(side 0/1 bytes, probabilities, certificate = capacity of the predicted cut)."""
        vol = self._volume(graph, volume)
        side = np.empty(graph._n, np.uint8)
        prob = np.empty(graph._n, np.float32)
        cert = ctypes.c_int64(0)
        self._check(_lib.cgp_cnn_predict(self._h, graph._h, _vol_ptr(vol), side.ctypes.data_as(_u8p),
                                         prob.ctypes.data_as(_fp), ctypes.byref(cert)))
        return side, prob, int(cert.value)

    def certificate(self, graph, volume=None):
        """Capacity of the predicted cut only (nothing copied back but the number)."""
        vol = self._volume(graph, volume)
        cert = ctypes.c_int64(0)
        self._check(_lib.cgp_cnn_predict(self._h, graph._h, _vol_ptr(vol), None, None, ctypes.byref(cert)))
        return int(cert.value)

    def gradients(self, graph, volume):
        """(loss, gradient) against the graph's S_min (the graph must be solved)."""
        vol = self._volume(graph, volume)
        grad = np.empty(CNN_PARAMS, np.float32)
        loss = ctypes.c_double(0)
        self._check(_lib.cgp_cnn_gradients(self._h, graph._h, _vol_ptr(vol), ctypes.byref(loss),
                                           grad.ctypes.data_as(_fp)))
        return float(loss.value), grad

    def train_step(self, graph, volume, learning_rate=1e-3):
        vol = self._volume(graph, volume)
        loss = ctypes.c_double(0)
        self._check(_lib.cgp_cnn_train_step(self._h, graph._h, _vol_ptr(vol), float(learning_rate),
                                            ctypes.byref(loss)))
        return float(loss.value)


def generate_volume(nx, ny, nz, seed, instance=0):
    """Synthetic blob volume, uint8 (nz, ny, nx)."""
    out = np.empty((nz, ny, nx), np.uint8)
    rc = _lib.cgp_generate_volume(nx, ny, nz, seed, instance, out.ctypes.data_as(_u8p))
    if rc != CGP_OK:
        raise CgpError(rc, "cgp_generate_volume failed")
    return out


def caps_from_volume(volume, conn=6):
    """Integer capacities of a volume: (cap_source, cap_sink, cap_n[conn/2])."""
    vol = np.ascontiguousarray(volume, dtype=np.uint8)
    nz, ny, nx = vol.shape
    cs = np.empty(vol.shape, np.int32)
    ct = np.empty(vol.shape, np.int32)
    cn = np.empty((conn // 2,) + vol.shape, np.int32)
    rc = _lib.cgp_caps_from_volume(nx, ny, nz, conn, vol.ctypes.data_as(_u8p), _p32(cs), _p32(ct), _p32(cn))
    if rc != CGP_OK:
        raise CgpError(rc, "cgp_caps_from_volume failed")
    return cs, ct, cn


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def batch_shapes(count, seed=0, lo=32, hi=128):
    """Shapes (nz, ny, nx) of the batch workload: each edge length log-uniform in [lo, hi], drawn
    per axis and per grid from SplitMix64 counters of (seed, grid, axis)."""
    import math
    key = _splitmix64(_splitmix64(seed) ^ 0xB5AD4ECEDA1CE2A9)
    shapes = []
    for i in range(count):
        dims = []
        for axis in range(3):
            u = (_splitmix64(key + 3 * i + axis) >> 11) / float(1 << 53)
            dims.append(int(min(hi, max(lo, round(lo * math.exp(u * math.log(hi / lo)))))))
        nx, ny, nz = dims
        shapes.append((nz, ny, nx))
    return shapes


def max_flow_noalloc(shape, conn, cs_ptr, ct_ptr, cn_ptr, cnb_ptr=None, source_set_ptr=None, stream=None):
    """cgp_grid_max_flow_noalloc: one grid solve on device capacity arrays (raw device pointers,
    layout of set_grid) whose working memory is the library's fixed per-device scratch; S_min goes
    to source_set_ptr (device bytes) when given. Returns the flow."""
    nz, ny, nx = shape
    flow = ctypes.c_int64(0)
    rc = _lib.cgp_grid_max_flow_noalloc(nx, ny, nz, conn, _vp(cs_ptr), _vp(ct_ptr), _vp(cn_ptr), _vp(cnb_ptr),
                                        _vp(source_set_ptr), ctypes.byref(flow), _vp(stream))
    if rc != CGP_OK:
        raise CgpError(rc, _lib.cgp_last_error(None).decode(errors="replace"))
    return int(flow.value)


def read_region_graph(graph_path, markers_path=None, weight_scale=1.0, marker_capacity=10000):
    """A region adjacency graph file (and marker file) of the reference's region benchmarks, as the
    arrays of Graph.set_nweights / set_tweights: (vertex_count, edges (m, 2) as (first, second),
    weights (m,), cap_source, cap_sink); cgp_region_graph_read."""
    n = ctypes.c_int32(0)
    m = ctypes.c_int32(0)
    _global_check(_lib.cgp_region_graph_info(os.fsencode(graph_path), ctypes.byref(n), ctypes.byref(m)))
    n, m = int(n.value), int(m.value)
    rec = (Edge * max(m, 1))()
    w = np.empty(max(m, 1), np.int32)
    cs = np.empty(n, np.int32)
    ct = np.empty(n, np.int32)
    _global_check(_lib.cgp_region_graph_read(os.fsencode(graph_path),
                                             os.fsencode(markers_path) if markers_path else None,
                                             float(weight_scale), int(marker_capacity), rec, _p32(w), _p32(cs),
                                             _p32(ct)))
    arr = np.frombuffer(rec, dtype=np.uint32, count=2 * max(m, 1)).reshape(-1, 2)[:m]
    edges = np.stack([arr[:, 1], arr[:, 0]], axis=1).astype(np.int64)
    return n, edges, w[:m].copy(), cs, ct


def make_instance(nx, ny, nz, seed, instance=0, conn=6):
    """Volume and capacities of the synthetic instance (seed, instance)."""
    vol = generate_volume(nx, ny, nz, seed, instance)
    cs, ct, cn = caps_from_volume(vol, conn)
    return vol, cs, ct, cn


def make_batch(count, seed=0, conn=6, threads=None, lo=32, hi=128):
    """The batch workload: `count` grids with edge lengths log-uniform in [lo, hi] per axis and grid
    (batch_shapes), grid i being the synthetic instance (seed, i) of its shape. Returns a list of
    (cap_source, cap_sink, cap_n). The generator runs in native code without the GIL, so the grids
    are built by a pool of host threads."""
    from concurrent.futures import ThreadPoolExecutor
    shapes = batch_shapes(count, seed, lo, hi)

    def one(i):
        nz, ny, nx = shapes[i]
        _, cs, ct, cn = make_instance(nx, ny, nz, seed, i, conn)
        return cs, ct, cn

    workers = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(one, range(count)))
