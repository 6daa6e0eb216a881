"""MFI instance files in numpy -- TEST INFRASTRUCTURE ONLY (the checker of the library's MFI code).

Restates the format the reference reads and writes (benchmarks/mfi.h:4-77 for the fields and
the arc enumeration, benchmarks/mfi.cpp:17-103 for the packed header and the write order,
:105-200 for the read order): the magic "MFI", a packed little-endian header, the source and
sink capacity planes, one capacity plane per arc (arc a of node v: v -> v + offset(a)), and one
active-mask byte per node. Nodes are numbered x + y*width + z*width*height.

arc_edges() turns a file's arcs into the generic edge list of the oracle's max-flow solvers,
one directed edge per arc, with no grid-layout conversion in between.
"""
import struct

import numpy as np

HEADER = struct.Struct("<5iqi4096s3i")  # dimension, connectivity, width, height, depth, maxflow,
#                                         symmetric_arcs, metadata, active_count, two capacity types
TYPES = [np.uint8, np.uint16, np.uint32]

ARCS_2D = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0),
           (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 1, 0)]
ARCS_3D = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
           (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 1, 0),
           (0, -1, -1), (0, 1, -1), (0, -1, 1), (0, 1, 1),
           (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
           (-1, -1, -1), (1, -1, -1), (-1, 1, -1), (1, 1, -1),
           (-1, -1, 1), (1, -1, 1), (-1, 1, 1), (1, 1, 1)]


def arc_offsets(dimension, connectivity):
    return (ARCS_2D if dimension == 2 else ARCS_3D)[:connectivity]


def write(path, dimension, connectivity, source, sink, arcs, maxflow=0, symmetric=False, metadata=b"",
          terminal_type=2, neighbor_type=2, active=None):
    """source, sink: (depth, height, width) arrays; arcs: (connectivity, depth, height, width)."""
    source = np.asarray(source)
    depth, height, width = source.shape
    n = source.size
    active = np.zeros(n, np.uint8) if active is None else np.asarray(active, np.uint8).reshape(-1)
    head = HEADER.pack(dimension, connectivity, width, height, depth if dimension == 3 else 1, int(maxflow),
                       1 if symmetric else 0, metadata.ljust(4096, b"\0")[:4096], int(active.sum()),
                       terminal_type, neighbor_type)
    with open(path, "wb") as f:
        f.write(b"MFI" + head)
        f.write(np.ascontiguousarray(source, TYPES[terminal_type]).tobytes())
        f.write(np.ascontiguousarray(sink, TYPES[terminal_type]).tobytes())
        for a in range(connectivity):
            f.write(np.ascontiguousarray(arcs[a], TYPES[neighbor_type]).tobytes())
        f.write(active.tobytes())


def read(path):
    data = open(path, "rb").read()
    if data[:3] != b"MFI":
        raise ValueError("not an MFI file")
    (dimension, connectivity, width, height, depth, maxflow, symmetric, metadata, active_count, tt,
     nt) = HEADER.unpack_from(data, 3)
    if dimension == 2:
        depth = 1
    n = width * height * depth
    pos = 3 + HEADER.size
    shape = (depth, height, width)

    def take(dtype):
        nonlocal pos
        arr = np.frombuffer(data, dtype=dtype, count=n, offset=pos).reshape(shape)
        pos += n * np.dtype(dtype).itemsize
        return arr.astype(np.int64)

    source = take(TYPES[tt])
    sink = take(TYPES[tt])
    arcs = np.stack([take(TYPES[nt]) for _ in range(connectivity)]) if connectivity else None
    active = np.frombuffer(data, np.uint8, count=n, offset=pos).reshape(shape)
    return dict(dimension=dimension, connectivity=connectivity, width=width, height=height, depth=depth,
                maxflow=maxflow, symmetric_arcs=symmetric, metadata=metadata.rstrip(b"\0"),
                active_count=active_count, type_terminal_cap=tt, type_neighbor_cap=nt,
                source=source, sink=sink, arcs=arcs, active=active)


def arc_edges(mfi):
    """Directed edges (ea, eb, cap) of every arc whose head lies inside the grid, capacity > 0."""
    depth, height, width = mfi["source"].shape
    z, y, x = np.meshgrid(np.arange(depth), np.arange(height), np.arange(width), indexing="ij")
    ea, eb, cap = [], [], []
    for a, (dx, dy, dz) in enumerate(arc_offsets(mfi["dimension"], mfi["connectivity"])):
        ux, uy, uz = x + dx, y + dy, z + dz
        ok = (ux >= 0) & (ux < width) & (uy >= 0) & (uy < height) & (uz >= 0) & (uz < depth)
        c = mfi["arcs"][a]
        ok &= c > 0
        ea.append(((z * height + y) * width + x)[ok])
        eb.append(((uz * height + uy) * width + ux)[ok])
        cap.append(c[ok])
    return np.concatenate(ea), np.concatenate(eb), np.concatenate(cap)


def solve_arcs(mfi, method="dinic"):
    """Max flow of an MFI instance straight from its arcs (one directed edge each)."""
    from oracle import bk, dinic
    ea, eb, cap = arc_edges(mfi)
    n = mfi["source"].size
    zero = np.zeros_like(cap)
    fn = dinic if method == "dinic" else bk
    out = fn(n, ea, eb, cap, zero, mfi["source"].reshape(-1), mfi["sink"].reshape(-1))
    return out[0] if isinstance(out, tuple) else out
