"""Known-answer graphs of the reference's own tests and benchmarks, rebuilt as edge lists.

Each builder returns (n, ea, eb, fwd, bwd, cs, ct) with fwd = capacity ea -> eb, and the
expected value the reference's code states or the closed form its construction implies:

  grid_4x4       test/graph_cut.cu:13-67: 4x4 grid, source row 0 (100), sink row 3 (100),
                 weight 5 except the four row-1 -> row-2 links (weight 1)       -> flow 4
  layered        unit_tests/graph_cut_test.cu:22-67 with MinCutLayeredGraph{,2} (:172-197):
                 100 layers x 1024 vertices x 8 neighbours, middle layer weight 1 forward
                 and 2 backward, others 20 / 40                                   -> 8192,
                 with source and sink swapped                                      -> 16384
  grid_1000      benchmarks/bench-CUGIP.cu:50-97 (run_test): 1000x1000, source row 0, sink
                 row 999, horizontal 50, vertical |y - 50| + 1                     -> 1000
  two_layer      unit_tests/graph_cut_test.cu:70-100 with bfsPropagation (:102-169): 32
                 vertices, both t-links 100000 on the first 16                   -> 16 BFS seeds
"""
import numpy as np


def _pack(n, edges, fwd, bwd, cs, ct):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    lo = np.minimum(e[:, 0], e[:, 1])
    hi = np.maximum(e[:, 0], e[:, 1])
    return (n, lo.astype(np.int32), hi.astype(np.int32), np.asarray(fwd, np.int32), np.asarray(bwd, np.int32),
            np.asarray(cs, np.int32), np.asarray(ct, np.int32))


def grid_4x4():
    v1 = [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]
    v2 = [1, 2, 3, 5, 6, 7, 9, 10, 11, 13, 14, 15, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]
    w = [1 if 16 <= i < 20 else 5 for i in range(24)]
    cs = [100] * 4 + [0] * 12
    ct = [0] * 12 + [100] * 4
    return _pack(16, list(zip(v1, v2)), w, w, cs, ct), 4


def layered(flip=False, layers=100, size=1024, neighbours=8):
    edges, fwd, bwd = [], [], []
    for layer in range(1, layers):
        w = 1 if layer == layers // 2 else 20
        for i in range(size):
            for j in range(neighbours):
                edges.append(((layer - 1) * size + i, layer * size + (i + j) % size))
                fwd.append(w)
                bwd.append(2 * w)
    n = layers * size
    ts = np.zeros(n, np.int32)
    tt = np.zeros(n, np.int32)
    ts[:size] = 100000
    tt[-size:] = 100000
    if flip:
        ts, tt = tt, ts
    # EdgeRecord keeps first = min, and the edges here already run from the lower index
    return _pack(n, edges, fwd, bwd, ts, tt), (16384 if flip else 8192)


def grid_1000(w=1000, h=1000):
    idx = np.arange(w * h).reshape(h, w)
    horiz = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    vert = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    hw = np.full(len(horiz), 50, np.int32)
    vw = (np.abs(np.repeat(np.arange(h - 1), w) - 50) + 1).astype(np.int32)
    edges = np.concatenate([horiz, vert])
    weights = np.concatenate([hw, vw])
    cs = np.zeros(w * h, np.int32)
    ct = np.zeros(w * h, np.int32)
    cs[:w] = 100000
    ct[-w:] = 100000
    return _pack(w * h, edges, weights, weights, cs, ct), 1000


def two_layer(vertex_count=32, layer_size=16, neighbours=4):
    edges = []
    for i in range(layer_size):
        for j in range(neighbours):
            edges.append((i, layer_size + (i + j) % layer_size))
    w = np.ones(len(edges), np.int32)
    t = np.zeros(vertex_count, np.int32)
    t[:layer_size] = 100000
    return _pack(vertex_count, edges, w, w, t, t), layer_size
