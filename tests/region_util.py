"""Random region adjacency graphs in the text format of the reference's region benchmarks
(benchmarks/graph_cut/graph.hpp:48-104, markers as in segmentation_regions.cpp:53-80), written for
the tests, and an independent Python reading of such files (the checker of the library's reader)."""
import numpy as np


def write_random_region_graph(path, markers_path, n, seed, extra_edges=3):
    """n regions on a ring with `extra_edges` random chords each, fractional boundary statistics,
    a few object and background markers; also comment lines, blank lines and one repeated edge."""
    rng = np.random.default_rng(seed)
    lines = ["# region graph written by tests/region_util.py", ""]
    order = rng.permutation(n) + 1
    for i in order:
        lines.append(f"n;{i};{int(rng.integers(1, 500))};{rng.random() * 1000:.6f}")
    edges = []
    for v in range(1, n + 1):
        edges.append((v, v % n + 1))
        for _ in range(extra_edges):
            u = int(rng.integers(1, n + 1))
            if u != v:
                edges.append((v, u))
    for a, b in edges:
        lines.append(f"e;{a};{b};{int(rng.integers(1, 50))};{rng.random() * 2.0:.6f}")
    # a later record of the same ordered pair replaces the earlier one
    a, b = edges[0]
    lines.append(f"e;{a};{b};7;0.125000")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    fg = rng.choice(n, max(1, n // 20), replace=False) + 1
    bg = rng.choice(np.setdiff1d(np.arange(1, n + 1), fg), max(1, n // 20), replace=False)
    with open(markers_path, "w") as f:
        f.write("# markers\n")
        f.writelines(f"f {i}\n" for i in fg)
        f.writelines(f"b\t{i}\n" for i in bg)


def parse_region_graph(path, markers_path, weight_scale, marker_capacity):
    """(n, ea, eb, w, cs, ct) read with plain Python string handling."""
    nodes = set()
    edges = {}
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        parts = line.split(";")
        if parts[0] == "n":
            nodes.add(int(parts[1]))
        else:
            edges[(int(parts[1]), int(parts[2]))] = float(parts[4])
    n = max(nodes)
    assert nodes == set(range(1, n + 1))
    ea = np.array([min(a, b) - 1 for a, b in edges], np.int32)
    eb = np.array([max(a, b) - 1 for a, b in edges], np.int32)
    w = np.array([int(np.rint(s * weight_scale)) for s in edges.values()], np.int32)
    cs = np.zeros(n, np.int32)
    ct = np.zeros(n, np.int32)
    for line in open(markers_path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        i = int(line[1:]) - 1
        if line[0] == "f":
            cs[i], ct[i] = 0, marker_capacity
        else:
            cs[i], ct[i] = marker_capacity, 0
    return n, ea, eb, w, cs, ct
