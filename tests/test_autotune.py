"""The autotuner's output on a machine without a GPU: a table of winners becomes a header that
compiles and whose lookup picks the bounded row for small grids and the unbounded row otherwise."""
import os
import subprocess

from cugip_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_from_table_compiles_and_matches(tmp_path):
    table = [
        dict(workload="big", conn=6, batch=False, max_voxels=None,
             best=dict(tile=[64, 4, 4], phases=8, iterations=2, ms=5.0)),
        dict(workload="small", conn=6, batch=False, max_voxels=1 << 21,
             best=dict(tile=[32, 8, 4], phases=4, iterations=1, ms=1.0)),
        dict(workload="b26", conn=26, batch=True, max_voxels=None,
             best=dict(tile=[16, 8, 4], phases=16, iterations=4, ms=9.0)),
    ]
    text = autotune.header_text(table)
    inc = tmp_path / "inc"
    inc.mkdir()
    (inc / "tuned_policy.h").write_text(text)
    src = tmp_path / "t.cpp"
    src.write_text(
        '#include "tuned_policy.h"\n#include <cstdio>\n'
        "int main() { using namespace cgp;\n"
        "  TunedPolicy a = tuned_policy(6, false, 1000), b = tuned_policy(6, false, 1LL << 24),\n"
        "              c = tuned_policy(26, true, 5), d = tuned_policy(26, false, 5);\n"
        '  std::printf("%d %d %d %d %d %d %d %d\\n", a.tile_x, a.iterations, b.tile_x, b.phases, c.tile_y,\n'
        "              c.iterations, d.tile_x, d.phases);\n  return 0; }\n")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-I", str(inc), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    # small grid -> bounded row; large -> unbounded row; 26-connected batch row; the single-grid 26C
    # rows were not tuned in this run, so the committed header's first one (grid26_small) is kept
    kept = [line for line in open(os.path.join(ROOT, "cugip_amd", "csrc", "tuned_policy.h"))
            if "// grid26_small:" in line][0]
    f = [x.strip() for x in kept.strip().lstrip("{").split("}")[0].split(",")]
    assert out == ["32", "1", "64", "8", "8", "4", f[3], f[6]]
    # rows of the tuned classes replaced, the others kept
    assert "// grid6_large" not in text and "// grid6_small" not in text and "// batch26" not in text
    assert "// batch6:" in text and "// grid26:" in text


def test_candidates_cover_compiled_shapes():
    src = open(os.path.join(ROOT, "cugip_amd", "csrc", "grid_solver.hip")).read()
    for conn, shapes in autotune.SHAPES.items():
        for tx, ty, tz in shapes:
            assert "F(%d, %d, %d, %d)" % (conn, tx, ty, tz) in src


def test_shapes_filter_restricts_candidates(monkeypatch):
    """--shapes keeps only the listed tile shapes of the workload's connectivity (the one-wave
    retune times 2 shapes × phases × iterations instead of every compiled shape)."""
    seen = []

    class FakeHandle:
        def set_policy(self, tile_x, tile_y, tile_z, discharges_per_relabel, max_tile_iterations):
            seen.append((tile_x, tile_y, tile_z))

        def close(self):
            pass

    monkeypatch.setitem(autotune.WORKLOADS, "fake", (6, False, None, lambda: []))
    monkeypatch.setattr(autotune, "_solver", lambda grids, conn, batch: (FakeHandle(), lambda: [1]))
    res = autotune.tune_workload("fake", repeats=1, log=lambda *a: None, shapes={(16, 4, 4), (32, 2, 4)})
    assert set(seen) == {(16, 4, 4), (32, 2, 4)}
    assert len(res["candidates"]) == 2 * len(autotune.PHASES[6]) * len(autotune.ITERATIONS)
    assert {tuple(c["tile"]) for c in res["candidates"]} == {(16, 4, 4), (32, 2, 4)}
