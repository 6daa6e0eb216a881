"""The C ABI boundary on a machine without a GPU: the library loads, exports every function that
include/cugip_hip.h declares, reports the absence of a device as an error, and the C++ drop-in
header compiles. No solver computation is attempted here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cugip_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cugip_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(cgp_[a-z0-9_]+)\s*\(", text, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 24
    out = subprocess.run(["nm", "-D", "--defined-only", cugip_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(cgp_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = cugip_amd.library()
    for n in names:
        assert hasattr(lib, n)


def test_abi_version():
    assert cugip_amd.abi_version() == 2


def test_struct_layouts_match_header():
    assert ctypes.sizeof(cugip_amd.Edge) == 8
    assert cugip_amd.Edge.second.offset == 0 and cugip_amd.Edge.first.offset == 4
    # cgp_policy: 6 int32 + float + 8 reserved
    assert ctypes.sizeof(cugip_amd.Policy) == 4 * 15
    # cgp_stats: 8 + 4*4 + 3*8 + 5*8 + 8 + 4*4 + 8*4
    assert ctypes.sizeof(cugip_amd.Stats) == 8 + 16 + 24 + 40 + 8 + 16 + 32
    # cgp_grid_instance: 4 int32 + 4 pointers
    assert ctypes.sizeof(cugip_amd.GridInstance) == 48
    assert cugip_amd.GridInstance.cap_source.offset == 16 and cugip_amd.GridInstance.cap_n_back.offset == 40
    # cgp_mfi_info: 6 int32, int64, 4 int32, 4 reserved, 4096 bytes of metadata
    assert ctypes.sizeof(cugip_amd.MfiInfo) == 24 + 8 + 16 + 16 + 4096
    assert cugip_amd.MfiInfo.maxflow.offset == 24 and cugip_amd.MfiInfo.metadata.offset == 64
    # cgp_volume_weights: 5 tables of 256 int32, 4 int32, 4 reserved
    assert ctypes.sizeof(cugip_amd.VolumeWeights) == 5 * 1024 + 16 + 16


def test_default_volume_weights_are_the_generator_tables():
    w = cugip_amd.default_volume_weights()
    vol = (np.arange(6 * 7 * 8) % 256).astype(np.uint8).reshape(6, 7, 8)
    for conn in (6, 26):
        cs, ct, cn = cugip_amd.caps_from_volume(vol, conn)
        assert np.array_equal(cs, np.array(w.tlink_source)[vol])
        assert np.array_equal(ct, np.array(w.tlink_sink)[vol])
    assert (w.marker_source, w.marker_sink, w.border) == (-1, -1, 0)


def test_no_device_is_an_error_not_a_crash():
    if cugip_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(cugip_amd.CgpError) as err:
        cugip_amd.GridGraph()
    assert err.value.code == cugip_amd.CGP_ERR_DEVICE
    lib = cugip_amd.library()
    assert lib.cgp_destroy(None) == cugip_amd.CGP_OK
    assert lib.cgp_set_vertex_count(None, 4) == cugip_amd.CGP_ERR_INVALID
    assert lib.cgp_max_flow(None, None, None) == cugip_amd.CGP_ERR_INVALID


def test_cpp_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(
        "#include <cugip/graph_cut.hpp>\n"
        "int main() { cugip::GraphData<float> d; d.setVertexCount(2); d.addEdge(0, 1, 1.f, 2.f);\n"
        "  cugip::GraphCutPolicy p; p.relabelPolicy.edgeTraversalCheck.minimalResidual = 0.f;\n"
        "  static_assert(sizeof(cugip::EdgeRecord) == 8, \"layout\");\n"
        "  static_assert(sizeof(cgp_grid_instance) == 48, \"layout\");\n"
        "  static_assert(sizeof(cgp_mfi_info) == 4160 && sizeof(cgp_volume_weights) == 5152, \"layout\");\n"
        "  return d.edges[0].first == 0 && d.edges[0].second == 1 ? 0 : 1; }\n")
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", os.path.dirname(cugip_amd.LIB_PATH), "-lcugip_hip",
                    f"-Wl,-rpath,{os.path.dirname(cugip_amd.LIB_PATH)}"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_cpp_programs_built():
    for name in ("graph_cut_kats", "graph_cut_trace", "graph_cut_grid"):
        assert os.access(os.path.join(ROOT, "tests", "cpp", name), os.X_OK), name


def test_product_does_not_use_the_oracle():
    """The product package and the HIP sources never load, import or link the oracle."""
    for base in ("cugip_amd", "include"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, base)):
            for f in files:
                if f.endswith((".py", ".hip", ".h", ".hpp", ".cpp", "Makefile")) or f == "Makefile":
                    text = open(os.path.join(dirpath, f), errors="ignore").read()
                    assert "liboracle" not in text and "import oracle" not in text and "orc_" not in text, f


def test_region_graph_reader(tmp_path):
    """cgp_region_graph_read gives the edges, weights and t-links an independent Python reading of
    the same files gives (ids shifted to 0, round(sum * scale) both ways, background markers on the
    source link, a repeated edge record replacing the earlier one); malformed files are refused."""
    import numpy as np

    from tests.region_util import parse_region_graph, write_random_region_graph
    g, mk = tmp_path / "g.txt", tmp_path / "m.txt"
    write_random_region_graph(g, mk, 200, seed=1)
    n, edges, w, cs, ct = cugip_amd.read_region_graph(str(g), str(mk), 1000.0, 10000)
    rn, ea, eb, rw, rcs, rct = parse_region_graph(g, mk, 1000.0, 10000)
    assert n == rn
    got = sorted(zip(edges[:, 0].tolist(), edges[:, 1].tolist(), w.tolist()))
    ref = sorted(zip(ea.tolist(), eb.tolist(), rw.tolist()))
    assert got == ref
    assert np.array_equal(cs, rcs) and np.array_equal(ct, rct)
    bad = tmp_path / "bad.txt"
    bad.write_text("n;1;3;2.0\nn;3;3;2.0\ne;1;3;1;0.5\n")  # region 2 has no node record
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.read_region_graph(str(bad))
    assert e.value.code == cugip_amd.CGP_ERR_INVALID
    bad.write_text("n;1;3;2.0\nn;2;3;2.0\ne;1;2;1;abc\n")
    with pytest.raises(cugip_amd.CgpError):
        cugip_amd.read_region_graph(str(bad))
