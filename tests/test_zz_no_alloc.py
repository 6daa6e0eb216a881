"""The no-alloc grid solve (cgp_grid_max_flow_noalloc, working memory in the library's fixed
per-device scratch) against the allocating entry point (cgp_set_grid_device + cgp_max_flow +
cgp_source_set) on the same device capacity arrays: identical flow and S_min bytes, over calls in a
row that reuse the scratch (the second smaller than the first, then a 26-connected grid), and a
clear error for a grid that does not fit."""
import ctypes

import numpy as np
import pytest

import cugip_amd
from tests.gpu_util import say
from tests.hip_rt import D2H, DeviceArray, _hip

pytestmark = pytest.mark.gpu


def device_caps(dims, conn, seed):
    nx, ny, nz = dims
    vol, cs, ct, cn = cugip_amd.make_instance(nx, ny, nz, seed, 0, conn)
    return (cs, ct, cn), [DeviceArray(a) for a in (cs, ct, cn)]


def read_bytes(dev, n):
    out = np.empty(n, np.uint8)
    assert _hip.hipMemcpy(out.ctypes.data, dev.ptr, n, D2H) == 0
    return out


def noalloc_solve(dims, conn, dev):
    nx, ny, nz = dims
    n = nx * ny * nz
    side = DeviceArray(np.full(n, 7, np.uint8))  # the caller's output buffer
    flow = cugip_amd.max_flow_noalloc((nz, ny, nx), conn, dev[0].address, dev[1].address, dev[2].address, None,
                                      side.address)
    smin = read_bytes(side, n)
    side.free()
    return flow, smin


def allocating_solve(dims, conn, dev):
    nx, ny, nz = dims
    with cugip_amd.GridGraph() as g:
        g.set_grid_device((nz, ny, nx), conn, dev[0].address, dev[1].address, dev[2].address)
        flow = g.max_flow()
        return flow, g.source_set().reshape(-1)


@pytest.mark.parametrize("order", ["large_first"])
def test_noalloc_matches_allocating_entry(order):
    cases = [((96, 80, 64), 6, 0), ((40, 36, 28), 6, 1), ((48, 40, 32), 26, 2), ((33, 17, 9), 6, 3)]
    for dims, conn, seed in cases:
        say(f"no-alloc solve {dims} conn {conn} seed {seed}")
        _, dev = device_caps(dims, conn, seed)
        flow, smin = noalloc_solve(dims, conn, dev)
        ref_flow, ref_smin = allocating_solve(dims, conn, dev)
        assert flow == ref_flow
        assert np.array_equal(smin, ref_smin), int((smin != ref_smin).sum())
        assert set(np.unique(smin)) <= {0, 1}
        for d in dev:
            d.free()


def test_noalloc_refuses_grids_beyond_its_scratch():
    dims = (128, 128, 128)  # 2.1 M voxels x 41 bytes > 64 MiB
    say(f"no-alloc too large {dims}")
    _, dev = device_caps(dims, 6, 4)
    with pytest.raises(cugip_amd.CgpError) as e:
        noalloc_solve(dims, 6, dev)
    assert e.value.code == cugip_amd.CGP_ERR_INVALID and "scratch" in str(e.value)
    for d in dev:
        d.free()
    # the scratch is usable again after the refusal
    small = (24, 20, 16)
    _, dev = device_caps(small, 6, 5)
    flow, smin = noalloc_solve(small, 6, dev)
    ref_flow, ref_smin = allocating_solve(small, 6, dev)
    assert flow == ref_flow and np.array_equal(smin, ref_smin)
    for d in dev:
        d.free()
    # invalid arguments are refused before the scratch is touched
    with pytest.raises(cugip_amd.CgpError) as e:
        cugip_amd.max_flow_noalloc((4, 4, 4), 6, None, None, None)
    assert e.value.code == cugip_amd.CGP_ERR_INVALID
    assert ctypes.sizeof(ctypes.c_int64) == 8
