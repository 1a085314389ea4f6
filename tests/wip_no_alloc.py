"""The allocation-free TDF entry point (rtr_tdf_noalloc) against the
allocating one (rtr_tdf) through the C ABI.

Inputs and outputs are device buffers the caller owns (torch tensors here);
the library keeps its intermediate grid in a per-device static array.  Two
calls in a row on one device, the second on a smaller grid, each must be
bit-identical to rtr_tdf on the same voxels (which is itself bit-identical
to the brute-force oracle of kernel.cu:9-31, tests/test_gpu_tdf.py); a grid
above the static array's size must be refused.
"""
import ctypes as C

import numpy as np
import pytest

import realtime_robot_amd as R
from realtime_robot_amd import _lib as L

pytestmark = pytest.mark.gpu


def _noalloc(torch, occ, dims, trunc2):
    dev = torch.device("cuda:0")
    occ = np.ascontiguousarray(occ, np.int32).reshape(-1, 3)
    occ_t = torch.from_numpy(occ.copy()).to(dev)
    out_t = torch.full((int(np.prod(dims)),), -1.0, dtype=torch.float32, device=dev)
    d = np.ascontiguousarray(dims, np.int32)
    stream = torch.cuda.current_stream(dev).cuda_stream
    L.check(L.lib.rtr_tdf_noalloc(0, C.c_void_p(occ_t.data_ptr()), len(occ), L.ptr(d), float(trunc2),
                                  C.c_void_p(out_t.data_ptr()), C.c_void_p(stream)))
    torch.cuda.synchronize(dev)
    return out_t.cpu().numpy().reshape(dims[2], dims[1], dims[0])


def test_noalloc_tdf_matches_allocating_entry_point():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    cases = [((120, 100, 90), 2000, 900.0), ((30, 30, 30), 400, 900.0)]   # the second call is smaller
    for dims, n_occ, trunc2 in cases:
        occ = rng.integers(-2, np.array(dims) + 2, size=(n_occ, 3)).astype(np.int32)
        print(f"rtr_tdf_noalloc: dims {dims}, {n_occ} occupied voxels", flush=True)
        got = _noalloc(torch, occ, dims, trunc2)
        ref = R.tdf(occ, dims, trunc2)
        assert np.array_equal(got, ref), f"dims {dims}: {np.count_nonzero(got != ref)} voxels differ"


def test_noalloc_tdf_refuses_grids_above_its_limit():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    occ_t = torch.zeros((1, 3), dtype=torch.int32, device=dev)
    out_t = torch.empty(1, dtype=torch.float32, device=dev)
    d = np.array([129, 128, 128], np.int32)                # 2,113,536 voxels > 2^21
    rc = L.lib.rtr_tdf_noalloc(0, C.c_void_p(occ_t.data_ptr()), 1, L.ptr(d), 900.0, C.c_void_p(out_t.data_ptr()),
                               None)
    assert rc == L.ERR_CAPACITY and "limit" in L.last_error()
