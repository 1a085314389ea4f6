"""The allocation-free TDF entry point (rtr_tdf_noalloc) against the
allocating one (rtr_tdf) through the C ABI.

Inputs and outputs are device buffers the caller owns, allocated here through
the HIP runtime the library itself links (libamdhip64, loaded once per
process); the library keeps its intermediate grid in a per-device static
array.  Two calls in a row on one device, the second on a smaller grid, each
must be bit-identical to rtr_tdf on the same voxels (itself bit-identical to
the brute-force oracle of kernel.cu:9-31, tests/test_gpu_tdf.py); a grid
above the static array's size must be refused.
"""
import ctypes as C

import numpy as np
import pytest

import realtime_robot_amd as R
from realtime_robot_amd import _lib as L

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")      # the runtime librtr.so is linked against (already loaded)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipSetDevice.argtypes = [C.c_int]
    return hip


def _ok(rc):
    assert rc == 0, f"HIP error {rc}"


def _noalloc(hip, stream, occ, dims, trunc2):
    occ = np.ascontiguousarray(occ, np.int32).reshape(-1, 3)
    n = int(np.prod(dims))
    d_occ, d_out = C.c_void_p(), C.c_void_p()
    _ok(hip.hipMalloc(C.byref(d_occ), max(occ.nbytes, 4)))
    _ok(hip.hipMalloc(C.byref(d_out), 4 * n))
    try:
        _ok(hip.hipMemcpy(d_occ, occ.ctypes.data, occ.nbytes, _H2D))
        out = np.full(n, -1.0, np.float32)
        _ok(hip.hipMemcpy(d_out, out.ctypes.data, out.nbytes, _H2D))
        d = np.ascontiguousarray(dims, np.int32)
        L.check(L.lib.rtr_tdf_noalloc(0, d_occ, len(occ), L.ptr(d), float(trunc2), d_out, stream))
        _ok(hip.hipStreamSynchronize(stream))
        _ok(hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, _D2H))
        return out.reshape(dims[2], dims[1], dims[0])
    finally:
        hip.hipFree(d_occ)
        hip.hipFree(d_out)


def test_noalloc_tdf_matches_allocating_entry_point():
    hip = _hip()
    _ok(hip.hipSetDevice(0))
    stream = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(stream)))
    try:
        rng = np.random.default_rng(11)
        cases = [((120, 100, 90), 2000, 900.0), ((30, 30, 30), 400, 900.0)]   # the second call is smaller
        for dims, n_occ, trunc2 in cases:
            occ = rng.integers(-2, np.array(dims) + 2, size=(n_occ, 3)).astype(np.int32)
            print(f"rtr_tdf_noalloc: dims {dims}, {n_occ} occupied voxels", flush=True)
            got = _noalloc(hip, stream, occ, dims, trunc2)
            ref = R.tdf(occ, dims, trunc2)
            assert np.array_equal(got, ref), f"dims {dims}: {np.count_nonzero(got != ref)} voxels differ"
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_tdf_refuses_grids_above_its_limit():
    hip = _hip()
    _ok(hip.hipSetDevice(0))
    d_occ, d_out = C.c_void_p(), C.c_void_p()
    _ok(hip.hipMalloc(C.byref(d_occ), 12))
    _ok(hip.hipMalloc(C.byref(d_out), 4))
    try:
        d = np.array([129, 128, 128], np.int32)                # 2,113,536 voxels > 2^21
        rc = L.lib.rtr_tdf_noalloc(0, d_occ, 1, L.ptr(d), 900.0, d_out, None)
        assert rc == L.ERR_CAPACITY and "limit" in L.last_error()
    finally:
        hip.hipFree(d_occ)
        hip.hipFree(d_out)
