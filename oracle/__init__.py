"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by the ``cpu_baseline`` leg of bench.py.  The product
(realtime_robot_amd) never imports this module.

Transforms are 4x4 float64 numpy arrays (row-major in Python, converted to
the column-major double[16] the C side uses).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

NOT_CONVERGED, CONV_ITERATIONS, CONV_TRANSFORM, CONV_ABS_MSE, CONV_REL_MSE, FAIL_NO_CORR = range(6)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


_lib = _load()

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class IcpParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("min_corr", C.c_int), ("mode", C.c_int), ("pad_", C.c_int),
                ("eps_T", C.c_double), ("eps_rot", C.c_double), ("eps_fit", C.c_double),
                ("r_max", C.c_double), ("anchor", C.c_double * 3)]


class IcpResult(C.Structure):
    _fields_ = [("T", C.c_double * 16), ("mse", C.c_double), ("n_corr", C.c_int64),
                ("iterations", C.c_int), ("converged", C.c_int), ("state", C.c_int), ("pad_", C.c_int)]


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_transform = _sig("orc_transform_f32", None, [_f64p, _f32p, C.c_int64, _f32p])
_r2f = _sig("orc_r2_to_float", C.c_float, [C.c_double])
_nn_brute = _sig("orc_nn_brute", None, [_f32p, C.c_int64, _f32p, C.c_int64, C.c_float, _i32p, _f32p])
_kd_build = _sig("orc_kd_build", C.c_void_p, [_f32p, C.c_int64, C.c_int])
_kd_free = _sig("orc_kd_free", None, [C.c_void_p])
_kd_nn = _sig("orc_kd_nn", None, [C.c_void_p, _f32p, C.c_int64, C.c_float, C.c_int, _i32p, _f32p])
_kd_knn = _sig("orc_kd_knn", None, [C.c_void_p, _f32p, C.c_int64, C.c_int, _i32p, _f32p])
_svd3 = _sig("orc_svd3", None, [_f64p, _f64p, _f64p, _f64p])
_umeyama = _sig("orc_umeyama", C.c_int, [_f64p, _f64p, C.c_int64, _f64p])
_eig3 = _sig("orc_eig3", None, [_f64p, _f64p, _f64p])
_chol6 = _sig("orc_cholesky6", C.c_int, [_f64p, _f64p, _f64p])
_euler = _sig("orc_euler_to_T", None, [_f64p, _f64p, _f64p])
_icp = _sig("orc_icp", C.c_int, [_f32p, C.c_int64, _f32p, C.c_void_p, C.c_int64, C.POINTER(IcpParams),
                                 _f64p, C.POINTER(IcpResult), C.c_void_p])
_fitness = _sig("orc_fitness", C.c_double, [_f32p, C.c_int64, _f32p, C.c_int64, _f64p, C.c_double])
_vox = _sig("orc_voxel_accumulate", C.c_int64, [_f32p, C.c_int64, C.c_void_p, C.c_float, _i64p, _f64p,
                                                _i64p, C.c_int64, _i64p])
_normals = _sig("orc_normals", None, [_f32p, C.c_int64, C.c_int, _f32p, _f32p, _f32p])
_downsample = _sig("orc_voxel_downsample", C.c_int64, [_f32p, C.c_int64, C.c_float, _f32p, C.c_int64, _i64p])
_tdf = _sig("orc_tdf", None, [_i32p, C.c_int64, _i32p, C.c_float, _f32p])
_fps = _sig("orc_fps", None, [_f32p, C.c_int64, C.c_int64, C.c_int, _i32p, _f64p])
_pcl_icp = _sig("orc_pcl_icp", C.c_int, [_f32p, C.c_int64, _f32p, C.c_int64, C.POINTER(IcpParams), _f64p,
                                         C.POINTER(IcpResult)])
_pcl_batch = _sig("orc_pcl_icp_batch", C.c_int, [C.c_int, C.c_void_p, _i64p, C.c_void_p, _i64p,
                                                 C.POINTER(IcpParams), C.c_void_p, C.c_void_p, C.c_int])


def _pts(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 3, a.shape
    return a


def _colmajor(T):
    return np.ascontiguousarray(np.asarray(T, dtype=np.float64).T).reshape(16)


def _rowmajor(cm):
    return np.array(cm, dtype=np.float64).reshape(4, 4).T.copy()


def r2_to_float(r2: float) -> np.float32:
    return np.float32(_r2f(float(r2)))


def transform(T, pts):
    p = _pts(pts)
    out = np.empty_like(p)
    _transform(_colmajor(T), p, len(p), out)
    return out


def nn_brute(q, t, r_max=None):
    q, t = _pts(q), _pts(t)
    r2f = np.finfo(np.float32).max if r_max is None else r2_to_float(r_max * r_max)
    idx = np.empty(len(q), np.int32)
    d2 = np.empty(len(q), np.float32)
    _nn_brute(q, len(q), t, len(t), r2f, idx, d2)
    return idx, d2


class KdTree:
    def __init__(self, t, leaf=15):
        self._t = _pts(t)
        self._h = _kd_build(self._t, len(self._t), leaf)

    def __del__(self):
        if getattr(self, "_h", None):
            _kd_free(self._h)
            self._h = None

    def nn(self, q, r_max=None, tie_lowest=True, bounded=False):
        """Exact NN (ties to the lowest index, or FLANN's first found);
        bounded: the search starts from the radius (same result)."""
        q = _pts(q)
        r2f = np.finfo(np.float32).max if r_max is None else r2_to_float(r_max * r_max)
        idx = np.empty(len(q), np.int32)
        d2 = np.empty(len(q), np.float32)
        _kd_nn(self._h, q, len(q), r2f, int(tie_lowest) | (2 if bounded else 0), idx, d2)
        return idx, d2

    def knn(self, q, k):
        q = _pts(q)
        idx = np.empty((len(q), k), np.int32)
        d2 = np.empty((len(q), k), np.float32)
        _kd_knn(self._h, q, len(q), k, idx, d2)
        return idx, d2


def svd3(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    U, S, V = np.empty(9), np.empty(3), np.empty(9)
    _svd3(A, U, S, V)
    return U.reshape(3, 3), S, V.reshape(3, 3)


def umeyama(src, tgt):
    s = np.ascontiguousarray(src, np.float64)
    t = np.ascontiguousarray(tgt, np.float64)
    T = np.empty(16)
    _umeyama(s, t, len(s), T)
    return _rowmajor(T)


def eig3(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    w, V = np.empty(3), np.empty(9)
    _eig3(A, w, V)
    return w, V.reshape(3, 3)


def pack_upper6(M):
    return np.array([M[r, c] for r in range(6) for c in range(r, 6)], dtype=np.float64)


def cholesky6(AtA, Atb):
    x = np.empty(6)
    rc = _chol6(pack_upper6(np.asarray(AtA, np.float64)), np.ascontiguousarray(Atb, np.float64), x)
    return x, rc


def euler_to_T(x, anchor=(0.0, 0.0, 0.0)):
    T = np.empty(16)
    _euler(np.ascontiguousarray(x, np.float64), np.ascontiguousarray(anchor, np.float64), T)
    return _rowmajor(T)


def params(max_iter=10, r_max=float(np.sqrt(np.finfo(np.float64).max)), eps_T=0.0, eps_fit=-np.finfo(np.float64).max,
           eps_rot=0.0, mode=0, anchor=(0.0, 0.0, 0.0), min_corr=3):
    p = IcpParams()
    p.max_iter, p.min_corr, p.mode = max_iter, min_corr, mode
    p.eps_T, p.eps_rot, p.eps_fit, p.r_max = eps_T, eps_rot, eps_fit, r_max
    for i in range(3):
        p.anchor[i] = float(anchor[i])
    return p


def _result(r: IcpResult):
    return dict(T=_rowmajor(list(r.T)), mse=r.mse, n_corr=r.n_corr, iterations=r.iterations,
                converged=bool(r.converged), state=r.state)


def icp(src, tgt, p: IcpParams, guess=None, tgt_normals=None, trace=False):
    s, t = _pts(src), _pts(tgt)
    g = _colmajor(np.eye(4) if guess is None else guess)
    r = IcpResult()
    nrm = None
    if tgt_normals is not None:
        nrm = _pts(tgt_normals)
    tr = np.zeros(16 * max(p.max_iter, 1)) if trace else None
    _icp(s, len(s), t, None if nrm is None else nrm.ctypes.data, len(t), C.byref(p), g, C.byref(r),
         None if tr is None else tr.ctypes.data)
    out = _result(r)
    if trace:
        out["trace"] = [_rowmajor(tr[16 * i:16 * i + 16]) for i in range(r.iterations)]
    return out


def fitness(src, tgt, T, max_range=np.finfo(np.float64).max):
    s, t = _pts(src), _pts(tgt)
    return _fitness(s, len(s), t, len(t), _colmajor(T), float(max_range))


def voxel_accumulate(pts, inv_h, T=None):
    p = _pts(pts)
    n = len(p)
    keys = np.empty(max(n, 1), np.int64)
    sums = np.empty(3 * max(n, 1), np.float64)
    counts = np.empty(max(n, 1), np.int64)
    dropped = np.zeros(1, np.int64)
    Tc = None if T is None else _colmajor(T)
    nv = _vox(p, n, None if Tc is None else Tc.ctypes.data, np.float32(inv_h), keys, sums, counts, n, dropped)
    return keys[:nv].copy(), sums[:3 * nv].reshape(nv, 3).copy(), counts[:nv].copy(), int(dropped[0])


def normals(pts, k=20, vp=(0.0, 0.0, 0.0)):
    p = _pts(pts)
    n = np.empty_like(p)
    c = np.empty(len(p), np.float32)
    _normals(p, len(p), k, np.asarray(vp, np.float32), n, c)
    return n, c


def voxel_downsample(pts, leaf):
    p = _pts(pts)
    out = np.empty_like(p)
    counts = np.empty(max(len(p), 1), np.int64)
    nv = _downsample(p, len(p), np.float32(leaf), out, len(p), counts)
    return out[:nv].copy(), counts[:nv].copy()


def tdf(occupied, dims, trunc2=900.0):
    """Brute-force truncated squared distance field, (dz, dy, dx) float32."""
    occ = np.ascontiguousarray(occupied, np.int32).reshape(-1, 3)
    d = np.ascontiguousarray(dims, np.int32).reshape(3)
    out = np.empty((int(d[2]), int(d[1]), int(d[0])), np.float32)
    _tdf(occ, len(occ), d, np.float32(trunc2), out)
    return out


def fps(pts, k=1024, max_sub=8192):
    """Farthest-point sample: (subsample indices (k,), FP64 centroid of the
    subsample, the subsample itself (m, 3))."""
    p = _pts(pts)
    sel = np.empty(k, np.int32)
    cen = np.empty(3)
    _fps(p, len(p), int(max_sub), int(k), sel, cen)
    m = min(len(p), max_sub)
    sub = p[(np.arange(m, dtype=np.uint64) * np.uint64(len(p)) // np.uint64(max(m, 1))).astype(np.int64)]
    return sel, cen, sub


def pcl_icp(src, tgt, p: IcpParams, guess=None):
    s, t = _pts(src), _pts(tgt)
    r = IcpResult()
    _pcl_icp(s, len(s), t, len(t), C.byref(p), _colmajor(np.eye(4) if guess is None else guess), C.byref(r))
    return _result(r)


def pcl_icp_batch(srcs, tgts, p: IcpParams, guesses=None, nthreads=1):
    srcs = [_pts(s) for s in srcs]
    tgts = [_pts(t) for t in tgts]
    n = len(srcs)
    sp = (C.c_void_p * n)(*[s.ctypes.data for s in srcs])
    tp = (C.c_void_p * n)(*[t.ctypes.data for t in tgts])
    ns = np.array([len(s) for s in srcs], np.int64)
    nt = np.array([len(t) for t in tgts], np.int64)
    g = None
    if guesses is not None:
        g = np.concatenate([_colmajor(x) for x in guesses])
    res = (IcpResult * n)()
    _pcl_batch(n, sp, ns, tp, nt, C.byref(p), None if g is None else g.ctypes.data, res, nthreads)
    return [_result(r) for r in res]
