"""PCL-shaped host interface over librtr (Python mirror of the C++ header
``include/rtr_registration.hpp``).

The reference registers a model cloud onto a scan with
``pcl::IterativeClosestPoint<PointXYZ, PointXYZ>`` (RealTimeRobot/function.h:
111-123): ``setInputCloud`` (= ``setInputSource``), ``setInputTarget``,
``align``, ``getFinalTransformation``; and merges transformed clouds into the
scene at RealTimeRobot.cpp:104-114.  The classes below keep those names and
meanings (camelCase as in PCL, snake_case aliases for Python callers).  Every
call goes through the C ABI; nothing here computes geometry.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

DBL_MAX = float(np.finfo(np.float64).max)


class IterativeClosestPoint:
    """``pcl::IterativeClosestPoint<PointXYZ, PointXYZ>`` on one GPU."""

    _mode = L.MODE_POINT_TO_POINT

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib.rtr_icp_create(C.byref(h), device))
        self._h = h
        self._final = np.eye(4)
        self._converged = False
        self._nsrc = 0
        if self._mode != L.MODE_POINT_TO_POINT:
            p = self.get_params()
            p.mode = self._mode
            self.set_params(p)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.rtr_icp_destroy(h)
            self._h = None

    # --- inputs -----------------------------------------------------------
    def setInputSource(self, cloud):
        a = L.pts3(cloud)
        L.check(L.lib.rtr_icp_set_input_source(self._h, L.ptr(a), len(a), 12))
        self._nsrc = len(a)

    setInputCloud = setInputSource  # deprecated PCL alias used at function.h:113

    def setInputTarget(self, cloud):
        a = L.pts3(cloud)
        L.check(L.lib.rtr_icp_set_input_target(self._h, L.ptr(a), len(a), 12))
        self._ntgt = len(a)

    def setTargetNormals(self, normals):
        a = L.pts3(normals)
        L.check(L.lib.rtr_icp_set_target_normals(self._h, L.ptr(a), len(a), 12))

    def computeTargetNormals(self, k: int = 20, viewpoint=(0.0, 0.0, 0.0)):
        vp = np.asarray(viewpoint, np.float32)
        L.check(L.lib.rtr_icp_compute_target_normals(self._h, k, L.ptr(vp)))

    def getTargetNormals(self) -> np.ndarray:
        out = np.empty((self._ntgt, 3), np.float32)
        L.check(L.lib.rtr_icp_get_target_normals(self._h, L.ptr(out)))
        return out

    # --- parameters -------------------------------------------------------
    def setMaxCorrespondenceDistance(self, d: float):
        L.check(L.lib.rtr_icp_set_max_correspondence_distance(self._h, float(d)))

    def setMaximumIterations(self, n: int):
        L.check(L.lib.rtr_icp_set_maximum_iterations(self._h, int(n)))

    def setTransformationEpsilon(self, e: float):
        L.check(L.lib.rtr_icp_set_transformation_epsilon(self._h, float(e)))

    def setEuclideanFitnessEpsilon(self, e: float):
        L.check(L.lib.rtr_icp_set_euclidean_fitness_epsilon(self._h, float(e)))

    def get_params(self) -> L.Params:
        p = L.Params()
        L.check(L.lib.rtr_icp_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, p: L.Params):
        L.check(L.lib.rtr_icp_set_params(self._h, C.byref(p)))

    def set_profiling(self, on: bool):
        L.check(L.lib.rtr_icp_set_profiling(self._h, int(on)))

    # --- registration -----------------------------------------------------
    def align(self, guess=None) -> np.ndarray:
        """Runs ICP; returns transformPointCloud(source, final) like ``align(Final)``."""
        out = np.empty((self._nsrc, 3), np.float32)
        T = np.empty(16)
        conv = C.c_int()
        g = None if guess is None else L.colmajor(guess)
        L.check(L.lib.rtr_icp_align(self._h, L.ptr(g), L.ptr(out), 12, L.ptr(T), C.byref(conv)))
        self._final = L.rowmajor(T)
        self._converged = bool(conv.value)
        return out

    def getFinalTransformation(self) -> np.ndarray:
        return self._final.copy()

    def hasConverged(self) -> bool:
        return self._converged

    def getFitnessScore(self, max_range: float = DBL_MAX) -> float:
        out = C.c_double()
        L.check(L.lib.rtr_icp_get_fitness_score(self._h, float(max_range), C.byref(out)))
        return out.value

    def getResult(self) -> dict:
        r = L.Result()
        L.check(L.lib.rtr_icp_get_result(self._h, C.byref(r)))
        return result_dict(r)

    def getCorrespondences(self, T=None):
        """(index, d2) per source point under T (default: the final transform)."""
        T = self._final if T is None else T
        idx = np.empty(self._nsrc, np.int32)
        d2 = np.empty(self._nsrc, np.float32)
        L.check(L.lib.rtr_icp_get_correspondences(self._h, L.ptr(L.colmajor(T)), L.ptr(idx), L.ptr(d2)))
        return idx, d2

    def getAnchor(self) -> np.ndarray:
        a = np.empty(3, np.float32)
        L.check(L.lib.rtr_icp_get_anchor(self._h, L.ptr(a)))
        return a

    def getTiming(self) -> dict:
        t = L.Timing()
        L.check(L.lib.rtr_icp_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.Timing._fields_}

    # snake_case aliases
    set_input_source = setInputSource
    set_input_target = setInputTarget
    set_max_correspondence_distance = setMaxCorrespondenceDistance
    set_maximum_iterations = setMaximumIterations
    set_transformation_epsilon = setTransformationEpsilon
    set_euclidean_fitness_epsilon = setEuclideanFitnessEpsilon
    get_final_transformation = getFinalTransformation
    has_converged = hasConverged
    get_fitness_score = getFitnessScore


class IterativeClosestPointWithNormals(IterativeClosestPoint):
    """Point-to-plane ICP (``pcl::IterativeClosestPointWithNormals`` with the
    linear least-squares point-to-plane estimator)."""

    _mode = L.MODE_POINT_TO_PLANE


def result_dict(r: L.Result) -> dict:
    return dict(T=L.rowmajor(r.T), mse=r.mse, fitness_rmse=r.fitness_rmse, n_corr=r.n_corr,
                iterations=r.iterations, converged=bool(r.converged), state=r.state,
                coarse_iterations=r.coarse_iterations)


_RESULT_DTYPE = None


def result_dicts(res) -> list[dict]:
    """result_dict of every entry of a ctypes Result array, converted through
    one numpy view of the array (per-entry ctypes access costs ~5 us each; the
    view's dtype is derived once, np.ctypeslib.as_array costs ~45 us a call)."""
    global _RESULT_DTYPE
    n = len(res)
    if n == 0:
        return []
    if _RESULT_DTYPE is None:
        _RESULT_DTYPE = np.ctypeslib.as_array(res).dtype
    arr = np.frombuffer(res, dtype=_RESULT_DTYPE, count=n)
    Ts = arr["T"].reshape(n, 4, 4).transpose(0, 2, 1).copy()
    mse, fr, nc, it, st, ci = (arr[k].tolist() for k in ("mse", "fitness_rmse", "n_corr", "iterations", "state",
                                                           "coarse_iterations"))
    cv = (arr["converged"] != 0).tolist()
    return [{"T": T, "mse": a, "fitness_rmse": b, "n_corr": c, "iterations": d, "converged": e, "state": f,
             "coarse_iterations": g} for T, a, b, c, d, e, f, g in zip(Ts, mse, fr, nc, it, cv, st, ci)]


def make_params(max_iterations=10, max_correspondence_distance=float(np.sqrt(DBL_MAX)), transformation_epsilon=0.0,
                euclidean_fitness_epsilon=-DBL_MAX, rotation_epsilon=0.0, mode=L.MODE_POINT_TO_POINT,
                min_correspondences=3, cell_size=0.0, cells_per_point=0.0, anderson_memory=0, coarse_stride=0,
                coarse_iterations=0) -> L.Params:
    """rtr_params from keyword arguments (PCL 1.8's defaults).  anderson_memory
    > 0 turns on Anderson acceleration of the iteration, coarse_stride > 1
    coarse-to-fine iterations (neither has a PCL counterpart)."""
    p = L.default_params()
    p.max_iterations = max_iterations
    p.max_correspondence_distance = max_correspondence_distance
    p.transformation_epsilon = transformation_epsilon
    p.euclidean_fitness_epsilon = euclidean_fitness_epsilon
    p.rotation_epsilon = rotation_epsilon
    p.mode = mode
    p.min_correspondences = min_correspondences
    p.cell_size = cell_size
    p.cells_per_point = cells_per_point
    p.anderson_memory = anderson_memory
    p.coarse_stride = coarse_stride
    p.coarse_iterations = coarse_iterations
    return p


class BatchRegistration:
    """Many registrations in one persistent launch, replayed from a hipGraph."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib.rtr_batch_create(C.byref(h), device))
        self._h = h
        self._owner = None
        self._keep = []
        self.n_problems = 0
        self._max_ids = (-1, -1)

    @classmethod
    def borrow(cls, handle, owner):
        """A view of an engine another handle owns (rtr_learned_batches); it
        keeps `owner` alive and never destroys the engine."""
        self = cls.__new__(cls)
        self._h = handle
        self._owner = owner
        self._keep = []
        self.n_problems = 0
        self._max_ids = (-1, -1)
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owner", None) is None:
            L.lib.rtr_batch_destroy(h)
        self._h = None

    def set_clouds(self, sources, targets):
        src = [L.pts3(s) for s in sources]
        tgt = [L.pts3(t) for t in targets]
        sp = (C.c_void_p * len(src))(*[s.ctypes.data for s in src])
        tp = (C.c_void_p * len(tgt))(*[t.ctypes.data for t in tgt])
        ns = (C.c_size_t * len(src))(*[len(s) for s in src])
        nt = (C.c_size_t * len(tgt))(*[len(t) for t in tgt])
        L.check(L.lib.rtr_batch_set_clouds(self._h, len(src), sp, ns, len(tgt), tp, nt))
        self._ns = [len(s) for s in src]
        if self._max_ids[0] >= len(src) or self._max_ids[1] >= len(tgt):
            self.n_problems = 0             # the library dropped the problem set

    def set_target_normals(self, normals):
        nr = [L.pts3(n) for n in normals]
        npt = (C.c_void_p * len(nr))(*[n.ctypes.data for n in nr])
        L.check(L.lib.rtr_batch_set_target_normals(self._h, npt))

    def compute_target_normals(self, k=20, viewpoint=(0.0, 0.0, 0.0)):
        vp = np.asarray(viewpoint, np.float32)
        L.check(L.lib.rtr_batch_compute_target_normals(self._h, k, L.ptr(vp)))

    def set_problems(self, src_ids, tgt_ids):
        s = np.ascontiguousarray(src_ids, np.int32)
        t = np.ascontiguousarray(tgt_ids, np.int32)
        L.check(L.lib.rtr_batch_set_problems(self._h, len(s), L.ptr(s), L.ptr(t)))
        self.n_problems = len(s)
        self._max_ids = (int(s.max(initial=-1)), int(t.max(initial=-1)))
        self._prob_src = s.copy()

    def set_params(self, p: L.Params):
        L.check(L.lib.rtr_batch_set_params(self._h, C.byref(p)))

    def set_graph(self, on: bool):
        L.check(L.lib.rtr_batch_set_graph(self._h, int(on)))

    def set_rebuild(self, always: bool):
        """True (default): every align rebuilds the grids; False: they are
        built once (first align or build()) and reused, also across
        set_problems."""
        L.check(L.lib.rtr_batch_set_rebuild(self._h, int(always)))

    def build(self):
        L.check(L.lib.rtr_batch_build(self._h))

    def set_tuning(self, cell_factor=0.0, cells_per_point=0.0, item_pts=0, blocks_per_cu=-1, use_tile=-1, retire=-1.0,
                   fine_from=-1):
        L.check(L.lib.rtr_batch_set_tuning(self._h, float(cell_factor), float(cells_per_point), int(item_pts),
                                           int(blocks_per_cu), int(use_tile)))
        if retire >= 0:
            L.check(L.lib.rtr_batch_set_retire(self._h, float(retire)))
        if fine_from >= 0:
            L.check(L.lib.rtr_batch_set_fine_items(self._h, int(fine_from)))

    def tuning(self):
        """(cell_factor, cells_per_point, item_pts, blocks_per_cu, use_tile, retire, fine_from) in effect."""
        cf, cp, rt = C.c_float(), C.c_float(), C.c_float()
        ip, bp, ut, ff = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.check(L.lib.rtr_batch_get_tuning(self._h, C.byref(cf), C.byref(cp), C.byref(ip), C.byref(bp),
                                           C.byref(ut)))
        L.check(L.lib.rtr_batch_get_retire(self._h, C.byref(rt)))
        L.check(L.lib.rtr_batch_get_fine_items(self._h, C.byref(ff)))
        return (cf.value, cp.value, ip.value, bp.value, ut.value, rt.value, ff.value)

    def set_profiling(self, on: bool):
        L.check(L.lib.rtr_batch_set_profiling(self._h, int(on)))

    def align_raw(self, guesses=None, results=None):
        """align into a caller-owned ctypes Result array (no per-call allocation)."""
        g = None
        if guesses is not None:
            g = np.ascontiguousarray(np.transpose(np.asarray(guesses, np.float64), (0, 2, 1))).reshape(-1)
        L.check(L.lib.rtr_batch_align(self._h, L.ptr(g), results))
        return results

    def align(self, guesses=None):
        res = (L.Result * self.n_problems)()
        self.align_raw(guesses, res)
        return result_dicts(res)

    def timing(self) -> dict:
        t = L.Timing()
        L.check(L.lib.rtr_batch_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.Timing._fields_}

    def grid(self, tgt: int):
        cell = C.c_float()
        dims = np.empty(3, np.int32)
        anchor = np.empty(3, np.float32)
        L.check(L.lib.rtr_batch_get_grid(self._h, tgt, C.byref(cell), L.ptr(dims), L.ptr(anchor)))
        return cell.value, dims, anchor

    def history(self, prob: int, max_iterations: int = 256):
        """(iterations + 1, 3, 4) float32 rows of the transform each iteration queried with."""
        buf = np.empty((max_iterations + 1, 12), np.float32)
        n = C.c_int()
        L.check(L.lib.rtr_batch_get_history(self._h, int(prob), L.ptr(buf), len(buf), C.byref(n)))
        return buf[:n.value].reshape(-1, 3, 4).copy()

    def set_timeline(self, on: bool):
        """Record each problem's solve start/end wall-clock stamps per iteration (diagnostic)."""
        L.check(L.lib.rtr_batch_set_timeline(self._h, int(on)))

    def timeline(self, prob: int, max_iterations: int = 256):
        """(iterations, 2) solve start/end times of problem `prob` in microseconds (device wall clock)."""
        buf = np.empty(2 * max_iterations, np.uint64)
        n = C.c_int()
        tick = C.c_double()
        L.check(L.lib.rtr_batch_get_timeline(self._h, int(prob), L.ptr(buf), len(buf), C.byref(n), C.byref(tick)))
        return buf[:n.value].reshape(-1, 2).astype(np.float64) * tick.value * 1e-3

    def last_correspondences(self, prob: int):
        """(idx int32, d2 float32) per source point: the correspondences the
        last iteration of problem `prob` used (-1 / inf: none within range)."""
        if not 0 <= prob < self.n_problems:
            raise IndexError(f"problem {prob} out of range")
        n = self._ns[int(self._prob_src[prob])]
        idx = np.empty(n, np.int32)
        d2 = np.empty(n, np.float32)
        L.check(L.lib.rtr_batch_get_last_correspondences(self._h, int(prob), L.ptr(idx), L.ptr(d2)))
        return idx, d2

    def fitness(self, max_range=DBL_MAX):
        out = np.empty(self.n_problems)
        L.check(L.lib.rtr_batch_fitness(self._h, float(max_range), L.ptr(out)))
        return out

    def judge(self, k: int = 20, max_points: int = 2048, max_curvature: float = 0.02):
        """(plane_rmse, inlier_frac, planar_frac) per problem at its final
        transform: RMS distance of the inliers (nearest target point within
        the max correspondence distance) to their k-NN PCA plane, counted
        where that neighbourhood is planar (surface variation <=
        max_curvature), over a strided sample of up to max_points source
        points."""
        rmse = np.empty(self.n_problems)
        frac = np.empty(self.n_problems)
        pfrac = np.empty(self.n_problems)
        L.check(L.lib.rtr_batch_judge(self._h, int(k), int(max_points), float(max_curvature), L.ptr(rmse),
                                      L.ptr(frac), L.ptr(pfrac)))
        return rmse, frac, pfrac


class VoxelMap:
    """Voxel-hash map of FP32 centroid sums + counts (merging transformed
    clouds into the scene map)."""

    def __init__(self, voxel_size: float, capacity: int = 1 << 20, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib.rtr_map_create(C.byref(h), device, float(voxel_size), int(capacity)))
        self._h = h
        self.voxel_size = voxel_size

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.rtr_map_destroy(h)
            self._h = None

    def integrate(self, cloud, T=None):
        a = L.pts3(cloud)
        Tc = None if T is None else L.colmajor(T)
        L.check(L.lib.rtr_map_integrate(self._h, L.ptr(a), len(a), 12, L.ptr(Tc)))

    def integrate_frames(self, clouds, poses=None):
        """Every cloud (moved by its pose, (n, 4, 4) or None) in one launch."""
        arrs = [L.pts3(c) for c in clouds]
        k = len(arrs)
        ptrs = (C.c_void_p * k)(*[a.ctypes.data for a in arrs])
        ns = (C.c_size_t * k)(*[len(a) for a in arrs])
        Tc = None
        if poses is not None:
            Tc = np.ascontiguousarray(np.transpose(np.asarray(poses, np.float64), (0, 2, 1))).reshape(-1)
        L.check(L.lib.rtr_map_integrate_frames(self._h, k, ptrs, ns, 12, L.ptr(Tc)))

    def set_profiling(self, on: bool):
        L.check(L.lib.rtr_map_set_profiling(self._h, int(on)))

    def last_ms(self) -> float:
        """Device time of the last integration kernel (profiling on)."""
        ms = C.c_float()
        L.check(L.lib.rtr_map_last_ms(self._h, C.byref(ms)))
        return ms.value

    def size(self) -> int:
        n = C.c_size_t()
        L.check(L.lib.rtr_map_size(self._h, C.byref(n)))
        return n.value

    def dropped(self) -> int:
        n = C.c_uint64()
        L.check(L.lib.rtr_map_dropped(self._h, C.byref(n)))
        return n.value

    def clear(self):
        L.check(L.lib.rtr_map_clear(self._h))

    def extract(self, centroids=True):
        n = self.size()
        keys = np.empty(max(n, 1), np.int64)
        xyz = np.empty((max(n, 1), 3), np.float32)
        counts = np.empty(max(n, 1), np.uint32)
        nout = C.c_size_t()
        L.check(L.lib.rtr_map_extract(self._h, L.ptr(keys), L.ptr(xyz), L.ptr(counts), max(n, 1),
                                      C.byref(nout), int(centroids)))
        m = nout.value
        return keys[:m].copy(), xyz[:m].copy(), counts[:m].copy()


def voxel_downsample(cloud, leaf: float, device: int = 0) -> np.ndarray:
    a = L.pts3(cloud)
    out = np.empty((max(len(a), 1), 3), np.float32)
    n = C.c_size_t()
    L.check(L.lib.rtr_voxel_downsample(device, L.ptr(a), len(a), 12, float(leaf), L.ptr(out), len(out), C.byref(n)))
    return out[:n.value].copy()


def load_pcd(path: str):
    """pcl::io::loadPCDFile for x/y/z (+ normals when present)."""
    n = C.c_size_t()
    hn = C.c_int()
    L.check(L.lib.rtr_pcd_info(path.encode(), C.byref(n), C.byref(hn)))
    xyz = np.empty((max(n.value, 1), 3), np.float32)
    nrm = np.empty((max(n.value, 1), 3), np.float32) if hn.value else None
    nout = C.c_size_t()
    L.check(L.lib.rtr_pcd_read(path.encode(), L.ptr(xyz), L.ptr(nrm), len(xyz), C.byref(nout)))
    m = nout.value
    return xyz[:m].copy(), (None if nrm is None else nrm[:m].copy())


def save_pcd_ascii(path: str, cloud):
    a = L.pts3(cloud)
    L.check(L.lib.rtr_pcd_write_ascii(path.encode(), L.ptr(a), len(a), 12))


def selfcheck_nn(target, queries, cell: float, max_dist: float = float(np.sqrt(DBL_MAX))):
    """The library's grid NN search run on the host (same source as the kernel)."""
    t = L.pts3(target)
    q = L.pts3(queries)
    idx = np.empty(len(q), np.int32)
    d2 = np.empty(len(q), np.float32)
    L.check(L.lib.rtr_selfcheck_nn(L.ptr(t), len(t), L.ptr(q), len(q), float(cell), float(max_dist), L.ptr(idx),
                                   L.ptr(d2)))
    return idx, d2


def selfcheck_solve(mode, a, b, normals=None, anchor=(0.0, 0.0, 0.0)):
    """The library's moment reduction + FP64 solve run on the host."""
    a = L.pts3(a)
    b = L.pts3(b)
    n = None if normals is None else L.pts3(normals)
    an = np.asarray(anchor, np.float32)
    T = np.empty(16)
    L.check(L.lib.rtr_selfcheck_solve(mode, L.ptr(a), L.ptr(b), L.ptr(n), len(a), L.ptr(an), L.ptr(T)))
    return L.rowmajor(T)


def tdf(occupied, dims, trunc2: float = 900.0, device: int = 0) -> np.ndarray:
    """Truncated squared distance field (voxel units) of an occupancy grid:
    occupied (n, 3) int voxel coordinates, dims (dx, dy, dz); returns a
    (dz, dy, dx) float32 array (x fastest in memory), the values of the
    reference's ComputeTDF (kernel.cu:9-31) for any grid size."""
    occ = np.ascontiguousarray(occupied, np.int32).reshape(-1, 3)
    d = np.ascontiguousarray(dims, np.int32).reshape(3)
    out = np.empty((int(d[2]), int(d[1]), int(d[0])), np.float32)
    L.check(L.lib.rtr_tdf(device, L.ptr(occ), len(occ), L.ptr(d), float(trunc2), L.ptr(out)))
    return out


def tdf_points(points, origin, resolution: float, dims, trunc2: float = 900.0, device: int = 0) -> np.ndarray:
    """tdf() of the voxels floor((p - origin) / resolution) the points occupy."""
    p = L.pts3(points)
    o = np.ascontiguousarray(origin, np.float32).reshape(3)
    d = np.ascontiguousarray(dims, np.int32).reshape(3)
    out = np.empty((int(d[2]), int(d[1]), int(d[0])), np.float32)
    L.check(L.lib.rtr_tdf_points(device, L.ptr(p), len(p), 12, L.ptr(o), float(resolution), L.ptr(d),
                                 float(trunc2), L.ptr(out)))
    return out
