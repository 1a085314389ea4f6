"""ctypes binding of librtr.so (the C ABI declared in include/rtr.h).

The shared library is built in-tree by ``realtime_robot_amd/csrc/Makefile``
(``__graft_entry__.build()``).  Loading fails loudly when it is missing:
there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "librtr.so")

OK = 0
ERR_INVALID_ARG, ERR_EMPTY_INPUT, ERR_NO_CORRESPONDENCES, ERR_CAPACITY, ERR_DEVICE, ERR_NOT_READY, ERR_IO, \
    ERR_FORMAT = range(1, 9)
MODE_POINT_TO_POINT, MODE_POINT_TO_PLANE = 0, 1
NOT_CONVERGED, CONV_ITERATIONS, CONV_TRANSFORM, CONV_ABS_MSE, CONV_REL_MSE, FAIL_NO_CORRESPONDENCES = range(6)


class RtrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"librtr error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("min_correspondences", C.c_int), ("mode", C.c_int),
                ("anderson_memory", C.c_int), ("max_correspondence_distance", C.c_double),
                ("transformation_epsilon", C.c_double), ("rotation_epsilon", C.c_double),
                ("euclidean_fitness_epsilon", C.c_double), ("cell_size", C.c_float),
                ("cells_per_point", C.c_float), ("coarse_stride", C.c_int), ("coarse_iterations", C.c_int)]


class Result(C.Structure):
    _fields_ = [("T", C.c_double * 16), ("mse", C.c_double), ("fitness_rmse", C.c_double),
                ("n_corr", C.c_int64), ("iterations", C.c_int), ("converged", C.c_int), ("state", C.c_int),
                ("coarse_iterations", C.c_int)]


class Timing(C.Structure):
    _fields_ = [("ms_total", C.c_float), ("ms_grid", C.c_float), ("ms_icp", C.c_float), ("ms_finish", C.c_float),
                ("icp_launches", C.c_int64), ("iterations_total", C.c_int64),
                ("queries_searched", C.c_int64)]


class LearnedConfig(C.Structure):
    _fields_ = [("hypotheses", C.c_int), ("n_starts", C.c_int), ("tau", C.c_float),
                ("rmse_threshold", C.c_double), ("min_inlier_frac", C.c_double), ("judge_k", C.c_int),
                ("judge_points", C.c_int), ("max_curvature", C.c_double), ("refine_anderson", C.c_int),
                ("refine_coarse_stride", C.c_int), ("refine_coarse_iterations", C.c_int),
                ("fallback_anderson", C.c_int), ("fallback_item_pts", C.c_int),
                ("fallback_blocks_per_cu", C.c_int)]


class LearnedStats(C.Structure):
    _fields_ = [("n_pairs", C.c_int), ("n_fallback", C.c_int), ("ms_propose", C.c_float),
                ("ms_build_wait", C.c_float), ("ms_refine", C.c_float), ("ms_judge", C.c_float),
                ("ms_fallback", C.c_float), ("ms_total", C.c_float), ("ms_propose_device", C.c_float)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C realtime_robot_amd/csrc)")
    return C.CDLL(LIB_PATH)


lib = _load()

_vp = C.c_void_p
_i = C.c_int
_sz = C.c_size_t
_d = C.c_double
_f = C.c_float

_SIGS = {
    "rtr_last_error": (C.c_char_p, []),
    "rtr_version": (_i, []),
    "rtr_params_default": (_i, [C.POINTER(Params)]),
    "rtr_icp_create": (_i, [C.POINTER(_vp), _i]),
    "rtr_icp_destroy": (_i, [_vp]),
    "rtr_icp_set_input_source": (_i, [_vp, _vp, _sz, _sz]),
    "rtr_icp_set_input_target": (_i, [_vp, _vp, _sz, _sz]),
    "rtr_icp_set_target_normals": (_i, [_vp, _vp, _sz, _sz]),
    "rtr_icp_compute_target_normals": (_i, [_vp, _i, _vp]),
    "rtr_icp_get_target_normals": (_i, [_vp, _vp]),
    "rtr_icp_set_max_correspondence_distance": (_i, [_vp, _d]),
    "rtr_icp_set_maximum_iterations": (_i, [_vp, _i]),
    "rtr_icp_set_transformation_epsilon": (_i, [_vp, _d]),
    "rtr_icp_set_euclidean_fitness_epsilon": (_i, [_vp, _d]),
    "rtr_icp_set_params": (_i, [_vp, C.POINTER(Params)]),
    "rtr_icp_get_params": (_i, [_vp, C.POINTER(Params)]),
    "rtr_icp_align": (_i, [_vp, _vp, _vp, _sz, _vp, _vp]),
    "rtr_icp_get_result": (_i, [_vp, C.POINTER(Result)]),
    "rtr_icp_get_fitness_score": (_i, [_vp, _d, C.POINTER(_d)]),
    "rtr_icp_get_correspondences": (_i, [_vp, _vp, _vp, _vp]),
    "rtr_icp_get_anchor": (_i, [_vp, _vp]),
    "rtr_icp_set_profiling": (_i, [_vp, _i]),
    "rtr_icp_get_timing": (_i, [_vp, C.POINTER(Timing)]),
    "rtr_batch_create": (_i, [C.POINTER(_vp), _i]),
    "rtr_batch_destroy": (_i, [_vp]),
    "rtr_batch_set_clouds": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "rtr_batch_set_target_normals": (_i, [_vp, _vp]),
    "rtr_batch_compute_target_normals": (_i, [_vp, _i, _vp]),
    "rtr_batch_set_problems": (_i, [_vp, _i, _vp, _vp]),
    "rtr_batch_set_params": (_i, [_vp, C.POINTER(Params)]),
    "rtr_batch_set_tuning": (_i, [_vp, _f, _f, _i, _i, _i]),
    "rtr_batch_get_tuning": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rtr_batch_set_retire": (_i, [_vp, _f]),
    "rtr_batch_set_fine_items": (_i, [_vp, _i]),
    "rtr_batch_get_fine_items": (_i, [_vp, _vp]),
    "rtr_batch_get_retire": (_i, [_vp, _vp]),
    "rtr_batch_set_graph": (_i, [_vp, _i]),
    "rtr_batch_set_rebuild": (_i, [_vp, _i]),
    "rtr_batch_build": (_i, [_vp]),
    "rtr_batch_set_profiling": (_i, [_vp, _i]),
    "rtr_batch_align": (_i, [_vp, _vp, _vp]),
    "rtr_batch_get_timing": (_i, [_vp, C.POINTER(Timing)]),
    "rtr_batch_get_grid": (_i, [_vp, _i, _vp, _vp, _vp]),
    "rtr_batch_fitness": (_i, [_vp, _d, _vp]),
    "rtr_batch_get_history": (_i, [_vp, _i, _vp, _i, C.POINTER(_i)]),
    "rtr_batch_set_timeline": (_i, [_vp, _i]),
    "rtr_batch_get_timeline": (_i, [_vp, _i, _vp, _i, C.POINTER(_i), C.POINTER(_d)]),
    "rtr_batch_get_last_correspondences": (_i, [_vp, _i, _vp, _vp]),
    "rtr_batch_judge": (_i, [_vp, _i, _i, _d, _vp, _vp, _vp]),
    "rtr_map_create": (_i, [C.POINTER(_vp), _i, _f, _sz]),
    "rtr_map_destroy": (_i, [_vp]),
    "rtr_map_clear": (_i, [_vp]),
    "rtr_map_integrate": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "rtr_map_integrate_frames": (_i, [_vp, _i, _vp, _vp, _sz, _vp]),
    "rtr_map_set_profiling": (_i, [_vp, _i]),
    "rtr_map_last_ms": (_i, [_vp, C.POINTER(_f)]),
    "rtr_map_size": (_i, [_vp, C.POINTER(_sz)]),
    "rtr_map_extract": (_i, [_vp, _vp, _vp, _vp, _sz, C.POINTER(_sz), _i]),
    "rtr_map_dropped": (_i, [_vp, C.POINTER(C.c_uint64)]),
    "rtr_voxel_downsample": (_i, [_i, _vp, _sz, _sz, _f, _vp, _sz, C.POINTER(_sz)]),
    "rtr_pcd_info": (_i, [C.c_char_p, C.POINTER(_sz), C.POINTER(_i)]),
    "rtr_pcd_read": (_i, [C.c_char_p, _vp, _vp, _sz, C.POINTER(_sz)]),
    "rtr_pcd_write_ascii": (_i, [C.c_char_p, _vp, _sz, _sz]),
    "rtr_copy_probe": (_i, [_i, _sz, _i, C.POINTER(_f)]),
    "rtr_tdf": (_i, [_i, _vp, _sz, _vp, _f, _vp]),
    "rtr_tdf_points": (_i, [_i, _vp, _sz, _sz, _vp, _f, _vp, _f, _vp]),
    "rtr_tdf_noalloc": (_i, [_i, _vp, _sz, _vp, _f, _vp, _vp]),
    "rtr_coarse_create": (_i, [C.POINTER(_vp), _i]),
    "rtr_coarse_destroy": (_i, [_vp]),
    "rtr_coarse_set_clouds": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "rtr_coarse_propose": (_i, [_vp, _i, _f, _f, _vp, _vp]),
    "rtr_coarse_pair_scores": (_i, [_vp, _f, _vp, _vp, _vp]),
    "rtr_coarse_occupancy": (_i, [_vp, _vp, _vp]),
    "rtr_coarse_last_ms": (_f, [_vp]),
    "rtr_prop_create": (_i, [C.POINTER(_vp), _i, _i]),
    "rtr_prop_destroy": (_i, [_vp]),
    "rtr_prop_num_params": (_i, [C.POINTER(_sz)]),
    "rtr_prop_set_weights": (_i, [_vp, _vp]),
    "rtr_prop_get_weights": (_i, [_vp, _vp]),
    "rtr_prop_set_clouds": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "rtr_prop_fps": (_i, [_vp, _vp, _vp, _vp]),
    "rtr_prop_propose": (_i, [_vp, _vp]),
    "rtr_prop_propose_multi": (_i, [_vp, _i, _vp, _f, _vp, _vp, _vp]),
    "rtr_prop_set_fused": (_i, [_vp, _i]),
    "rtr_prop_last_ms": (_i, [_vp, C.POINTER(_f)]),
    "rtr_prop_forward_backward": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rtr_prop_train_step": (_i, [_vp, _i, _vp, _vp, _f, C.POINTER(_f)]),
    "rtr_gemm_bf16": (_i, [_i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _i]),
    "rtr_bf16_conversion_check": (_i, [_i, C.POINTER(C.c_ulonglong)]),
    "rtr_learned_config_default": (_i, [C.POINTER(LearnedConfig)]),
    "rtr_learned_create": (_i, [C.POINTER(_vp), _i, _i, C.POINTER(Params), C.POINTER(LearnedConfig)]),
    "rtr_learned_destroy": (_i, [_vp]),
    "rtr_learned_set_weights": (_i, [_vp, _vp]),
    "rtr_learned_load_weights": (_i, [_vp, C.c_char_p]),
    "rtr_learned_set_rotations": (_i, [_vp, _vp, _i, _vp, _i]),
    "rtr_learned_get_rotations": (_i, [_vp, _vp, _vp]),
    "rtr_learned_set_clouds": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rtr_learned_align": (_i, [_vp, _vp, _vp, _vp]),
    "rtr_learned_align_from": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "rtr_learned_get_stats": (_i, [_vp, C.POINTER(LearnedStats)]),
    "rtr_learned_get_proposals": (_i, [_vp, _vp]),
    "rtr_learned_get_judge": (_i, [_vp, _vp, _vp, _vp]),
    "rtr_learned_use_plain_params": (_i, [_vp, _i]),
    "rtr_learned_batches": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "rtr_selfcheck_nn": (_i, [_vp, _sz, _vp, _sz, _f, _d, _vp, _vp]),
    "rtr_selfcheck_nn_stats": (_i, [_vp, _sz, _vp, _sz, _f, _d, _vp, _vp, _vp, _vp]),
    "rtr_selfcheck_solve": (_i, [_i, _vp, _vp, _vp, _sz, _vp, _vp]),
    "rtr_selfcheck_anderson": (_i, [_i, _vp, _vp, C.c_double, _i, _vp, _vp]),
    "rtr_selfcheck_so3": (_i, [_vp, _vp, _vp]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def last_error() -> str:
    return lib.rtr_last_error().decode("utf-8", "replace")


def check(rc: int) -> None:
    if rc != OK:
        raise RtrError(rc, last_error())


def ptr(a) -> int | None:
    return None if a is None else a.ctypes.data


def pts3(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"expected an (n, 3) point array, got {a.shape}")
    return a


def colmajor(T) -> np.ndarray:
    T = np.asarray(T, dtype=np.float64)
    if T.shape != (4, 4):
        raise ValueError(f"expected a 4x4 transform, got {T.shape}")
    return np.ascontiguousarray(T.T).reshape(16)


def rowmajor(cm) -> np.ndarray:
    return np.array(list(cm), dtype=np.float64).reshape(4, 4).T.copy()


def default_params() -> Params:
    p = Params()
    check(lib.rtr_params_default(C.byref(p)))
    return p
