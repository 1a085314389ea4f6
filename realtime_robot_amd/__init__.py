"""realtime_robot_amd: MI355X (gfx950) ICP registration and voxel fusion.

Drop-in for the registration path of ICCD/RealTime_Robot
(pcl::IterativeClosestPoint at RealTimeRobot/function.h:111-123, and the map
merge at RealTimeRobot.cpp:104-114).  The compute lives in librtr.so (HIP
kernels + C++ engine behind a C ABI, include/rtr.h); this package is its
Python host mirror.
"""
from ._lib import RtrError, Params, Result, MODE_POINT_TO_POINT, MODE_POINT_TO_PLANE  # noqa: F401
from .registration import (BatchRegistration, IterativeClosestPoint, IterativeClosestPointWithNormals,  # noqa: F401
                           VoxelMap, load_pcd, make_params, save_pcd_ascii, selfcheck_nn, selfcheck_solve, tdf,
                           tdf_points, voxel_downsample)

__all__ = ["IterativeClosestPoint", "IterativeClosestPointWithNormals", "BatchRegistration", "VoxelMap",
           "voxel_downsample", "tdf", "tdf_points", "load_pcd", "save_pcd_ascii", "make_params", "selfcheck_nn", "selfcheck_solve",
           "RtrError", "Params", "Result"]
