"""Multi-start initial poses and best-start selection for batched ICP.

Plain ICP only converges from inside its basin; the reference's answer to
that is its coarse keypoint/RANSAC stage (RealTimeRobot/function.h:35-109).
Multi-start runs ICP from several rotated initial poses about the source
centroid (centroids aligned) and keeps the start whose final alignment has
the most correspondences, breaking near-ties by the lower fitness MSE.
"""
from __future__ import annotations

import numpy as np

# rotation axes of the non-identity starts
_AXES = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 1]], np.float64)


def _rot(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    th = np.deg2rad(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * (K @ K)


def start_poses(src_centroid, tgt_centroid, n_starts: int = 8, angle_deg: float = 25.0, base=None):
    """n_starts poses: `base` (default: centroid alignment) then rotations of
    angle_deg about fixed axes through the source centroid."""
    cs = np.asarray(src_centroid, np.float64)
    ct = np.asarray(tgt_centroid, np.float64)
    if base is None:
        base = np.eye(4)
        base[:3, 3] = ct - cs
    poses = [np.array(base, np.float64)]
    for k in range(n_starts - 1):
        R = _rot(_AXES[k % len(_AXES)], angle_deg * (1 + k // len(_AXES)))
        D = np.eye(4)
        D[:3, :3] = R
        # rotate about the (moved) source centroid
        c = base[:3, :3] @ cs + base[:3, 3]
        D[:3, 3] = c - R @ c
        poses.append(D @ base)
    return poses


def select_best(results, rel_corr: float = 0.98):
    """Index of the best start: most correspondences, then lowest MSE among
    starts within rel_corr of the best count."""
    ncorr = np.array([r["n_corr"] if r["state"] != 5 else 0 for r in results], np.float64)
    if ncorr.max() <= 0:
        return 0
    ok = np.where(ncorr >= rel_corr * ncorr.max())[0]
    mse = np.array([results[i]["mse"] for i in ok])
    return int(ok[np.argmin(mse)])
