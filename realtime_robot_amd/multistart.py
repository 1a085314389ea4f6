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


def _rodrigues(axes, angles):
    a = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    K = np.zeros((len(a), 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -a[:, 2], a[:, 1], -a[:, 0]
    K[:, 1, 0], K[:, 2, 0], K[:, 2, 1] = a[:, 2], -a[:, 1], a[:, 0]
    s, c = np.sin(angles)[:, None, None], np.cos(angles)[:, None, None]
    return np.eye(3) + s * K + (1 - c) * (K @ K)


def so3_rotations(n: int = 8, pool: int = 4096, seed: int = 12345):
    """n rotations spread over SO(3): identity, then greedy farthest-point
    picks (geodesic angle) from a fixed pseudo-random pool of rotations
    (uniform on SO(3))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = rng.normal(size=(pool, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    chosen = [np.array([1.0, 0.0, 0.0, 0.0])]
    dist = 2 * np.arccos(np.clip(np.abs(q @ chosen[0]), 0, 1))
    for _ in range(n - 1):
        i = int(np.argmax(dist))
        chosen.append(q[i])
        dist = np.minimum(dist, 2 * np.arccos(np.clip(np.abs(q @ q[i]), 0, 1)))
    Q = np.array(chosen)
    w, x, y, z = Q[:, 0], Q[:, 1], Q[:, 2], Q[:, 3]
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    return R


def so3_start_poses(src_centroid, tgt_centroid, n_starts: int = 8):
    """Starts for arbitrary orientation: the source rotated about its
    centroid by each of so3_rotations(n_starts), centroid moved onto the
    target's."""
    return start_poses_from(so3_rotations(n_starts), src_centroid, tgt_centroid)


def start_poses_from(rotations, src_centroid, tgt_centroid):
    """T = [R | ct - R cs] for each rotation, with R cs summed left to right
    in FP64 (the order the library's fallback uses, so both give the same
    bits)."""
    cs = [float(v) for v in np.asarray(src_centroid, np.float64)]
    ct = [float(v) for v in np.asarray(tgt_centroid, np.float64)]
    out = []
    for R in np.asarray(rotations, np.float64):
        T = np.eye(4)
        T[:3, :3] = R
        for r in range(3):
            rc = float(R[r, 0]) * cs[0]
            rc = rc + float(R[r, 1]) * cs[1]
            rc = rc + float(R[r, 2]) * cs[2]
            T[r, 3] = ct[r] - rc
        out.append(T)
    return out
