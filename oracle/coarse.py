"""CPU restatement of the keypoint + TDF + yaw-search coarse proposer's scoring.

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker of rtr_coarse_*.
It restates the reference's initial-alignment scoring as the library applies
it (RealTimeRobot.cpp:45-105):
  * keypoints: farthest-point sampling (oracle.fps, itself pinned bit-exact)
    standing in for PCL's Harris 3D keypoints;
  * source keypoint occupancy: 1 cm voxels within +-10 cm (key_point.h:112-161);
  * target keypoint TDF: 30^3 voxels about the keypoint, squared voxel
    distance capped at 900 (key_point.h:251-334, kernel.cu:9-31; oracle.tdf);
  * score(i, j, yaw) = mean TDF over the source voxel centres rotated about z
    (matching.h:122-222's 36 yaw steps), outside the grid 900, pairs whose
    +-10 cm occupied-voxel counts differ by more than 2x rejected
    (function.h:171-178).
Float32 expressions follow the kernel's evaluation order, so voxel indices
agree exactly; the means are summed in float64 here.
"""
from __future__ import annotations

import numpy as np

from . import fps, tdf

KEYS, OCC, TDFD, YAW, MAX_VOX, SUB = 32, 20, 30, 36, 2048, 8192
CAP = np.float32(900.0)
REJECT = np.float32(3.0e38)


def subsample(cloud, m_max=SUB):
    p = np.ascontiguousarray(cloud, np.float32).reshape(-1, 3)
    n = len(p)
    m = min(n, m_max)
    idx = (np.arange(m, dtype=np.uint64) * np.uint64(n) // np.uint64(max(m, 1))).astype(np.int64)
    return p[idx]


def keypoints(cloud):
    sel, _, sub = fps(cloud, KEYS, SUB)
    return sub[sel]


def _voxel_coords(sub, kp, half_cells, res):
    res32 = np.float32(res)
    h = np.float32(np.float32(0.5) * np.float32(half_cells * 2)) * res32
    inv = np.float32(1.0) / res32
    f = ((sub - kp.astype(np.float32)) + h) * inv
    n = 2 * half_cells
    ok = np.all((f >= 0) & (f < np.float32(n)), axis=1)
    return f[ok].astype(np.int64)


def occupancy(cloud, kp, res=0.01):
    """Ascending voxel indices x + 20 y + 400 z of the source keypoint (at most 2048)."""
    v = _voxel_coords(subsample(cloud), kp, OCC // 2, res)
    idx = np.unique(v[:, 0] + OCC * (v[:, 1] + OCC * v[:, 2]))
    return idx[:MAX_VOX]


def target_grid(cloud, kp, res=0.01):
    """(TDF (30, 30, 30) indexed [z, y, x], occupied voxels of the inner +-10 cm box)."""
    v = _voxel_coords(subsample(cloud), kp, TDFD // 2, res)
    occ = np.unique(v, axis=0) if len(v) else np.zeros((0, 3), np.int64)
    grid = tdf(occ, (TDFD, TDFD, TDFD), 900.0)
    lo, hi = (TDFD - OCC) // 2, (TDFD - OCC) // 2 + OCC
    inner = int(np.sum(np.all((occ >= lo) & (occ < hi), axis=1))) if len(occ) else 0
    return grid, inner


def yaw_table():
    k = np.arange(YAW, dtype=np.float64)
    return np.cos(2 * np.pi * k / YAW).astype(np.float32), np.sin(2 * np.pi * k / YAW).astype(np.float32)


def score(vox, grid, n_inner, res=0.01):
    """(36,) mean TDF of the source voxel centres rotated by each yaw step."""
    nv = len(vox)
    if nv == 0 or n_inner == 0 or n_inner > 2 * nv or nv > 2 * n_inner:
        return np.full(YAW, REJECT, np.float32)
    res32 = np.float32(res)
    vx = (vox % OCC).astype(np.float32)
    vy = ((vox // OCC) % OCC).astype(np.float32)
    vz = (vox // (OCC * OCC)).astype(np.float32)
    off = np.float32(np.float32(0.5) * np.float32(OCC)) * res32
    cx = (vx + np.float32(0.5)) * res32 - off
    cy = (vy + np.float32(0.5)) * res32 - off
    cz = (vz + np.float32(0.5)) * res32 - off
    h = np.float32(np.float32(0.5) * np.float32(TDFD)) * res32
    inv = np.float32(1.0) / res32
    ca, sa = yaw_table()
    out = np.empty(YAW, np.float64)
    for y in range(YAW):
        rx = ca[y] * cx - sa[y] * cy
        ry = sa[y] * cx + ca[y] * cy
        fx, fy, fz = (rx + h) * inv, (ry + h) * inv, (cz + h) * inv
        ok = (fx >= 0) & (fy >= 0) & (fz >= 0) & (fx < TDFD) & (fy < TDFD) & (fz < TDFD)
        val = np.full(nv, CAP, np.float64)
        val[ok] = grid[fz[ok].astype(np.int64), fy[ok].astype(np.int64), fx[ok].astype(np.int64)]
        out[y] = val.sum() / nv
    return out


def pair_scores(src, tgt, res=0.01):
    """(scores (32, 32, 36) [target kp i, source kp j, yaw], source keypoints, target keypoints)."""
    ks, kt = keypoints(src), keypoints(tgt)
    occ = [occupancy(src, ks[j], res) for j in range(KEYS)]
    out = np.empty((KEYS, KEYS, YAW))
    for i in range(KEYS):
        grid, inner = target_grid(tgt, kt[i], res)
        for j in range(KEYS):
            out[i, j] = score(occ[j], grid, inner, res)
    return out, ks, kt
