"""Keypoint + TDF + yaw-search coarse proposer (rtr_coarse_*).

The reference's own initial alignment (RealTimeRobot.cpp:45-105,
get_Distance at matching.h:122-222, Ransac at function.h:35-109), restated on
the GPU as the baseline the learned proposer is measured against: 32
farthest-point keypoints per cloud, source keypoint occupancy (1 cm voxels
within +-10 cm), target keypoint TDF (30^3 voxels, cap 900), the best of 36
yaw steps about z for every keypoint pair, and the pose of the best pairs with
the lowest truncated chamfer.  It searches yaw only, as the reference does.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

KEYS = 32
YAW_STEPS = 36
OCC = 20
MAX_VOX = 2048


class CoarseProposer:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib.rtr_coarse_create(C.byref(h), device))
        self._h = h
        self.n_pairs = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.rtr_coarse_destroy(h)
            self._h = None

    def set_clouds(self, sources, targets):
        src = [L.pts3(s) for s in sources]
        tgt = [L.pts3(t) for t in targets]
        if len(src) != len(tgt):
            raise ValueError("sources and targets must pair up")
        self._keep = (src, tgt)
        sp = (C.c_void_p * len(src))(*[s.ctypes.data for s in src])
        tp = (C.c_void_p * len(tgt))(*[t.ctypes.data for t in tgt])
        ns = (C.c_size_t * len(src))(*[len(s) for s in src])
        nt = (C.c_size_t * len(tgt))(*[len(t) for t in tgt])
        L.check(L.lib.rtr_coarse_set_clouds(self._h, len(src), sp, ns, tp, nt))
        self.n_pairs = len(src)

    def propose(self, candidates: int = 64, resolution: float = 0.01, tau: float = 0.1):
        """((n, 4, 4) row-major poses source -> target, (n,) truncated-chamfer scores)."""
        T = np.empty((self.n_pairs, 16))
        s = np.empty(self.n_pairs)
        L.check(L.lib.rtr_coarse_propose(self._h, int(candidates), float(resolution), float(tau), L.ptr(T), L.ptr(s)))
        return np.ascontiguousarray(np.transpose(T.reshape(-1, 4, 4), (0, 2, 1))), s

    def pair_scores(self, resolution: float = 0.01):
        """(scores (n, 32, 32) [pair, target kp, source kp], yaw steps (n, 32, 32),
        keypoints (2n, 32, 3): cloud 2p the source, 2p + 1 the target)."""
        n = self.n_pairs
        sc = np.empty((n, KEYS, KEYS), np.float32)
        yw = np.empty((n, KEYS, KEYS), np.int32)
        kp = np.empty((2 * n, KEYS, 3), np.float32)
        L.check(L.lib.rtr_coarse_pair_scores(self._h, float(resolution), L.ptr(sc), L.ptr(yw), L.ptr(kp)))
        return sc, yw, kp

    def occupancy(self):
        """Source keypoint occupancy of the last scoring: list per (pair, kp) of voxel indices x + 20 y + 400 z."""
        n = self.n_pairs
        vox = np.empty((n, KEYS, MAX_VOX), np.uint16)
        cnt = np.empty((n, KEYS), np.int32)
        L.check(L.lib.rtr_coarse_occupancy(self._h, L.ptr(vox), L.ptr(cnt)))
        return [[vox[p, j, :cnt[p, j]].copy() for j in range(KEYS)] for p in range(n)]

    def last_ms(self) -> float:
        return float(L.lib.rtr_coarse_last_ms(self._h))
