"""Algorithmic byte counts of the correspondence search (roofline accounting).

Per ICP iteration of one pair, with a reference grid of cell edge r_max over
the target (independent of the engine's own tuned grid) and the queries
grouped by reference cell:

    B_iter = 12 * Ns  +  sum over occupied query cells c of 12 * |N27(c)|

where |N27(c)| is the number of target points in the 3 x 3 x 3 reference
cells around c: the query coordinates are read once and every candidate tile
once per query cell.  The correspondence output (8 B per query) is fused into
the moment reduction and is not counted.  The compulsory floor is
B_min = 12 * Ns + 12 * Nt.  Both follow from the inputs and the transform of
each iteration alone, so they can be recomputed on any CPU.
"""
from __future__ import annotations

import numpy as np


class TargetCounts:
    """Per-target reference grid: points per cell and their 3x3x3 sums."""

    def __init__(self, tgt, r_max: float):
        t = np.asarray(tgt, np.float64)
        t = t[np.all(np.isfinite(t), axis=1)]
        self.h = float(r_max)
        self.lo = t.min(axis=0)
        c = np.floor((t - self.lo) / self.h).astype(np.int64)
        self.dims = c.max(axis=0) + 1
        cnt = np.zeros(self.dims + 2, np.int64)  # one cell of padding on each side
        np.add.at(cnt, (c[:, 0] + 1, c[:, 1] + 1, c[:, 2] + 1), 1)
        # 3x3x3 box sum over the padded grid, then pad again so that query
        # cells one outside the target bbox still have neighbours counted
        n27 = np.zeros_like(cnt)
        P = np.pad(cnt, 1)
        for dx in range(3):
            for dy in range(3):
                for dz in range(3):
                    n27 += P[dx:dx + cnt.shape[0], dy:dy + cnt.shape[1], dz:dz + cnt.shape[2]]
        self.n27 = n27
        self.nt = len(t)

    def tile_points(self, q) -> int:
        """sum over occupied query cells of |N27(c)| for queries q (n, 3)."""
        q = np.asarray(q, np.float64)
        q = q[np.all(np.isfinite(q), axis=1)]
        c = np.floor((q - self.lo) / self.h).astype(np.int64) + 1
        inside = np.all((c >= 0) & (c < np.array(self.n27.shape)), axis=1)
        c = c[inside]
        if len(c) == 0:
            return 0
        key = (c[:, 0] * self.n27.shape[1] + c[:, 1]) * self.n27.shape[2] + c[:, 2]
        u = np.unique(key)
        zz = u % self.n27.shape[2]
        yy = (u // self.n27.shape[2]) % self.n27.shape[1]
        xx = u // (self.n27.shape[1] * self.n27.shape[2])
        return int(self.n27[xx, yy, zz].sum())


def iteration_bytes(src, counts: TargetCounts, rows34) -> tuple[int, int]:
    """(B_iter, B_min) of one iteration whose queries are rows34 (3x4) * src."""
    s = np.asarray(src, np.float64)
    R = np.asarray(rows34, np.float64)
    q = s @ R[:, :3].T + R[:, 3]
    ns = len(s)
    return 12 * ns + 12 * counts.tile_points(q), 12 * ns + 12 * counts.nt
