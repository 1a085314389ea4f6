// Geometry core shared by the HIP kernels and the host-side self-check hooks.
//
// Everything here is __host__ __device__ so that the exact same source runs
// inside the gfx950 kernels and, through the rtr_selfcheck_* entry points, on
// the CPU.  The library is compiled with -ffp-contract=off: every float
// expression below rounds exactly as written, on both sides.
//
// Reference behaviour restated here (PCL 1.8.0 is not part of the reference
// tree; it is reached from RealTimeRobot/function.h:111-123):
//   * squared distance: FLANN L2_Simple<float>, ((dx*dx) + dy*dy) + dz*dz
//   * correspondence rejection: distance > max_dist^2 rejects (strict)
//   * transformPointCloud: row-wise ((m0*x + m1*y) + m2*z) + m3 in float
//   * DefaultConvergenceCriteria with max_iterations_similar_transforms_ = 0
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtr_types.hpp"

namespace rtr {

#define RTR_HD __host__ __device__ __forceinline__

RTR_HD float d2_f32(float qx, float qy, float qz, float tx, float ty, float tz) {
  float dx = qx - tx;
  float dy = qy - ty;
  float dz = qz - tz;
  float s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}

RTR_HD void apply_f32(const float* m, float x, float y, float z, float& ox, float& oy, float& oz) {
  float a = m[0] * x;
  a = a + m[1] * y;
  a = a + m[2] * z;
  ox = a + m[3];
  float b = m[4] * x;
  b = b + m[5] * y;
  b = b + m[6] * z;
  oy = b + m[7];
  float c = m[8] * x;
  c = c + m[9] * y;
  c = c + m[10] * z;
  oz = c + m[11];
}

// rows 0..2 of a column-major double 4x4 as float (PCL's Matrix4f)
RTR_HD void tf_rows_f32(const double* T, float* m) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) m[r * 4 + c] = (float)T[c * 4 + r];
}

// Lexicographic (d2, original index) order packed in one 64-bit key: d2 >= 0
// so its IEEE bits order like the value.
RTR_HD uint64_t nn_key(float d2, int32_t idx) {
  union {
    float f;
    uint32_t u;
  } v;
  v.f = d2;
  return ((uint64_t)v.u << 32) | (uint32_t)idx;
}

struct NNBest {
  uint64_t key;  // (d2 bits, original index); ~0 = none
  float tx, ty, tz;
  uint32_t slot;  // position in the cell-sorted target array
};

RTR_HD void scan_range(const float4* __restrict__ pts, uint32_t b, uint32_t e, float qx, float qy,
                       float qz, NNBest& best) {
  for (uint32_t j = b; j < e; ++j) {
    float4 t = pts[j];
    float d = d2_f32(qx, qy, qz, t.x, t.y, t.z);
    union {
      float f;
      int32_t i;
    } w;
    w.f = t.w;
    uint64_t k = nn_key(d, w.i);
    if (k < best.key) {
      best.key = k;
      best.tx = t.x;
      best.ty = t.y;
      best.tz = t.z;
      best.slot = j;
    }
  }
}

RTR_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Cell of a query, clamped to the grid; false for a NaN query or an empty grid.
RTR_HD bool query_cell(const GridView& g, float qx, float qy, float qz, int& cx, int& cy, int& cz) {
  float fx = (qx - g.ox) * g.inv_h;
  float fy = (qy - g.oy) * g.inv_h;
  float fz = (qz - g.oz) * g.inv_h;
  if (!(fx == fx) || !(fy == fy) || !(fz == fz)) return false;  // NaN query
  if (g.n_pts == 0) return false;
  // clamp in float before converting (far queries must not overflow int)
  const float lim = 1.0e9f;
  fx = fx < -1.f ? -1.f : (fx > lim ? lim : fx);
  fy = fy < -1.f ? -1.f : (fy > lim ? lim : fy);
  fz = fz < -1.f ? -1.f : (fz > lim ? lim : fz);
  cx = clampi((int)floorf(fx), 0, g.dx - 1);
  cy = clampi((int)floorf(fy), 0, g.dy - 1);
  cz = clampi((int)floorf(fz), 0, g.dz - 1);
  return true;
}

// After the box of cells within Chebyshev distance k of (cx, cy, cz) has been
// searched: true when every unsearched cell is provably farther than
// vis.bound() or than the acceptance bound r2f (or none is left).  The bound
// is the distance from q to the faces of the searched box, less g.margin,
// which absorbs the float rounding of the cell assignment
// floor((p - o) * inv_h); squared and scaled by (1 - 2e-6) it undercuts the
// float d2 of any unsearched point (float d2 is within 6 ulp of exact).
template <class V>
RTR_HD bool ring_done(const GridView& g, float qx, float qy, float qz, int cx, int cy, int cz, int k,
                      float r2f, const V& vis) {
  float lb = 3.0e38f;
  bool any = false;
  if (cx - k > 0) {
    lb = fminf(lb, qx - (g.ox + (float)(cx - k) * g.h));
    any = true;
  }
  if (cx + k < g.dx - 1) {
    lb = fminf(lb, (g.ox + (float)(cx + k + 1) * g.h) - qx);
    any = true;
  }
  if (cy - k > 0) {
    lb = fminf(lb, qy - (g.oy + (float)(cy - k) * g.h));
    any = true;
  }
  if (cy + k < g.dy - 1) {
    lb = fminf(lb, (g.oy + (float)(cy + k + 1) * g.h) - qy);
    any = true;
  }
  if (cz - k > 0) {
    lb = fminf(lb, qz - (g.oz + (float)(cz - k) * g.h));
    any = true;
  }
  if (cz + k < g.dz - 1) {
    lb = fminf(lb, (g.oz + (float)(cz + k + 1) * g.h) - qz);
    any = true;
  }
  if (!any) return true;  // the whole grid has been searched
  lb = lb - g.margin;
  if (lb > 0.f) {
    const float lb2 = (lb * lb) * 0.999998f;
    if (vis.bound() < lb2) return true;
    if (lb2 > r2f) return true;
  }
  return false;
}

// Distance from q to the slab [ox + a*h, ox + (b+1)*h] along one axis
// (0 inside), shortened by g.margin and clamped at 0 -- the per-axis bound of
// ring_done.
RTR_HD float slab_gap(float q, float o, float h, float margin, int a, int b) {
  const float lo = o + (float)a * h, hi = o + (float)(b + 1) * h;
  float d = q < lo ? lo - q : (q > hi ? q - hi : 0.f);
  d = d - margin;
  return d > 0.f ? d : 0.f;
}

// Lower bound on the float d2 from q to every point stored in the cells
// x0..x1 of row (y, z): the Euclidean distance to the box of those cells with
// each axis gap shortened by g.margin (float rounding of the cell assignment,
// as in ring_done), squared and scaled by (1 - 2e-6), which undercuts the
// float d2 of any point in the box.  The three squared gaps are separable, so
// walks compute each axis once and combine them with lb2_of.
RTR_HD float lb2_of(float gx2, float gy2, float gz2) { return ((gx2 + gy2) + gz2) * 0.999998f; }

RTR_HD float row_lb2(const GridView& g, float qx, float qy, float qz, int x0, int x1, int y, int z) {
  const float ax = slab_gap(qx, g.ox, g.h, g.margin, x0, x1);
  const float ay = slab_gap(qy, g.oy, g.h, g.margin, y, y);
  const float az = slab_gap(qz, g.oz, g.h, g.margin, z, z);
  return lb2_of(ax * ax, ay * ay, az * az);
}

// A row segment can be skipped when its lower bound exceeds both the best
// found so far (no point there can win, ties included: the comparison is
// strict) and the acceptance bound r2f (no point there can be accepted; the
// accepted nearest neighbour, and whether there is one, stay exact).
RTR_HD bool row_skip(float lb2, float best, float r2f) { return lb2 > (best < r2f ? best : r2f); }

// The 3x3x3 block around (cx, cy, cz) (rings 0 and 1) as nine x-rows; each
// x-row of cells is one contiguous range of the cell-sorted array.  The
// query's own row goes first, then the four side rows, then the corners, each
// skipped when its box is beyond the best found so far (or beyond r2f).
constexpr int kBlockRowOrder[9][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1},
                                      {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};  // (dy, dz)

template <class V>
RTR_HD void walk_block(const GridView& g, float qx, float qy, float qz, int cx, int cy, int cz, float r2f,
                       V& vis) {
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < g.dx - 1 ? cx + 1 : g.dx - 1;
  const float ax = slab_gap(qx, g.ox, g.h, g.margin, x0, x1);
  float gy2[3], gz2[3];
#pragma unroll
  for (int d = -1; d <= 1; ++d) {
    const float ay = slab_gap(qy, g.oy, g.h, g.margin, cy + d, cy + d);
    const float az = slab_gap(qz, g.oz, g.h, g.margin, cz + d, cz + d);
    gy2[d + 1] = ay * ay;
    gz2[d + 1] = az * az;
  }
  const float gx2 = ax * ax;
  for (int k = 0; k < 9; ++k) {
    const int dy = kBlockRowOrder[k][0], dz = kBlockRowOrder[k][1];
    const int y = cy + dy, z = cz + dz;
    if (z < 0 || z >= g.dz || y < 0 || y >= g.dy) continue;
    if (k > 0 && row_skip(lb2_of(gx2, gy2[dy + 1], gz2[dz + 1]), vis.bound(), r2f)) continue;
    const uint32_t row = ((uint32_t)z * g.dy + y) * g.dx;
    vis.scan(g.cell_start[row + x0], g.cell_start[row + x1 + 1]);
  }
}

// Cells at Chebyshev distance exactly k (k >= 2): full x-rows on the y/z
// faces of the shell, the two end cells elsewhere.
template <class V>
RTR_HD void walk_shell(const GridView& g, float qx, float qy, float qz, int cx, int cy, int cz, int k,
                       float r2f, V& vis) {
  const int fx0 = clampi(cx - k, 0, g.dx - 1), fx1 = clampi(cx + k, 0, g.dx - 1);
  const float afx = slab_gap(qx, g.ox, g.h, g.margin, fx0, fx1);
  const float ae0 = slab_gap(qx, g.ox, g.h, g.margin, cx - k, cx - k);
  const float ae1 = slab_gap(qx, g.ox, g.h, g.margin, cx + k, cx + k);
  const float gfx2 = afx * afx, ge02 = ae0 * ae0, ge12 = ae1 * ae1;
  for (int dz = -k; dz <= k; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= g.dz) continue;
    const float az = slab_gap(qz, g.oz, g.h, g.margin, z, z);
    const float gz2 = az * az;
    // the whole plane is beyond the bound: every row in it is
    if (row_skip(lb2_of(gfx2, 0.f, gz2), vis.bound(), r2f)) continue;
    const bool zface = (dz == -k || dz == k);
    for (int dy = -k; dy <= k; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= g.dy) continue;
      const float ay = slab_gap(qy, g.oy, g.h, g.margin, y, y);
      const float gy2 = ay * ay;
      const uint32_t row = ((uint32_t)z * g.dy + y) * g.dx;
      if (zface || dy == -k || dy == k) {
        if (row_skip(lb2_of(gfx2, gy2, gz2), vis.bound(), r2f)) continue;
        vis.scan(g.cell_start[row + fx0], g.cell_start[row + fx1 + 1]);
      } else {
        if (cx - k >= 0 && !row_skip(lb2_of(ge02, gy2, gz2), vis.bound(), r2f))
          vis.scan(g.cell_start[row + cx - k], g.cell_start[row + cx - k + 1]);
        if (cx + k < g.dx && !row_skip(lb2_of(ge12, gy2, gz2), vis.bound(), r2f))
          vis.scan(g.cell_start[row + cx + k], g.cell_start[row + cx + k + 1]);
      }
    }
  }
}

// Thick shell: every cell within Chebyshev distance kn of (cx, cy, cz) that is
// not within distance k (already searched).  Rows crossing the inner box
// contribute their two outer x-segments, the others their full x-range.  Row
// bounds are read four rows at a time so their loads overlap.
template <class V>
RTR_HD void walk_thick_shell(const GridView& g, int cx, int cy, int cz, int k, int kn, V& vis) {
  const int x0 = cx - kn > 0 ? cx - kn : 0;
  const int x1 = cx + kn < g.dx - 1 ? cx + kn : g.dx - 1;
  const int ix0 = cx - k > 0 ? cx - k : 0;
  const int ix1 = cx + k < g.dx - 1 ? cx + k : g.dx - 1;
  const int y0 = cy - kn > 0 ? cy - kn : 0;
  const int y1 = cy + kn < g.dy - 1 ? cy + kn : g.dy - 1;
  const int z0 = cz - kn > 0 ? cz - kn : 0;
  const int z1 = cz + kn < g.dz - 1 ? cz + kn : g.dz - 1;
  const int ny = y1 - y0 + 1;
  const int nrows = ny * (z1 - z0 + 1);
  for (int r0 = 0; r0 < nrows; r0 += 4) {
    uint32_t b[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      b[u][0] = b[u][1] = b[u][2] = b[u][3] = 0u;
      const int r = r0 + u;
      if (r < nrows) {
        const int y = y0 + r % ny, z = z0 + r / ny;
        const uint32_t row = ((uint32_t)z * g.dy + y) * g.dx;
        const int ay = y > cy ? y - cy : cy - y, az = z > cz ? z - cz : cz - z;
        if (ay > k || az > k) {
          b[u][0] = g.cell_start[row + x0];
          b[u][1] = g.cell_start[row + x1 + 1];
        } else {
          if (x0 < ix0) {
            b[u][0] = g.cell_start[row + x0];
            b[u][1] = g.cell_start[row + ix0];
          }
          if (ix1 < x1) {
            b[u][2] = g.cell_start[row + ix1 + 1];
            b[u][3] = g.cell_start[row + x1 + 1];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      vis.scan(b[u][0], b[u][1]);
      vis.scan(b[u][2], b[u][3]);
    }
  }
}

// Search beyond the box of radius k0 (searched, or known empty): while the
// stopping test fails, scan the next shell (row bounds batched).  Growing one
// ring at a time stops as early as the plain ring walk; while nothing at all
// has been found and the acceptance radius is unbounded, the radius doubles.
template <class V>
RTR_HD void walk_far(const GridView& g, float qx, float qy, float qz, int cx, int cy, int cz,
                     float r2f, V& vis, int k0) {
  const int kmax = (g.dx > g.dy ? (g.dx > g.dz ? g.dx : g.dz) : (g.dy > g.dz ? g.dy : g.dz));
  int k = k0;
  while (k < kmax && !ring_done(g, qx, qy, qz, cx, cy, cz, k, r2f, vis)) {
    if (vis.bound() > 1.0e30f && r2f > 1.0e30f) {
      const int kn = 2 * k + 1 < kmax ? 2 * k + 1 : kmax;
      walk_thick_shell(g, cx, cy, cz, k, kn, vis);
      k = kn;
    } else {
      walk_shell(g, qx, qy, qz, cx, cy, cz, k + 1, r2f, vis);
      k = k + 1;
    }
  }
}

// Shells k0 + 1, k0 + 2, ... once the box of radius k0 has been searched
// (or is known to be empty), each preceded by the stopping test.
template <class V>
RTR_HD void walk_outer(const GridView& g, float qx, float qy, float qz, int cx, int cy, int cz,
                       float r2f, V& vis, int k0 = 1) {
  const int kmax = (g.dx > g.dy ? (g.dx > g.dz ? g.dx : g.dz) : (g.dy > g.dz ? g.dy : g.dz));
  for (int k = k0; k <= kmax; ++k) {
    if (ring_done(g, qx, qy, qz, cx, cy, cz, k, r2f, vis)) return;
    walk_shell(g, qx, qy, qz, cx, cy, cz, k + 1, r2f, vis);
  }
}

// Distance-field parameters for a grid of cell edge h: a query whose cell is
// at Chebyshev distance >= R from every occupied cell is at least
// (R - 1) * h - margin from every target point; skip_ring is the smallest R
// for which that exceeds r_max (255: never), dt_cap the largest stored value
// below the cap marker dt_cap + 1.
RTR_HD void dist_caps(float r_max, float h, float margin, int* skip_ring, int* dt_cap) {
  const double rr = ((double)r_max + (double)margin) / (double)h;
  if (!(rr < 12.0)) {
    *skip_ring = 255;
    *dt_cap = 14;
  } else {
    *skip_ring = (int)floor(rr) + 2;
    *dt_cap = *skip_ring < 14 ? *skip_ring : 14;
  }
}

// Ring walk that first consults the cell's distance D (a lower bound on the
// Chebyshev distance to the nearest occupied cell): D >= skip_ring means no
// point within r_max; D >= 2 means the 3x3x3 block is empty and the walk
// starts at shell D.
template <class V>
RTR_HD void ring_walk_dt(const GridView& g, const uint8_t* dist, int skip_ring, float qx, float qy,
                         float qz, float r2f, V& vis) {
  int cx, cy, cz;
  if (!query_cell(g, qx, qy, qz, cx, cy, cz)) return;
  const int D = dist[((uint32_t)cz * g.dy + cy) * g.dx + cx];
  if (D >= skip_ring) return;
  int k0 = 1;
  if (D >= 2)
    k0 = D - 1;
  else
    walk_block(g, qx, qy, qz, cx, cy, cz, r2f, vis);
  walk_far(g, qx, qy, qz, cx, cy, cz, r2f, vis, k0);
}

// Exact search by rings of cells at growing Chebyshev distance around the
// query's (clamped) cell, stopping as soon as ring_done() holds.
template <class V>
RTR_HD void ring_walk(const GridView& g, float qx, float qy, float qz, float r2f, V& vis) {
  int cx, cy, cz;
  if (!query_cell(g, qx, qy, qz, cx, cy, cz)) return;
  walk_block(g, qx, qy, qz, cx, cy, cz, r2f, vis);
  walk_far(g, qx, qy, qz, cx, cy, cz, r2f, vis, 1);
}

RTR_HD float key_d2(uint64_t key) {
  union {
    uint32_t u;
    float f;
  } bd;
  bd.u = (uint32_t)(key >> 32);
  return bd.f;
}

struct NNVisitor {
  const float4* pts;
  float qx, qy, qz;
  NNBest best;
  RTR_HD void scan(uint32_t b, uint32_t e) { scan_range(pts, b, e, qx, qy, qz, best); }
  RTR_HD float bound() const { return best.key == ~0ull ? 3.4e38f : key_d2(best.key); }
};

// NNVisitor that also counts the ranges and candidates it was handed (the
// search work of one query; used by the host-side cost model).
struct NNCountVisitor : NNVisitor {
  uint32_t ranges = 0, cands = 0;
  RTR_HD void scan(uint32_t b, uint32_t e) {
    ++ranges;
    cands += e - b;
    NNVisitor::scan(b, e);
  }
};

// Exact nearest neighbour (lexicographic (d2, index) minimum) accepted iff
// d2 <= r2f, which equals PCL's (double)d2 <= r_max^2 when r2f is r_max^2
// rounded down to float.  key == ~0 when none.
RTR_HD NNBest nn_search(const GridView& g, float qx, float qy, float qz, float r2f) {
  NNVisitor v;
  v.pts = g.pts;
  v.qx = qx;
  v.qy = qy;
  v.qz = qz;
  v.best.key = ~0ull;
  v.best.tx = v.best.ty = v.best.tz = 0.f;
  v.best.slot = 0xffffffffu;
  ring_walk(g, qx, qy, qz, r2f, v);
  if (v.best.key != ~0ull && !(key_d2(v.best.key) <= r2f)) v.best.key = ~0ull;
  return v.best;
}

// ---------------------------------------------------------------------------
// FP64 small solves
// ---------------------------------------------------------------------------

// Cyclic Jacobi eigen-decomposition of a symmetric 4x4 (row-major, in place):
// on return the diagonal holds eigenvalues, columns of V the eigenvectors.
RTR_HD void jacobi4(double* a, double* V) {
  for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < 4; ++p) {
      diag += fabs(a[p * 5]);
      for (int q = p + 1; q < 4; ++q) off += fabs(a[p * 4 + q]);
    }
    if (off == 0.0 || off <= 1e-20 * diag) break;
#pragma unroll 1
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 4; ++q) {
        const double apq = a[p * 4 + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * 5] - a[p * 5]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {
          const double akp = a[k * 4 + p], akq = a[k * 4 + q];
          a[k * 4 + p] = c * akp - s * akq;
          a[k * 4 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {
          const double apk = a[p * 4 + k], aqk = a[q * 4 + k];
          a[p * 4 + k] = c * apk - s * aqk;
          a[q * 4 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 4; ++k) {
          const double vkp = V[k * 4 + p], vkq = V[k * 4 + q];
          V[k * 4 + p] = c * vkp - s * vkq;
          V[k * 4 + q] = s * vkp + c * vkq;
        }
      }
  }
}

// Determinant of the 3x3 [[a, b, c], [d, e, f], [g, h, i]].
RTR_HD double det3(double a, double b, double c, double d, double e, double f, double g, double h, double i) {
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// Largest eigenpair of a symmetric 4x4 (row-major) from its characteristic
// polynomial (the QCP approach of Theobald 2005): the largest root by Newton's
// method started at the Frobenius norm, an upper bound, from where the
// iteration descends monotonically; the eigenvector is the column of
// adj(N - lambda I) with the largest diagonal cofactor, since for a simple
// eigenvalue adj(N - lambda I) = prod_j (lambda_j - lambda) v v^T.  Returns
// false, leaving q untouched, when that cofactor is below 1e-6 |N|_F^3 (a
// repeated or nearly repeated largest eigenvalue, where the cofactors lose
// their relative precision): the caller then diagonalises N by Jacobi.  All
// values stay in registers.
RTR_HD bool max_eig4(const double* N, double* q) {
  const double a00 = N[0], a01 = N[1], a02 = N[2], a03 = N[3];
  const double a11 = N[5], a12 = N[6], a13 = N[7];
  const double a22 = N[10], a23 = N[11], a33 = N[15];
  const double fro2 = a00 * a00 + a11 * a11 + a22 * a22 + a33 * a33 +
                      2.0 * (a01 * a01 + a02 * a02 + a03 * a03 + a12 * a12 + a13 * a13 + a23 * a23);
  const double fro = sqrt(fro2);
  if (!(fro > 0.0) || !(fro < 1e150)) return false;
  // det(l I - N) = l^4 - e1 l^3 + e2 l^2 - e3 l + e4 (sums of principal minors)
  const double e1 = a00 + a11 + a22 + a33;
  const double e2 = (a00 * a11 - a01 * a01) + (a00 * a22 - a02 * a02) + (a00 * a33 - a03 * a03) +
                    (a11 * a22 - a12 * a12) + (a11 * a33 - a13 * a13) + (a22 * a33 - a23 * a23);
  const double e3 = det3(a00, a01, a02, a01, a11, a12, a02, a12, a22) + det3(a00, a01, a03, a01, a11, a13, a03, a13, a33) +
                    det3(a00, a02, a03, a02, a22, a23, a03, a23, a33) + det3(a11, a12, a13, a12, a22, a23, a13, a23, a33);
  const double s0 = a00 * a11 - a01 * a01, s1 = a00 * a12 - a01 * a02, s2 = a00 * a13 - a01 * a03;
  const double s3 = a01 * a12 - a11 * a02, s4 = a01 * a13 - a11 * a03, s5 = a02 * a13 - a12 * a03;
  const double c5 = a22 * a33 - a23 * a23, c4 = a12 * a33 - a13 * a23, c3 = a12 * a23 - a13 * a22;
  const double c2 = a02 * a33 - a03 * a23, c1 = a02 * a23 - a03 * a22, c0 = a02 * a13 - a03 * a12;
  const double e4 = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  double l = fro;
#pragma unroll 1
  for (int it = 0; it < 64; ++it) {
    const double f = (((l - e1) * l + e2) * l - e3) * l + e4;
    const double df = ((4.0 * l - 3.0 * e1) * l + 2.0 * e2) * l - e3;
    if (!(df > 0.0)) break;
    const double step = f / df;
    l -= step;
    if (!(fabs(step) > 1e-15 * fro)) break;
  }
  // adjugate of B = N - l I (symmetric): the ten distinct cofactors
  const double b00 = a00 - l, b11 = a11 - l, b22 = a22 - l, b33 = a33 - l;
  const double C00 = det3(b11, a12, a13, a12, b22, a23, a13, a23, b33);
  const double C11 = det3(b00, a02, a03, a02, b22, a23, a03, a23, b33);
  const double C22 = det3(b00, a01, a03, a01, b11, a13, a03, a13, b33);
  const double C33 = det3(b00, a01, a02, a01, b11, a12, a02, a12, b22);
  const double C01 = -det3(a01, a12, a13, a02, b22, a23, a03, a23, b33);
  const double C02 = det3(a01, b11, a13, a02, a12, a23, a03, a13, b33);
  const double C03 = -det3(a01, b11, a12, a02, a12, b22, a03, a13, a23);
  const double C12 = -det3(b00, a01, a03, a02, a12, a23, a03, a13, b33);
  const double C13 = det3(b00, a01, a02, a02, a12, b22, a03, a13, a23);
  const double C23 = -det3(b00, a01, a02, a01, b11, a12, a03, a13, a23);
  double v0 = C00, v1 = C01, v2 = C02, v3 = C03, best = fabs(C00);
  if (fabs(C11) > best) {
    v0 = C01; v1 = C11; v2 = C12; v3 = C13; best = fabs(C11);
  }
  if (fabs(C22) > best) {
    v0 = C02; v1 = C12; v2 = C22; v3 = C23; best = fabs(C22);
  }
  if (fabs(C33) > best) {
    v0 = C03; v1 = C13; v2 = C23; v3 = C33; best = fabs(C33);
  }
  if (!(best > 1e-6 * fro2 * fro)) return false;
  const double nv = sqrt(v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3);
  q[0] = v0 / nv;
  q[1] = v1 / nv;
  q[2] = v2 / nv;
  q[3] = v3 / nv;
  return true;
}

// Rigid fit from first and second moments (Horn 1987, closed form via unit
// quaternions): n pairs (a_i -> b_i) with sums Sa, Sb and Sab[r*3+c] =
// sum a_r b_c, all relative to a common origin.  Produces R, t with
// b ~ R a + t in that origin's frame.  The quaternion of the largest
// eigenvalue of Horn's symmetric 4x4 N is a proper rotation, which is the
// reflection-guarded Kabsch/Umeyama solution.  ws: 32 doubles of scratch
// (N and its eigenvectors; LDS inside the kernels, keeping VGPRs free).
RTR_HD void rigid_horn(double n, const double* Sa, const double* Sb, const double* Sab, double* R,
                       double* t, double* ws) {
  double ma[3], mb[3], M[9];
  for (int i = 0; i < 3; ++i) {
    ma[i] = Sa[i] / n;
    mb[i] = Sb[i] / n;
  }
  // centred cross-covariance M[r][c] = sum (a_r - ma_r)(b_c - mb_c) / n
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M[r * 3 + c] = Sab[r * 3 + c] / n - ma[r] * mb[c];
  const double Sxx = M[0], Sxy = M[1], Sxz = M[2];
  const double Syx = M[3], Syy = M[4], Syz = M[5];
  const double Szx = M[6], Szy = M[7], Szz = M[8];
  double* N = ws;
  double* V = ws + 16;
  N[0] = Sxx + Syy + Szz;
  N[1] = N[4] = Syz - Szy;
  N[2] = N[8] = Szx - Sxz;
  N[3] = N[12] = Sxy - Syx;
  N[5] = Sxx - Syy - Szz;
  N[6] = N[9] = Sxy + Syx;
  N[7] = N[13] = Szx + Sxz;
  N[10] = -Sxx + Syy - Szz;
  N[11] = N[14] = Syz + Szy;
  N[15] = -Sxx - Syy + Szz;
  double qv[4];
  if (!max_eig4(N, qv)) {
    jacobi4(N, V);
    int best = 0;
    for (int i = 1; i < 4; ++i)
      if (N[i * 5] > N[best * 5]) best = i;
    for (int i = 0; i < 4; ++i) qv[i] = V[i * 4 + best];
  }
  double q0 = qv[0], qx = qv[1], qy = qv[2], qz = qv[3];
  const double nq = sqrt(q0 * q0 + qx * qx + qy * qy + qz * qz);
  q0 /= nq;
  qx /= nq;
  qy /= nq;
  qz /= nq;
  R[0] = q0 * q0 + qx * qx - qy * qy - qz * qz;
  R[1] = 2.0 * (qx * qy - q0 * qz);
  R[2] = 2.0 * (qx * qz + q0 * qy);
  R[3] = 2.0 * (qy * qx + q0 * qz);
  R[4] = q0 * q0 - qx * qx + qy * qy - qz * qz;
  R[5] = 2.0 * (qy * qz - q0 * qx);
  R[6] = 2.0 * (qz * qx - q0 * qy);
  R[7] = 2.0 * (qz * qy + q0 * qx);
  R[8] = q0 * q0 - qx * qx - qy * qy + qz * qz;
  for (int r = 0; r < 3; ++r) t[r] = mb[r] - (R[r * 3] * ma[0] + R[r * 3 + 1] * ma[1] + R[r * 3 + 2] * ma[2]);
}

RTR_HD int up6(int r, int c) { return r * 6 - r * (r - 1) / 2 + (c - r); }

// 6x6 SPD solve by Cholesky.  A pivot below 1e-15 of the largest diagonal
// marks a direction the correspondences do not constrain; its unknown is 0.
// ws: 36 doubles of scratch for L.
RTR_HD int cholesky6(const double* AtA, const double* Atb, double* x, double* ws) {
  double* L = ws;
  double y[6];
  int active[6];
  double dmax = 0;
  for (int i = 0; i < 36; ++i) L[i] = 0.0;
  for (int i = 0; i < 6; ++i) dmax = fmax(dmax, AtA[up6(i, i)]);
  int ok = 1;
#pragma unroll 1
  for (int j = 0; j < 6; ++j) {
    double s = AtA[up6(j, j)];
    for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
    if (!(s > 1e-15 * dmax)) {
      active[j] = 0;
      ok = 0;
      L[j * 6 + j] = 1.0;
      continue;
    }
    active[j] = 1;
    const double ljj = sqrt(s);
    L[j * 6 + j] = ljj;
    for (int i = j + 1; i < 6; ++i) {
      double t = AtA[up6(j, i)];
      for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
      L[i * 6 + j] = t / ljj;
    }
  }
  for (int i = 0; i < 6; ++i) {
    double s = Atb[i];
    for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * y[k];
    y[i] = active[i] ? s / L[i * 6 + i] : 0.0;
  }
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
    x[i] = active[i] ? s / L[i * 6 + i] : 0.0;
  }
  return ok;
}

// Low (2004) small-angle parameterisation: R = Rz(g) Ry(b) Rx(a), t in the
// anchor frame; returns the world-frame 4x4 (column-major).
RTR_HD void euler_anchor_T(const double* x, const double* anchor, double* T) {
  const double ca = cos(x[0]), sa = sin(x[0]);
  const double cb = cos(x[1]), sb = sin(x[1]);
  const double cg = cos(x[2]), sg = sin(x[2]);
  const double R[9] = {cg * cb, -sg * ca + cg * sb * sa, sg * sa + cg * sb * ca,
                       sg * cb, cg * ca + sg * sb * sa,  -cg * sa + sg * sb * ca,
                       -sb,     cb * sa,                 cb * ca};
  for (int i = 0; i < 16; ++i) T[i] = 0.0;
  T[15] = 1.0;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[c * 4 + r] = R[r * 3 + c];
    const double Ra = R[r * 3] * anchor[0] + R[r * 3 + 1] * anchor[1] + R[r * 3 + 2] * anchor[2];
    T[12 + r] = x[3 + r] + anchor[r] - Ra;
  }
}

// C = A B for a C that aliases neither input (one column at a time).
RTR_HD void mat4_mul_to(const double* A, const double* B, double* C) {
#pragma unroll 1
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A[k * 4 + r] * B[c * 4 + k];
      C[c * 4 + r] = s;
    }
}

RTR_HD void mat4_mul(const double* A, const double* B, double* C) {
  double Rm[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A[k * 4 + r] * B[c * 4 + k];
      Rm[c * 4 + r] = s;
    }
  for (int i = 0; i < 16; ++i) C[i] = Rm[i];
}

// Incremental transform from reduced FP64 sums (layout in rtr_types.hpp).
// Returns the number of correspondences; Tinc is column-major.  ws: 64
// doubles of scratch.
RTR_HD double solve_increment(int mode, const double* v, const float* anchor_f, double* Tinc,
                              double* mse, double* ws) {
  const double n = v[SUM_N];
  *mse = n > 0 ? v[SUM_D2] / n : 0.0;
  const double anchor[3] = {(double)anchor_f[0], (double)anchor_f[1], (double)anchor_f[2]};
  for (int i = 0; i < 16; ++i) Tinc[i] = 0.0;
  Tinc[0] = Tinc[5] = Tinc[10] = Tinc[15] = 1.0;
  if (n <= 0) return n;
  if (mode == RTR_MODE_POINT_TO_POINT) {
    double R[9], t[3];
    rigid_horn(n, v + SUM_A, v + SUM_B, v + SUM_AB, R, t, ws);
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) Tinc[c * 4 + r] = R[r * 3 + c];
      const double Ra = R[r * 3] * anchor[0] + R[r * 3 + 1] * anchor[1] + R[r * 3 + 2] * anchor[2];
      Tinc[12 + r] = t[r] + anchor[r] - Ra;
    }
  } else {
    double x[6];
    cholesky6(v + SUM_ATA, v + SUM_ATB, x, ws);
    euler_anchor_T(x, anchor, Tinc);
  }
  return n;
}

// PCL 1.8 DefaultConvergenceCriteria (max_iterations_similar_transforms_ = 0,
// mse_threshold_absolute_ = 1e-12), thresholds as IterativeClosestPoint sets
// them: rotation 1 - transformation_epsilon (or the rotation epsilon when
// set), translation transformation_epsilon, relative MSE the euclidean fitness
// epsilon.  prev_mse starts at DBL_MAX.
RTR_HD int convergence(const IcpParams& p, int iterations, const double* Tinc, double mse,
                       double* prev_mse) {
  if (iterations >= p.max_iter) return RTR_CONV_ITERATIONS;
  const double rot_thr = p.eps_rot > 0 ? p.eps_rot : 1.0 - p.eps_T;
  const double cos_angle = 0.5 * (Tinc[0] + Tinc[5] + Tinc[10] - 1.0);
  const double tsq = Tinc[12] * Tinc[12] + Tinc[13] * Tinc[13] + Tinc[14] * Tinc[14];
  if (cos_angle >= rot_thr && tsq <= p.eps_T) return RTR_CONV_TRANSFORM;
  // With Anderson acceleration the iterates are extrapolated, and two of them
  // can share an MSE far from the fixed point: only the increment decides.
  if (p.anderson > 0) {
    *prev_mse = mse;
    return RTR_NOT_CONVERGED;
  }
  const double dm = fabs(mse - *prev_mse);
  if (dm < 1e-12) return RTR_CONV_ABS_MSE;
  if (dm / *prev_mse < p.eps_fit) return RTR_CONV_REL_MSE;
  *prev_mse = mse;
  return RTR_NOT_CONVERGED;
}

// ---------------------------------------------------------------------------
// Anderson acceleration of the ICP iteration (opt-in; with memory 0 the
// iteration is PCL's plain one).  ICP is a fixed-point map T -> G(T) = dT(T) T
// that converges linearly, slowly along weakly constrained directions.
// Type-II Anderson mixing (Walker & Ni, SIAM J. Numer. Anal. 49(4), 2011)
// extrapolates from the last m + 1 pairs (x_j, g_j = G(x_j)):
//   gamma = argmin || f_k - dF gamma ||,  f_j = g_j - x_j,  dF_j = f_(j+1) - f_j
//   x_(k+1) = g_k - dG gamma,             dG_j = g_(j+1) - g_j
// on a 6-vector x = (v, t) of a pose: v charts the rotation relative to a base
// rotation R_b (R R_b^T, below), t is the translation.  Safeguard as in AA-ICP (Pavlov,
// Ovchinnikov, Shipulin, Parfenov, ICRA 2018): when the objective at an
// iterate is not below the last accepted one's, the next iterate is the plain
// step from that accepted iterate and the history restarts; a mixed step
// under a tenth of the plain one (stagnation) takes the plain step and
// restarts; one beyond ten plain steps is shortened to ten.  Convergence is judged on the plain ICP increment only (the
// transformation criterion and the iteration limit; the MSE criteria are
// off, see convergence()), so the final pose is the plain step from the last
// iterate.
// ---------------------------------------------------------------------------
// kAndersonMax / kAndersonStride are in rtr_types.hpp.  State layout:
// [0..8] R_b (row-major), [9] pairs held, [10] best MSE, [12 + 12 j ..] x_j (6)
// then g_j (6), oldest first.  The update runs on one lane; every array it
// touches is passed in (LDS on the device), and its loops are kept rolled, so
// it adds neither private memory nor register pressure to the kernel.
constexpr int kAndersonWs = 168;  // doubles of workspace of anderson_pose_step (the mix uses 0..71 and 120..167)

// Rotations are charted by the vector part v of their unit quaternion with
// positive scalar part (v = sin(th/2) * axis): matrix -> v needs only a square
// root and v -> matrix is polynomial, so the update calls no transcendental
// function.  so3_to_qv fails beyond about 174 degrees (scalar part below
// 0.05), where the chart is ill-conditioned; the step is then not accelerated.
RTR_HD bool so3_to_qv(const double* R, double* v) {
  const double s = 1.0 + ((R[0] + R[4]) + R[8]);  // 4 w^2
  if (!(s > 0.01)) return false;
  const double w = 0.5 * sqrt(s);
  const double f = 0.25 / w;
  v[0] = (R[7] - R[5]) * f;
  v[1] = (R[2] - R[6]) * f;
  v[2] = (R[3] - R[1]) * f;
  return true;
}

// Row-major rotation of the unit quaternion (sqrt(1 - |v|^2), v); false when
// |v| > 1.
RTR_HD bool qv_to_so3(const double* v, double* R) {
  const double x = v[0], y = v[1], z = v[2];
  const double n2 = (x * x + y * y) + z * z;
  if (!(n2 <= 1.0)) return false;
  const double w = sqrt(1.0 - n2);
  R[0] = 1.0 - 2.0 * (y * y + z * z);
  R[1] = 2.0 * (x * y - w * z);
  R[2] = 2.0 * (x * z + w * y);
  R[3] = 2.0 * (x * y + w * z);
  R[4] = 1.0 - 2.0 * (x * x + z * z);
  R[5] = 2.0 * (y * z - w * x);
  R[6] = 2.0 * (x * z - w * y);
  R[7] = 2.0 * (y * z + w * x);
  R[8] = 1.0 - 2.0 * (x * x + y * y);
  return true;
}

// Mixing step over k >= 2 pairs H[12 j .. 12 j + 11] = (x_j, g_j), oldest
// first.  ws: 72 doubles.  False when the least-squares system is singular.
RTR_HD bool anderson_mix(const double* H, int k, double* x_new, double* ws) {
  const int n = k - 1;
  double* A = ws;        // n x n
  double* b = ws + 64;   // n
  double* D = ws + 120;  // n x 6: dF_j = f_(j+1) - f_j, f_j = g_j - x_j (each difference taken once)
#pragma unroll 1
  for (int j = 0; j < n; ++j)
#pragma unroll
    for (int i = 0; i < 6; ++i)
      D[6 * j + i] = (H[12 * (j + 1) + 6 + i] - H[12 * (j + 1) + i]) - (H[12 * j + 6 + i] - H[12 * j + i]);
#pragma unroll 1
  for (int r = 0; r < n; ++r) {
#pragma unroll 1
    for (int c = r; c < n; ++c) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) s += D[6 * r + i] * D[6 * c + i];
      A[r * n + c] = s;
      A[c * n + r] = s;
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) s += D[6 * r + i] * (H[12 * (k - 1) + 6 + i] - H[12 * (k - 1) + i]);
    b[r] = s;
  }
  // Tikhonov term relative to the diagonal keeps near-collinear histories solvable
  double tr = 0.0;
#pragma unroll 1
  for (int r = 0; r < n; ++r) tr += A[r * n + r];
  if (!(tr > 0.0)) return false;
#pragma unroll 1
  for (int r = 0; r < n; ++r) A[r * n + r] += 1e-10 * tr / n;
  // Cholesky in place (lower), then the two triangular solves into b
#pragma unroll 1
  for (int j = 0; j < n; ++j) {
    double s = A[j * n + j];
#pragma unroll 1
    for (int q = 0; q < j; ++q) s -= A[j * n + q] * A[j * n + q];
    if (!(s > 0.0)) return false;
    const double d = sqrt(s);
    A[j * n + j] = d;
#pragma unroll 1
    for (int r = j + 1; r < n; ++r) {
      double t = A[r * n + j];
#pragma unroll 1
      for (int q = 0; q < j; ++q) t -= A[r * n + q] * A[j * n + q];
      A[r * n + j] = t / d;
    }
  }
#pragma unroll 1
  for (int r = 0; r < n; ++r) {
    double s = b[r];
#pragma unroll 1
    for (int q = 0; q < r; ++q) s -= A[r * n + q] * b[q];
    b[r] = s / A[r * n + r];
  }
#pragma unroll 1
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
#pragma unroll 1
    for (int q = r + 1; q < n; ++q) s -= A[q * n + r] * b[q];
    b[r] = s / A[r * n + r];
  }
#pragma unroll 1
  for (int i = 0; i < 6; ++i) {
    double s = H[12 * (k - 1) + 6 + i];
#pragma unroll 1
    for (int j = 0; j < n; ++j) s -= (H[12 * (j + 1) + 6 + i] - H[12 * j + 6 + i]) * b[j];
    if (!(fabs(s) < 1e300)) return false;  // NaN or overflow
    x_new[i] = s;
  }
  return true;
}

// Pose (column-major 4x4) -> x relative to the row-major base rotation Rb.
// M: 9 doubles of workspace.
RTR_HD bool pose_to_x(const double* T, const double* Rb, double* x, double* M) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int r = i / 3, c = i % 3;  // (R Rb^T)[r][c]
    M[i] = (T[r] * Rb[c * 3] + T[4 + r] * Rb[c * 3 + 1]) + T[8 + r] * Rb[c * 3 + 2];
  }
  if (!so3_to_qv(M, x)) return false;
  x[3] = T[12];
  x[4] = T[13];
  x[5] = T[14];
  return true;
}

// x relative to Rb -> pose (column-major); false outside the chart.  E: 9
// doubles of workspace.
RTR_HD bool x_to_pose(const double* x, const double* Rb, double* T, double* E) {
  if (!qv_to_so3(x, E)) return false;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int r = i / 3, c = i % 3;  // (E Rb)[r][c]
    T[c * 4 + r] = (E[r * 3] * Rb[c] + E[r * 3 + 1] * Rb[3 + c]) + E[r * 3 + 2] * Rb[6 + c];
  }
  T[3] = T[7] = T[11] = 0.0;
  T[12] = x[3];
  T[13] = x[4];
  T[14] = x[5];
  T[15] = 1.0;
  return true;
}

// One accelerated update.  st: the problem's state (layout above, set up by
// anderson_reset); Tq: the transform this iteration queried with, mse the
// objective there (the caller's choice: non-increasing along plain steps);
// Tg: the plain ICP step from Tq.  Writes the next
// iterate to Tout (Tg itself when the step is not accelerated).  ws:
// kAndersonWs doubles.  Returns 1 when the step was accelerated.
RTR_HD int anderson_pose_step(double* st, int m, const double* Tq, const double* Tg, double mse, double* Tout,
                              double* ws) {
  double* x = ws + 80;  // 6
  double* M = ws + 88;  // 9
#pragma unroll
  for (int i = 0; i < 16; ++i) Tout[i] = Tg[i];
  if (m < 1) return 0;
  if (m > kAndersonMax) m = kAndersonMax;
  int k = (int)st[9];
  if (k > 0 && !(mse < st[10])) {
    // the extrapolated iterate is worse than the last accepted one: go back
    // to the plain step from that one (the last pair's g) and start a new
    // history there
    st[9] = 0.0;
    st[10] = 1.7976931348623157e308;
    if (x_to_pose(st + 12 + 12 * (k - 1) + 6, st, ws + 100, M)) {
#pragma unroll
      for (int i = 0; i < 16; ++i) Tout[i] = ws[100 + i];
      return 0;
    }
    k = 0;
  }
  if (k == 0) {
#pragma unroll
    for (int i = 0; i < 9; ++i) st[i] = Tq[(i % 3) * 4 + i / 3];
  }
  if (k == m + 1) {  // drop the oldest pair
    // loads in groups of 4, so they are in flight together
#pragma unroll 1
    for (int i = 12; i < 12 + 12 * m; i += 4) {
      const double t0 = st[i + 12], t1 = st[i + 13], t2 = st[i + 14], t3 = st[i + 15];
      st[i] = t0;
      st[i + 1] = t1;
      st[i + 2] = t2;
      st[i + 3] = t3;
    }
    k = m;
  }
  double* pr = st + 12 + 12 * k;
  if (!pose_to_x(Tq, st, pr, M) || !pose_to_x(Tg, st, pr + 6, M)) {
    st[9] = 0.0;
    st[10] = mse;
    return 0;
  }
  k += 1;
  st[9] = (double)k;
  st[10] = mse;  // after a restart the first MSE, else an improvement
  if (k < 2) return 0;
  if (!anderson_mix(st + 12, k, x, ws)) return 0;
  // stagnation: a mixed step far shorter than the plain one (noise in the
  // map dominates the history) would stall the iteration; take the plain
  // step and start a new history from it
  double da = 0.0, dp = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    da += (x[i] - pr[i]) * (x[i] - pr[i]);
    dp += (pr[6 + i] - pr[i]) * (pr[6 + i] - pr[i]);
  }
  if (da < 0.01 * dp) {
    st[9] = 0.0;
    return 0;
  }
  // trust region: an extrapolation beyond ten plain steps is shortened to ten
  if (da > 100.0 * dp) {
    const double f = 10.0 * sqrt(dp / da);
#pragma unroll
    for (int i = 0; i < 6; ++i) x[i] = pr[i] + f * (x[i] - pr[i]);
  }
  if (!x_to_pose(x, st, ws + 100, M)) return 0;  // outside the chart: plain step
#pragma unroll
  for (int i = 0; i < 16; ++i) Tout[i] = ws[100 + i];
  return 1;
}

RTR_HD void anderson_reset(double* st) {
  st[9] = 0.0;
  st[10] = 1.7976931348623157e308;  // DBL_MAX
}

// One correspondence added to FP64 moment sums (layout in rtr_types.hpp).
// q is the transformed source point, c its target point and nrm c's normal,
// all floats, and so is the anchor (A): q - A and c - A are exact in FP64
// (the difference of two floats needs at most 25 + the exponent gap bits),
// every product rounds once, and the sums carry no FP32 rounding at all.
// The moments are then those of the exact correspondences up to FP64
// rounding, whatever order lanes, waves and work items add them in.
template <int MODE>
RTR_HD void accumulate_pair(float qx, float qy, float qz, float cx, float cy, float cz, float nx, float ny,
                            float nz, float d2, double Ax, double Ay, double Az, double* acc) {
  const double s0 = (double)qx - Ax, s1 = (double)qy - Ay, s2 = (double)qz - Az;
  const double t0 = (double)cx - Ax, t1 = (double)cy - Ay, t2 = (double)cz - Az;
  acc[SUM_N] += 1.0;
  acc[SUM_D2] += (double)d2;
  if (MODE == RTR_MODE_POINT_TO_POINT) {
    acc[SUM_A + 0] += s0;
    acc[SUM_A + 1] += s1;
    acc[SUM_A + 2] += s2;
    acc[SUM_B + 0] += t0;
    acc[SUM_B + 1] += t1;
    acc[SUM_B + 2] += t2;
    const double a[3] = {s0, s1, s2}, b[3] = {t0, t1, t2};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[SUM_AB + r * 3 + c] = fma(a[r], b[c], acc[SUM_AB + r * 3 + c]);
  } else {
    // PCL's TransformationEstimationPointToPlaneLLS row (s x n, n) and
    // right-hand side n . (c - q), about the anchor
    const double n0 = nx, n1 = ny, n2 = nz;
    const double row[6] = {n2 * s1 - n1 * s2, n0 * s2 - n2 * s0, n1 * s0 - n0 * s1, n0, n1, n2};
    double rhs = n0 * (t0 - s0);
    rhs = rhs + n1 * (t1 - s1);
    rhs = rhs + n2 * (t2 - s2);
    int k = 0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = r; c < 6; ++c, ++k) acc[SUM_ATA + k] = fma(row[r], row[c], acc[SUM_ATA + k]);
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[SUM_ATB + r] = fma(row[r], rhs, acc[SUM_ATB + r]);
  }
}

}  // namespace rtr
