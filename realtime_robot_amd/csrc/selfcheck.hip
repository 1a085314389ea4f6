// Self-check hooks: the device geometry code (geom.hpp, __host__ __device__)
// run on the host, so that tests can compare the exact grid search, the FP64
// moment sums and solves, and the Anderson mixing with the CPU oracle without
// a GPU.  Host code only: this file also builds without device code (the
// sanitizer build of the host paths).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rtr.h"
#include "geom.hpp"
#include "guard.hpp"

using rtr::set_error;

extern "C" {

// ---------------------------------------------------------------------------
// self-check hooks: the same geometry code on the host
// ---------------------------------------------------------------------------

int rtr_selfcheck_nn_stats(const float* tgt, size_t nt, const float* q, size_t nq, float cell,
                           double max_dist, int32_t* idx, float* d2, uint32_t* ranges,
                           uint32_t* cands) {
  RTR_GUARD_BEGIN
  using namespace rtr;
  if ((nt && !tgt) || (nq && (!q || !idx || !d2)) || !(cell > 0)) return RTR_ERR_INVALID_ARG;
  // bbox + grid exactly as the device build chooses it (cell edge fixed)
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  size_t nv = 0;
  for (size_t i = 0; i < nt; ++i) {
    const float* p = tgt + 3 * i;
    if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) continue;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], p[a]);
      hi[a] = std::fmax(hi[a], p[a]);
    }
    ++nv;
  }
  GridView g;
  if (nv == 0) {
    lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.f;
  }
  g.h = cell;
  g.inv_h = 1.0f / cell;
  g.dx = (int)std::floor((hi[0] - lo[0]) * g.inv_h) + 1;
  g.dy = (int)std::floor((hi[1] - lo[1]) * g.inv_h) + 1;
  g.dz = (int)std::floor((hi[2] - lo[2]) * g.inv_h) + 1;
  if ((double)g.dx * g.dy * g.dz > 64.0e6) {
    set_error("selfcheck grid too large");
    return RTR_ERR_CAPACITY;
  }
  g.ox = lo[0];
  g.oy = lo[1];
  g.oz = lo[2];
  float mag = 0.f, ext = 0.f;
  for (int a = 0; a < 3; ++a) {
    mag = std::fmax(mag, std::fmax(std::fabs(lo[a]), std::fabs(hi[a])));
    ext = std::fmax(ext, hi[a] - lo[a]);
  }
  g.margin = 1e-3f * cell + 4e-6f * (mag + ext);
  const size_t ncell = (size_t)g.dx * g.dy * g.dz;
  std::vector<uint32_t> cnt(ncell + 1, 0), start(ncell + 1, 0), cid(nt, 0xffffffffu);
  for (size_t i = 0; i < nt; ++i) {
    const float* p = tgt + 3 * i;
    if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) continue;
    int ix = clampi((int)std::floor((p[0] - g.ox) * g.inv_h), 0, g.dx - 1);
    int iy = clampi((int)std::floor((p[1] - g.oy) * g.inv_h), 0, g.dy - 1);
    int iz = clampi((int)std::floor((p[2] - g.oz) * g.inv_h), 0, g.dz - 1);
    cid[i] = ((uint32_t)iz * g.dy + iy) * g.dx + ix;
    cnt[cid[i]]++;
  }
  for (size_t c = 0; c < ncell; ++c) start[c + 1] = start[c] + cnt[c];
  std::vector<float4> pts(nv ? nv : 1);
  std::vector<uint32_t> fill(start.begin(), start.end());
  for (size_t i = 0; i < nt; ++i) {
    if (cid[i] == 0xffffffffu) continue;
    union {
      int32_t i;
      float f;
    } w;
    w.i = (int32_t)i;
    pts[fill[cid[i]]++] = make_float4(tgt[3 * i], tgt[3 * i + 1], tgt[3 * i + 2], w.f);
  }
  g.pts = pts.data();
  g.cell_start = start.data();
  g.n_pts = (uint32_t)nv;
  g.pt_lo = 0;
  const double r2 = max_dist * max_dist;
  float r2f;
  if (!(r2 < (double)FLT_MAX))
    r2f = FLT_MAX;
  else {
    r2f = (float)r2;
    if ((double)r2f > r2) r2f = std::nextafter(r2f, 0.0f);
  }
  // distance field exactly as the device grid build makes it
  int skip_ring, dt_cap;
  dist_caps(max_dist < 3.0e38 ? (float)max_dist : 3.0e38f, g.h, g.margin, &skip_ring, &dt_cap);
  std::vector<uint8_t> da(ncell), db(ncell);
  const int dims[3] = {g.dx, g.dy, g.dz};
  const int strides[3] = {1, g.dx, g.dx * g.dy};
  for (int axis = 0; axis < 3; ++axis) {
    const std::vector<uint8_t>& src = axis == 1 ? da : db;
    std::vector<uint8_t>& dst = axis == 1 ? db : da;
    for (size_t c = 0; c < ncell; ++c) {
      const int pos = axis == 0 ? (int)(c % g.dx) : (axis == 1 ? (int)((c / g.dx) % g.dy) : (int)(c / ((size_t)g.dx * g.dy)));
      int best = dt_cap + 1;
      for (int d = -dt_cap; d <= dt_cap; ++d) {
        if (pos + d < 0 || pos + d >= dims[axis]) continue;
        const size_t cc = (size_t)((long long)c + (long long)d * strides[axis]);
        int v = axis == 0 ? (start[cc + 1] > start[cc] ? 0 : dt_cap + 1) : src[cc];
        const int ad = d < 0 ? -d : d;
        v = v > ad ? v : ad;
        best = v < best ? v : best;
      }
      dst[c] = (uint8_t)best;
    }
  }
  for (size_t i = 0; i < nq; ++i) {
    NNCountVisitor v;
    v.pts = g.pts;
    v.qx = q[3 * i];
    v.qy = q[3 * i + 1];
    v.qz = q[3 * i + 2];
    v.best.key = ~0ull;
    v.best.tx = v.best.ty = v.best.tz = 0.f;
    v.best.slot = 0xffffffffu;
    ring_walk_dt(g, da.data(), skip_ring, v.qx, v.qy, v.qz, r2f, v);
    NNBest b = v.best;
    if (b.key != ~0ull && !(key_d2(b.key) <= r2f)) b.key = ~0ull;
    if (ranges) ranges[i] = v.ranges;
    if (cands) cands[i] = v.cands;
    if (b.key != ~0ull) {
      idx[i] = (int32_t)(uint32_t)(b.key & 0xffffffffull);
      d2[i] = key_d2(b.key);
    } else {
      idx[i] = -1;
      d2[i] = INFINITY;
    }
  }
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_selfcheck_nn(const float* tgt, size_t nt, const float* q, size_t nq, float cell,
                     double max_dist, int32_t* idx, float* d2) {
  return rtr_selfcheck_nn_stats(tgt, nt, q, nq, cell, max_dist, idx, d2, nullptr, nullptr);
}

int rtr_selfcheck_solve(int mode, const float* a, const float* b, const float* nrm, size_t n,
                        const float anchor[3], double Tinc[16]) {
  RTR_GUARD_BEGIN
  using namespace rtr;
  if (!a || !b || !anchor || !Tinc || (mode == RTR_MODE_POINT_TO_PLANE && !nrm))
    return RTR_ERR_INVALID_ARG;
  double v[NSUM_MAX] = {0};
  const double A[3] = {(double)anchor[0], (double)anchor[1], (double)anchor[2]};
  for (size_t i = 0; i < n; ++i) {
    const float nx = nrm ? nrm[3 * i] : 0.f, ny = nrm ? nrm[3 * i + 1] : 0.f, nz = nrm ? nrm[3 * i + 2] : 0.f;
    if (mode == RTR_MODE_POINT_TO_POINT)
      accumulate_pair<RTR_MODE_POINT_TO_POINT>(a[3 * i], a[3 * i + 1], a[3 * i + 2], b[3 * i], b[3 * i + 1],
                                               b[3 * i + 2], nx, ny, nz, 0.f, A[0], A[1], A[2], v);
    else
      accumulate_pair<RTR_MODE_POINT_TO_PLANE>(a[3 * i], a[3 * i + 1], a[3 * i + 2], b[3 * i], b[3 * i + 1],
                                               b[3 * i + 2], nx, ny, nz, 0.f, A[0], A[1], A[2], v);
  }
  double mse, ws[64];
  solve_increment(mode, v, anchor, Tinc, &mse, ws);
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_selfcheck_anderson(int m, const double A[36], const double c[6], double tol, int max_iter,
                           double x_out[6], int* iters) {
  RTR_GUARD_BEGIN
  using namespace rtr;
  if (!A || !c || !x_out || !iters || m < 1 || m > kAndersonMax || max_iter < 1) return RTR_ERR_INVALID_ARG;
  // history of (x_j, g_j), oldest first, as anderson_pose_step keeps it
  double H[12 * (kAndersonMax + 1)], ws[kAndersonWs], x[6] = {0, 0, 0, 0, 0, 0};
  int k = 0, n_eval = 0;
  for (int it = 0; it < max_iter; ++it) {
    double g[6];
    for (int r = 0; r < 6; ++r) {
      double s = c[r];
      for (int q = 0; q < 6; ++q) s += A[r * 6 + q] * x[q];
      g[r] = s;
    }
    ++n_eval;
    double res = 0.0;
    for (int r = 0; r < 6; ++r) res = fmax(res, fabs(g[r] - x[r]));
    for (int r = 0; r < 6; ++r) x_out[r] = g[r];
    if (res <= tol) break;
    if (k == m + 1) {
      for (int i = 0; i < 12 * m; ++i) H[i] = H[i + 12];
      k = m;
    }
    for (int r = 0; r < 6; ++r) {
      H[12 * k + r] = x[r];
      H[12 * k + 6 + r] = g[r];
    }
    ++k;
    double xn[6];
    if (k >= 2 && anderson_mix(H, k, xn, ws)) {
      for (int r = 0; r < 6; ++r) x[r] = xn[r];
    } else {
      for (int r = 0; r < 6; ++r) x[r] = g[r];
    }
  }
  *iters = n_eval;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_selfcheck_so3(const double v[3], double R[9], double v_back[3]) {
  RTR_GUARD_BEGIN
  using namespace rtr;
  if (!v || !R || !v_back) return RTR_ERR_INVALID_ARG;
  if (!qv_to_so3(v, R) || !so3_to_qv(R, v_back)) {
    set_error("rotation outside the chart (|v| > 1 or angle beyond ~174 degrees)");
    return RTR_ERR_INVALID_ARG;
  }
  return RTR_OK;
  RTR_GUARD_END
}

}  // extern "C"
