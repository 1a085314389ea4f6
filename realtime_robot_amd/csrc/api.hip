// C ABI of librtr (declared in include/rtr.h).  Thin: argument checks,
// handle bookkeeping and status codes around rtr::Engine.  No C++ exception
// crosses this boundary.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rtr.h"
#include "engine.hpp"
#include "geom.hpp"

namespace rtr {
const char* last_error_str();
}

using rtr::Engine;
using rtr::set_error;

#define RTR_GUARD_BEGIN try {
#define RTR_GUARD_END                                 \
  }                                                   \
  catch (const std::bad_alloc&) {                     \
    set_error("host allocation failed");              \
    return RTR_ERR_CAPACITY;                          \
  }                                                   \
  catch (...) {                                       \
    set_error("unexpected host exception");           \
    return RTR_ERR_DEVICE;                            \
  }

extern "C" {

const char* rtr_last_error(void) { return rtr::last_error_str(); }

int rtr_version(void) { return 10000; }

int rtr_params_default(rtr_params* p) {
  if (!p) return RTR_ERR_INVALID_ARG;
  memset(p, 0, sizeof(*p));
  p->max_iterations = 10;
  p->min_correspondences = 3;
  p->mode = RTR_MODE_POINT_TO_POINT;
  p->max_correspondence_distance = std::sqrt(DBL_MAX);
  p->transformation_epsilon = 0.0;
  p->rotation_epsilon = 0.0;
  p->euclidean_fitness_epsilon = -DBL_MAX;
  p->cell_size = 0.f;
  p->cells_per_point = 0.f;
  return RTR_OK;
}

// ---------------------------------------------------------------------------
// single registration
// ---------------------------------------------------------------------------

struct rtr_icp {
  Engine eng;
  std::vector<float> src, tgt, nrm;  // packed host copies (PCL keeps shared_ptrs)
  bool have_src = false, have_tgt = false, have_nrm = false, aligned = false;
  bool src_dirty = true, tgt_dirty = true, problem_set = false;
  rtr_result last;
  explicit rtr_icp(int dev) : eng(dev) { memset(&last, 0, sizeof(last)); }
};

static int icp_sync_clouds(rtr_icp* h) {
  if (!h->have_src || h->src.empty()) {
    set_error("source cloud is empty or not set");
    return RTR_ERR_EMPTY_INPUT;
  }
  if (!h->have_tgt || h->tgt.empty()) {
    set_error("target cloud is empty or not set");
    return RTR_ERR_EMPTY_INPUT;
  }
  if (!h->src_dirty && !h->tgt_dirty) return RTR_OK;
  const float* s = h->src.data();
  const float* t = h->tgt.data();
  size_t ns = h->src.size() / 3, nt = h->tgt.size() / 3;
  // only the clouds that changed are re-uploaded (device-computed target
  // normals survive a new source)
  int rc = h->eng.set_clouds(h->src_dirty ? 1 : -1, h->src_dirty ? &s : nullptr, &ns, nullptr,
                             h->tgt_dirty ? 1 : -1, h->tgt_dirty ? &t : nullptr, &nt, nullptr);
  if (rc) return rc;
  if (h->tgt_dirty && h->have_nrm) {
    const float* n = h->nrm.data();
    rc = h->eng.set_target_normals(&n, nullptr);
    if (rc) return rc;
  }
  if (!h->problem_set) {
    int zero = 0;
    rc = h->eng.set_problems(1, &zero, &zero);
    if (rc) return rc;
    h->problem_set = true;
  }
  h->src_dirty = h->tgt_dirty = false;
  return RTR_OK;
}

static int copy_cloud(std::vector<float>& dst, const float* xyz, size_t n, size_t stride) {
  if (n && !xyz) {
    set_error("null point array with n = %zu", n);
    return RTR_ERR_INVALID_ARG;
  }
  if (stride != 0 && stride < 12) {
    set_error("stride %zu is smaller than 12 bytes", stride);
    return RTR_ERR_INVALID_ARG;
  }
  if (stride == 0) stride = 12;
  dst.resize(n * 3);
  const char* b = reinterpret_cast<const char*>(xyz);
  for (size_t i = 0; i < n; ++i) memcpy(dst.data() + 3 * i, b + i * stride, 12);
  return RTR_OK;
}

int rtr_icp_create(rtr_icp** out, int device) {
  RTR_GUARD_BEGIN
  if (!out) return RTR_ERR_INVALID_ARG;
  *out = nullptr;
  rtr_icp* h = new rtr_icp(device);
  int rc = h->eng.init();
  if (rc) {
    delete h;
    return rc;
  }
  h->eng.set_rebuild_always(false);
  h->eng.set_graph(false);
  *out = h;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_icp_destroy(rtr_icp* h) {
  RTR_GUARD_BEGIN
  delete h;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_icp_set_input_source(rtr_icp* h, const float* xyz, size_t n, size_t stride) {
  RTR_GUARD_BEGIN
  if (!h) return RTR_ERR_INVALID_ARG;
  int rc = copy_cloud(h->src, xyz, n, stride);
  if (rc) return rc;
  h->have_src = true;
  h->src_dirty = true;
  h->aligned = false;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_icp_set_input_target(rtr_icp* h, const float* xyz, size_t n, size_t stride) {
  RTR_GUARD_BEGIN
  if (!h) return RTR_ERR_INVALID_ARG;
  int rc = copy_cloud(h->tgt, xyz, n, stride);
  if (rc) return rc;
  h->have_tgt = true;
  h->have_nrm = false;
  h->tgt_dirty = true;
  h->aligned = false;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_icp_set_target_normals(rtr_icp* h, const float* nxyz, size_t n, size_t stride) {
  RTR_GUARD_BEGIN
  if (!h) return RTR_ERR_INVALID_ARG;
  if (!h->have_tgt || n != h->tgt.size() / 3) {
    set_error("normals count %zu does not match the target (%zu)", n, h->tgt.size() / 3);
    return RTR_ERR_INVALID_ARG;
  }
  int rc = copy_cloud(h->nrm, nxyz, n, stride);
  if (rc) return rc;
  h->have_nrm = true;
  h->tgt_dirty = true;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_icp_compute_target_normals(rtr_icp* h, int k, const float vp[3]) {
  RTR_GUARD_BEGIN
  if (!h) return RTR_ERR_INVALID_ARG;
  if (h->have_nrm) {
    h->have_nrm = false;
    h->tgt_dirty = true;
  }
  int rc = icp_sync_clouds(h);
  if (rc) return rc;
  const float zero[3] = {0.f, 0.f, 0.f};
  return h->eng.compute_target_normals(k, vp ? vp : zero);
  RTR_GUARD_END
}

int rtr_icp_get_target_normals(rtr_icp* h, float* nxyz) {
  RTR_GUARD_BEGIN
  if (!h || !nxyz) return RTR_ERR_INVALID_ARG;
  int rc = icp_sync_clouds(h);
  if (rc) return rc;
  return h->eng.target_normals(0, nxyz);
  RTR_GUARD_END
}

static int icp_update_param(rtr_icp* h, void (*f)(rtr_params&, double), double v) {
  if (!h) return RTR_ERR_INVALID_ARG;
  rtr_params p = h->eng.params();
  f(p, v);
  return h->eng.set_params(p);
}

int rtr_icp_set_max_correspondence_distance(rtr_icp* h, double d) {
  RTR_GUARD_BEGIN
  return icp_update_param(h, [](rtr_params& p, double v) { p.max_correspondence_distance = v; }, d);
  RTR_GUARD_END
}
int rtr_icp_set_maximum_iterations(rtr_icp* h, int n) {
  RTR_GUARD_BEGIN
  return icp_update_param(h, [](rtr_params& p, double v) { p.max_iterations = (int)v; }, (double)n);
  RTR_GUARD_END
}
int rtr_icp_set_transformation_epsilon(rtr_icp* h, double e) {
  RTR_GUARD_BEGIN
  return icp_update_param(h, [](rtr_params& p, double v) { p.transformation_epsilon = v; }, e);
  RTR_GUARD_END
}
int rtr_icp_set_euclidean_fitness_epsilon(rtr_icp* h, double e) {
  RTR_GUARD_BEGIN
  return icp_update_param(h, [](rtr_params& p, double v) { p.euclidean_fitness_epsilon = v; }, e);
  RTR_GUARD_END
}
int rtr_icp_set_params(rtr_icp* h, const rtr_params* p) {
  RTR_GUARD_BEGIN
  if (!h || !p) return RTR_ERR_INVALID_ARG;
  return h->eng.set_params(*p);
  RTR_GUARD_END
}
int rtr_icp_get_params(rtr_icp* h, rtr_params* p) {
  if (!h || !p) return RTR_ERR_INVALID_ARG;
  *p = h->eng.params();
  return RTR_OK;
}

int rtr_icp_align(rtr_icp* h, const double guess[16], float* out_xyz, size_t out_stride,
                  double final_T[16], int* converged) {
  RTR_GUARD_BEGIN
  if (!h) return RTR_ERR_INVALID_ARG;
  int rc = icp_sync_clouds(h);
  if (rc) return rc;
  rtr_result r;
  rc = h->eng.align(guess, &r);
  if (rc) return rc;
  h->last = r;
  h->aligned = true;
  if (final_T) memcpy(final_T, r.T, sizeof(r.T));
  if (converged) *converged = r.converged;
  if (out_xyz) {
    rc = h->eng.transformed_source(0, out_xyz, out_stride);
    if (rc) return rc;
  }
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_icp_get_result(rtr_icp* h, rtr_result* out) {
  if (!h || !out) return RTR_ERR_INVALID_ARG;
  if (!h->aligned) {
    set_error("align has not run since the inputs changed");
    return RTR_ERR_NOT_READY;
  }
  *out = h->last;
  return RTR_OK;
}

int rtr_icp_get_fitness_score(rtr_icp* h, double max_range, double* out) {
  RTR_GUARD_BEGIN
  if (!h || !out) return RTR_ERR_INVALID_ARG;
  if (!h->aligned) {
    set_error("align has not run since the inputs changed");
    return RTR_ERR_NOT_READY;
  }
  return h->eng.fitness(max_range, out);
  RTR_GUARD_END
}

int rtr_icp_get_correspondences(rtr_icp* h, const double T[16], int32_t* idx, float* d2) {
  RTR_GUARD_BEGIN
  if (!h || !T || !idx || !d2) return RTR_ERR_INVALID_ARG;
  int rc = icp_sync_clouds(h);
  if (rc) return rc;
  return h->eng.correspondences(0, T, h->eng.params().max_correspondence_distance, idx, d2);
  RTR_GUARD_END
}

int rtr_icp_get_anchor(rtr_icp* h, float anchor[3]) {
  RTR_GUARD_BEGIN
  if (!h || !anchor) return RTR_ERR_INVALID_ARG;
  int rc = icp_sync_clouds(h);
  if (rc) return rc;
  double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  std::vector<int32_t> idx(h->src.size() / 3);
  std::vector<float> d2(h->src.size() / 3);
  // make sure the grid (and with it the anchor) exists
  rc = h->eng.correspondences(0, T, 1e-30, idx.data(), d2.data());
  if (rc) return rc;
  return h->eng.grid_info(0, nullptr, nullptr, anchor);
  RTR_GUARD_END
}

int rtr_icp_set_profiling(rtr_icp* h, int on) {
  if (!h) return RTR_ERR_INVALID_ARG;
  h->eng.set_profiling(on != 0);
  return RTR_OK;
}

int rtr_icp_get_timing(rtr_icp* h, rtr_timing* out) {
  if (!h || !out) return RTR_ERR_INVALID_ARG;
  *out = h->eng.timing();
  return RTR_OK;
}

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------

struct rtr_batch {
  Engine eng;
  std::vector<rtr_result> last;
  explicit rtr_batch(int dev) : eng(dev) {}
};

int rtr_batch_create(rtr_batch** out, int device) {
  RTR_GUARD_BEGIN
  if (!out) return RTR_ERR_INVALID_ARG;
  *out = nullptr;
  rtr_batch* b = new rtr_batch(device);
  int rc = b->eng.init();
  if (rc) {
    delete b;
    return rc;
  }
  b->eng.set_rebuild_always(true);
  b->eng.set_graph(true);
  *out = b;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_batch_destroy(rtr_batch* b) {
  RTR_GUARD_BEGIN
  delete b;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_batch_set_clouds(rtr_batch* b, int n_src, const float* const* src, const size_t* ns,
                         int n_tgt, const float* const* tgt, const size_t* nt) {
  RTR_GUARD_BEGIN
  if (!b || n_src < 1 || n_tgt < 1 || !src || !ns || !tgt || !nt) return RTR_ERR_INVALID_ARG;
  for (int i = 0; i < n_src; ++i)
    if (ns[i] && !src[i]) return RTR_ERR_INVALID_ARG;
  for (int j = 0; j < n_tgt; ++j)
    if (nt[j] && !tgt[j]) return RTR_ERR_INVALID_ARG;
  return b->eng.set_clouds(n_src, src, ns, nullptr, n_tgt, tgt, nt, nullptr);
  RTR_GUARD_END
}

int rtr_batch_set_target_normals(rtr_batch* b, const float* const* nrm) {
  RTR_GUARD_BEGIN
  if (!b || !nrm) return RTR_ERR_INVALID_ARG;
  return b->eng.set_target_normals(nrm, nullptr);
  RTR_GUARD_END
}

int rtr_batch_compute_target_normals(rtr_batch* b, int k, const float vp[3]) {
  RTR_GUARD_BEGIN
  if (!b) return RTR_ERR_INVALID_ARG;
  const float zero[3] = {0.f, 0.f, 0.f};
  return b->eng.compute_target_normals(k, vp ? vp : zero);
  RTR_GUARD_END
}

int rtr_batch_set_problems(rtr_batch* b, int n, const int* src_id, const int* tgt_id) {
  RTR_GUARD_BEGIN
  if (!b || n < 1 || !src_id || !tgt_id) return RTR_ERR_INVALID_ARG;
  return b->eng.set_problems(n, src_id, tgt_id);
  RTR_GUARD_END
}

int rtr_batch_set_params(rtr_batch* b, const rtr_params* p) {
  RTR_GUARD_BEGIN
  if (!b || !p) return RTR_ERR_INVALID_ARG;
  return b->eng.set_params(*p);
  RTR_GUARD_END
}

int rtr_batch_set_tuning(rtr_batch* b, float cell_factor, float cells_per_point, int item_pts,
                         int blocks_per_cu, int use_tile) {
  RTR_GUARD_BEGIN
  if (!b) return RTR_ERR_INVALID_ARG;
  if (item_pts > 0 && (item_pts % 256 != 0 || item_pts > (1 << 20))) {
    set_error("item_pts must be a positive multiple of 256 (got %d)", item_pts);
    return RTR_ERR_INVALID_ARG;
  }
  rtr::TuneParams& t = b->eng.tune();
  if ((cell_factor > 0 && cell_factor != t.cell_factor) || (cells_per_point > 0 && cells_per_point != t.cells_per_point))
    b->eng.invalidate_grids();
  if (cell_factor > 0) t.cell_factor = cell_factor;
  if (cells_per_point > 0) t.cells_per_point = cells_per_point;
  if (item_pts > 0) t.item_pts = item_pts;
  if (blocks_per_cu >= 0) t.blocks_per_cu = blocks_per_cu;
  if (use_tile >= 0) t.use_tile = use_tile != 0;
  return b->eng.set_params(b->eng.params());
  RTR_GUARD_END
}

int rtr_batch_get_tuning(rtr_batch* b, float* cell_factor, float* cells_per_point, int* item_pts,
                         int* blocks_per_cu, int* use_tile) {
  if (!b) return RTR_ERR_INVALID_ARG;
  const rtr::TuneParams& t = b->eng.tune();
  if (cell_factor) *cell_factor = t.cell_factor;
  if (cells_per_point) *cells_per_point = t.cells_per_point;
  if (item_pts) *item_pts = t.item_pts;
  if (blocks_per_cu) *blocks_per_cu = t.blocks_per_cu;
  if (use_tile) *use_tile = t.use_tile ? 1 : 0;
  return RTR_OK;
}

int rtr_batch_set_retire(rtr_batch* b, float below) {
  RTR_GUARD_BEGIN
  if (!b || !(below >= 0.f && below <= 1.f)) {
    set_error("rtr_batch_set_retire: the fraction must lie in [0, 1]");
    return RTR_ERR_INVALID_ARG;
  }
  b->eng.tune().retire = below;
  b->eng.invalidate_graph();
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_batch_get_retire(rtr_batch* b, float* below) {
  if (!b || !below) return RTR_ERR_INVALID_ARG;
  *below = b->eng.tune().retire;
  return RTR_OK;
}

int rtr_batch_set_rebuild(rtr_batch* b, int always) {
  if (!b) return RTR_ERR_INVALID_ARG;
  b->eng.set_rebuild_always(always != 0);
  return RTR_OK;
}

int rtr_batch_build(rtr_batch* b) {
  RTR_GUARD_BEGIN
  if (!b) return RTR_ERR_INVALID_ARG;
  return b->eng.build();
  RTR_GUARD_END
}

int rtr_batch_set_graph(rtr_batch* b, int on) {
  if (!b) return RTR_ERR_INVALID_ARG;
  b->eng.set_graph(on != 0);
  return RTR_OK;
}

int rtr_batch_set_profiling(rtr_batch* b, int on) {
  if (!b) return RTR_ERR_INVALID_ARG;
  b->eng.set_profiling(on != 0);
  return RTR_OK;
}

int rtr_batch_align(rtr_batch* b, const double* guesses, rtr_result* results) {
  RTR_GUARD_BEGIN
  if (!b) return RTR_ERR_INVALID_ARG;
  return b->eng.align(guesses, results);
  RTR_GUARD_END
}

int rtr_batch_get_timing(rtr_batch* b, rtr_timing* out) {
  if (!b || !out) return RTR_ERR_INVALID_ARG;
  *out = b->eng.timing();
  return RTR_OK;
}

int rtr_batch_get_grid(rtr_batch* b, int tgt, float* cell, int dims[3], float anchor[3]) {
  RTR_GUARD_BEGIN
  if (!b) return RTR_ERR_INVALID_ARG;
  return b->eng.grid_info(tgt, cell, dims, anchor);
  RTR_GUARD_END
}

int rtr_batch_get_history(rtr_batch* b, int prob, float* rows12, int cap, int* n_out) {
  RTR_GUARD_BEGIN
  if (!b || !rows12 || !n_out) return RTR_ERR_INVALID_ARG;
  return b->eng.history(prob, rows12, cap, n_out);
  RTR_GUARD_END
}

int rtr_batch_get_last_correspondences(rtr_batch* b, int prob, int32_t* idx, float* d2) {
  RTR_GUARD_BEGIN
  if (!b || !idx || !d2) return RTR_ERR_INVALID_ARG;
  return b->eng.last_correspondences(prob, idx, d2);
  RTR_GUARD_END
}

int rtr_batch_fitness(rtr_batch* b, double max_range, double* out) {
  RTR_GUARD_BEGIN
  if (!b || !out) return RTR_ERR_INVALID_ARG;
  return b->eng.fitness(max_range, out);
  RTR_GUARD_END
}

int rtr_batch_judge(rtr_batch* b, int k, int max_points, double max_curvature, double* plane_rmse,
                    double* inlier_frac, double* planar_frac) {
  RTR_GUARD_BEGIN
  if (!b) return RTR_ERR_INVALID_ARG;
  return b->eng.judge(k, max_points, max_curvature, plane_rmse, inlier_frac, planar_frac);
  RTR_GUARD_END
}

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
  const int ns = mode == RTR_MODE_POINT_TO_POINT ? SUM_P2P_COUNT : SUM_P2L_COUNT;
  double v[NSUM_MAX] = {0};
  for (size_t i = 0; i < n; ++i) {
    float t[NSUM_MAX];
    point_terms(mode, a[3 * i] - anchor[0], a[3 * i + 1] - anchor[1], a[3 * i + 2] - anchor[2],
                b[3 * i] - anchor[0], b[3 * i + 1] - anchor[1], b[3 * i + 2] - anchor[2],
                nrm ? nrm[3 * i] : 0.f, nrm ? nrm[3 * i + 1] : 0.f, nrm ? nrm[3 * i + 2] : 0.f,
                0.f, t);
    for (int k = 0; k < ns; ++k) v[k] += t[k];
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
