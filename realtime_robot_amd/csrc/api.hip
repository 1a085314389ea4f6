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
#include "internal.h"
#include "geom.hpp"
#include "guard.hpp"

using rtr::Engine;
using rtr::set_error;


extern "C" {

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
  // One problem per launch: 256-point work items spread each iteration over
  // four times as many workgroups as 1024-point ones (a 10k-point source:
  // 40 items, 8.0 -> 4.3 ms for 50 iterations), and one persistent
  // workgroup per CU leaves fewer idle ones polling the queue.  The result
  // differs from the 1024-point layout only in the order of the FP64 moment
  // sums.
  h->eng.tune().item_pts = 256;
  h->eng.tune().blocks_per_cu = 1;
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

int rtr_batch_device_clouds(const rtr_batch* b, const float** src_xyz, const float** tgt_xyz) {
  if (!b || !src_xyz || !tgt_xyz) return RTR_ERR_INVALID_ARG;
  *src_xyz = b->eng.source_xyz();
  *tgt_xyz = b->eng.target_xyz();
  if (!*src_xyz || !*tgt_xyz) {
    set_error("rtr_batch_device_clouds: no clouds are resident");
    return RTR_ERR_INVALID_ARG;
  }
  return RTR_OK;
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

int rtr_batch_set_fine_items(rtr_batch* b, int from_iteration) {
  RTR_GUARD_BEGIN
  if (!b || from_iteration < 0 || from_iteration > 100000) {
    set_error("rtr_batch_set_fine_items: the iteration must lie in [0, 100000] (got %d)", from_iteration);
    return RTR_ERR_INVALID_ARG;
  }
  b->eng.tune().fine_from = from_iteration;
  return b->eng.set_params(b->eng.params());
  RTR_GUARD_END
}

int rtr_batch_get_fine_items(rtr_batch* b, int* from_iteration) {
  if (!b || !from_iteration) return RTR_ERR_INVALID_ARG;
  *from_iteration = b->eng.tune().fine_from;
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

int rtr_batch_set_timeline(rtr_batch* b, int on) {
  if (!b) return RTR_ERR_INVALID_ARG;
  b->eng.set_timeline(on != 0);
  b->eng.set_params(b->eng.params());
  return RTR_OK;
}

int rtr_batch_get_timeline(rtr_batch* b, int prob, uint64_t* stamps, int cap, int* n_out, double* tick_ns) {
  RTR_GUARD_BEGIN
  if (!b || !stamps || !n_out || !tick_ns) return RTR_ERR_INVALID_ARG;
  return b->eng.timeline(prob, stamps, cap, n_out, tick_ns);
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

}  // extern "C"
