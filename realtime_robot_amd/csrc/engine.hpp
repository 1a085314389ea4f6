// Host-side registration engine: owns the device arena of one batch of
// clouds/problems, builds grids, drives the persistent ICP kernel and replays
// the whole align from a hipGraph.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/rtr.h"
#include "engine_dev.hpp"
#include "guard.hpp"

namespace rtr {

void set_error(const char* fmt, ...);

#define RTR_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess) {                                                                \
      ::rtr::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
      return RTR_ERR_DEVICE;                                                               \
    }                                                                                      \
  } while (0)


// Grow-only device allocation (allocations happen in set_* calls, never in a
// captured or timed align).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n);
  void release();
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct TuneParams {
  float cell_factor = 0.25f;      // h_req = cell_factor * max correspondence distance
  float cells_per_point = 2.0f;   // grid cell budget per target point
  int item_pts = 1024;            // source points per work item
  int blocks_per_cu = 0;          // persistent workgroups per CU (0 = occupancy limit)
  int use_tile = 0;               // LDS candidate tiles on/off
  float retire = 0.f;             // surplus workgroups leave below this fraction of unfinished problems
  int fine_from = 0;              // passes from this one on run 256-point work items (0: never)
};

// Host clouds (xyz, `stride` bytes apart, 12 when stride is null or 0) to one
// packed xyz array at dst on the device, cloud c at point offset off[c]
// (ascending), through the caller's pinned staging buffer (grown as needed):
// host threads pack it a chunk at a time while the previous chunk's copy
// runs on `stream`; returns once every copy has finished.
int upload_xyz_staged(hipStream_t stream, void** staging, size_t* staging_bytes, int count,
                      const float* const* clouds, const size_t* n, const size_t* stride, const size_t* off,
                      size_t total, void* dst);

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  int init();

  int set_clouds(int n_src, const float* const* src, const size_t* ns, const size_t* s_stride,
                 int n_tgt, const float* const* tgt, const size_t* nt, const size_t* t_stride);
  int set_target_normals(const float* const* nrm, const size_t* stride);
  int compute_target_normals(int k, const float vp[3]);
  // the resident clouds as packed xyz (12 B per point): sources in their
  // order, then targets in theirs (set_clouds)
  const float* source_xyz() const { return static_cast<const float*>(src_raw_.p); }
  const float* target_xyz() const { return static_cast<const float*>(tgt_raw_.p); }
  int set_problems(int n, const int* src_id, const int* tgt_id);
  int set_params(const rtr_params& p);
  const rtr_params& params() const { return prm_; }
  void set_graph(bool on) { use_graph_ = on; }
  void set_profiling(bool on) { profiling_ = on; }
  void set_rebuild_always(bool on) {
    if (on != rebuild_always_) graph_dirty_ = true;
    rebuild_always_ = on;
  }
  // Builds the grids and source orders now (aligns then reuse them until the
  // clouds, parameters or tuning change, unless rebuild-always is on).
  int build();
  TuneParams& tune() { return tune_; }
  // the next align (or build) rebuilds the grids: cell settings changed
  void invalidate_grids() { grid_dirty_ = true; }
  void invalidate_graph() { graph_dirty_ = true; }

  int align(const double* guesses, rtr_result* results);
  int correspondences(int prob, const double T[16], double max_dist, int32_t* idx, float* d2);
  int transformed_source(int prob, float* out, size_t stride);
  int fitness(double max_range, double* out);
  int judge(int k, int max_pts, double max_curv, double* plane_rmse, double* inlier_frac, double* planar_frac);
  int grid_info(int tgt, float* cell, int dims[3], float anchor[3]);
  int history(int prob, float* rows12, int cap, int* n_out);
  void set_timeline(bool on) {
    if (on != timeline_) graph_dirty_ = true;
    timeline_ = on;
  }
  int timeline(int prob, uint64_t* stamps, int cap, int* n_out, double* tick_ns);
  int last_correspondences(int prob, int32_t* idx, float* d2);
  int target_normals(int tgt, float* out);
  const rtr_timing& timing() const { return timing_; }
  int n_problems() const { return (int)probs_.size(); }
  size_t source_size(int prob) const;
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }

 private:
  int ensure_built();
  int enqueue_build();
  int enqueue_align_body();
  DevEngine dev_view() const;
  int upload_problems();
  int layout_targets();
  int layout_problems();
  int grid_blocks() const;

  int device_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  bool graph_dirty_ = true;
  bool use_graph_ = true;
  bool profiling_ = false;
  bool rebuild_always_ = true;
  bool timeline_ = false;
  bool grid_dirty_ = true;
  bool results_valid_ = false;  // states/grids hold a completed align of the current layout
  bool normals_valid_ = false;
  int cu_count_ = 0;
  rtr_params prm_;
  TuneParams tune_;
  rtr_timing timing_;

  // host mirrors
  std::vector<TargetDesc> tgts_;
  std::vector<SourceDesc> srcs_;
  std::vector<Problem> probs_;
  std::vector<uint2> items_;
  size_t tgt_total_ = 0, src_total_ = 0, pt_total_ = 0;
  int max_tgt_ = 0, max_src_ = 0;
  uint32_t n_cells_ = 0, n_bins_ = 0;
  int max_cells_ = 0;
  uint32_t xcd_off_[kXcds + 1] = {0};
  uint32_t fxcd_off_[kXcds + 1] = {0};
  // [0] problems done, [32] error, [64..65] searches (a 64-bit count),
  // [96 + (pass * kXcds + x) * stride] claim counters
  int claim_stride() const { return prm_.max_iterations <= 256 ? 32 : 1; }
  // the 8-deep search kernel (2 waves per SIMD) serves launches tuned to at
  // most 2 workgroups per CU, where a third wave would not be resident
  // ... and the 4-wave variant launches tuned to 4 workgroups per CU
  int variant() const {
    return (tune_.blocks_per_cu == 1 || tune_.blocks_per_cu == 2) ? 1 : (tune_.blocks_per_cu == 4 ? 2 : 0);
  }
  size_t counter_words() const { return 96 + (size_t)prm_.max_iterations * kXcds * claim_stride(); }

  // device arena
  DevBuf tgt_raw_, tgt_nrm_raw_, tgt_sorted_, tgt_nrm_sorted_, cell_count_, cell_start_, cell_id_, cell_rank_;
  DevBuf cell_dist_, cell_dist_tmp_;
  DevBuf tgt_desc_, bbox_tmp_;
  DevBuf src_raw_, sx_, sy_, sz_, perm_, bin_count_, bin_start_, bin_id_, bin_rank_, src_desc_;
  DevBuf prob_desc_, state_, guesses_, items_buf_, partials_, done_, counters_, scan_tmp_;
  DevBuf out_i32_, out_f32_, out_T_, hist_, aa_, nn_seed_, tline_;
  // pinned host staging
  void* pin_ = nullptr;
  size_t pin_bytes_ = 0;
  int ensure_pinned(size_t bytes);
  int upload_xyz(int count, const float* const* clouds, const size_t* n, const size_t* stride,
                 const size_t* off, size_t total, void* dst);
};

}  // namespace rtr
