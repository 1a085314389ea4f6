// Host-side registration engine (see engine.hpp).
#include "engine.hpp"

#include <cfloat>
#include "geom.hpp"
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

namespace rtr {

int DevBuf::ensure(size_t n) {
  if (n == 0) n = 256;
  if (p && n <= bytes) return RTR_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  RTR_HIP(hipMalloc(&p, n));
  bytes = n;
  return RTR_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}


Engine::Engine(int device) : device_(device) {
  rtr_params_default(&prm_);
  memset(&timing_, 0, sizeof(timing_));
}

int Engine::init() {
  int n = 0;
  RTR_HIP(hipGetDeviceCount(&n));
  if (device_ < 0 || device_ >= n) {
    set_error("device %d not available (%d devices)", device_, n);
    return RTR_ERR_DEVICE;
  }
  RTR_HIP(hipSetDevice(device_));
  RTR_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  for (auto& e : ev_) RTR_HIP(hipEventCreate(&e));
  RTR_HIP(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, device_));
  return RTR_OK;
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  DevBuf* bufs[] = {&tgt_raw_, &tgt_nrm_raw_, &tgt_sorted_, &tgt_nrm_sorted_, &cell_count_,
                    &cell_start_, &cell_id_, &cell_rank_, &cell_dist_, &cell_dist_tmp_, &tgt_desc_, &bbox_tmp_, &src_raw_, &sx_, &sy_, &sz_, &perm_,
                    &bin_count_, &bin_start_, &bin_id_, &bin_rank_, &src_desc_, &prob_desc_, &state_,
                    &guesses_, &items_buf_, &partials_, &done_, &counters_, &scan_tmp_,
                    &out_i32_, &out_f32_, &out_T_, &hist_, &aa_, &nn_seed_, &tline_};
  for (DevBuf* b : bufs) b->release();
  if (pin_) (void)hipHostFree(pin_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::ensure_pinned(size_t bytes) {
  if (pin_ && bytes <= pin_bytes_) return RTR_OK;
  if (pin_) (void)hipHostFree(pin_);
  pin_ = nullptr;
  pin_bytes_ = 0;
  RTR_HIP(hipHostMalloc(&pin_, bytes < 4096 ? 4096 : bytes, hipHostMallocDefault));
  pin_bytes_ = bytes < 4096 ? 4096 : bytes;
  return RTR_OK;
}

static void pack_xyz(const float* src, size_t n, size_t stride, float* dst) {
  if (stride == 0 || stride == 12) {
    memcpy(dst, src, n * 12);
    return;
  }
  const char* b = reinterpret_cast<const char*>(src);
  for (size_t i = 0; i < n; ++i) memcpy(dst + 3 * i, b + i * stride, 12);
}

// Host clouds to one packed xyz array on the device (cloud c at point offset
// off[c], offsets ascending, total points in all): the clouds are packed into
// the pinned staging buffer (grown to the total when smaller) a chunk at a time, by several host threads at
// once, and each chunk's copy is queued on the stream as soon as it is
// packed, so it runs while the next chunk is packed.  Returns once the last
// copy has finished.
int upload_xyz_staged(hipStream_t stream, void** staging, size_t* staging_bytes, int count,
                      const float* const* clouds, const size_t* n, const size_t* stride, const size_t* off,
                      size_t total, void* dst) {
  if (total == 0) return RTR_OK;
  if (!*staging || *staging_bytes < total * 12) {
    if (*staging) (void)hipHostFree(*staging);
    *staging = nullptr;
    *staging_bytes = 0;
    RTR_HIP(hipHostMalloc(staging, total * 12, hipHostMallocDefault));
    *staging_bytes = total * 12;
  }
  float* const pin = static_cast<float*>(*staging);
  constexpr size_t kChunkPts = (size_t)2 << 20;   // 24 MB per chunk
  constexpr size_t kMinPerThread = (size_t)1 << 17;
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned max_threads = hw == 0 ? 1u : std::min(hw, 8u);
  // points [lo, hi) of the concatenation, from every cloud they cover
  auto pack_range = [&](size_t lo, size_t hi) {
    int c = (int)(std::upper_bound(off, off + count, lo) - off) - 1;
    for (; c < count && lo < hi; ++c) {
      if (c < 0 || off[c] + n[c] <= lo) continue;
      const size_t a = lo - off[c], e = std::min(n[c], hi - off[c]);
      const size_t s = (stride && stride[c]) ? stride[c] : 12;
      pack_xyz(reinterpret_cast<const float*>(reinterpret_cast<const char*>(clouds[c]) + a * s), e - a, s,
               pin + 3 * lo);
      lo = off[c] + e;
    }
  };
  std::vector<std::thread> pool;
  for (size_t a = 0; a < total; a += kChunkPts) {
    const size_t b = std::min(total, a + kChunkPts);
    const unsigned t = (unsigned)std::max<size_t>(1, std::min<size_t>(max_threads, (b - a) / kMinPerThread));
    const size_t per = (b - a + t - 1) / t;
    pool.clear();
    for (unsigned i = 1; i < t; ++i) {
      const size_t lo = a + i * per, hi = std::min(b, lo + per);
      if (lo >= hi) continue;
      try {
        pool.emplace_back(pack_range, lo, hi);
      } catch (...) {   // no thread to be had: this one packs the part
        pack_range(lo, hi);
      }
    }
    pack_range(a, std::min(b, a + per));
    for (std::thread& th : pool) th.join();
    RTR_HIP(hipMemcpyAsync(static_cast<char*>(dst) + 12 * a, pin + 3 * a, (b - a) * 12, hipMemcpyHostToDevice,
                           stream));
  }
  RTR_HIP(hipStreamSynchronize(stream));
  return RTR_OK;
}

int Engine::upload_xyz(int count, const float* const* clouds, const size_t* n, const size_t* stride,
                       const size_t* off, size_t total, void* dst) {
  return upload_xyz_staged(stream_, &pin_, &pin_bytes_, count, clouds, n, stride, off, total, dst);
}

static uint32_t source_bins(size_t n) {
  int P = 1;
  double want = std::cbrt((double)n / 4.0);
  while (P < want && P < 32) P <<= 1;
  return (uint32_t)(P * P * P);
}

int Engine::set_clouds(int n_src, const float* const* src, const size_t* ns, const size_t* s_stride,
                       int n_tgt, const float* const* tgt, const size_t* nt,
                       const size_t* t_stride) {
  RTR_HIP(hipSetDevice(device_));
  RTR_HIP(hipStreamSynchronize(stream_));
  // sizes
  size_t st = 0, tt = 0;
  for (int i = 0; i < n_src; ++i) st += ns[i];
  for (int j = 0; j < n_tgt; ++j) tt += nt[j];
  if (st >= (1ull << 31) || tt >= (1ull << 31)) {
    set_error("clouds too large (%zu source, %zu target points)", st, tt);
    return RTR_ERR_CAPACITY;
  }
  if (n_src >= 0 && src) {
    srcs_.assign(n_src, SourceDesc{});
    size_t off = 0;
    uint32_t boff = 0;
    max_src_ = 0;
    for (int i = 0; i < n_src; ++i) {
      SourceDesc& d = srcs_[i];
      d.off = (uint32_t)off;
      d.n = (uint32_t)ns[i];
      d.bin_off = boff;
      d.bin_budget = source_bins(ns[i]);
      boff += d.bin_budget;
      off += ns[i];
      if ((int)ns[i] > max_src_) max_src_ = (int)ns[i];
    }
    src_total_ = st;
    n_bins_ = boff;
    RTR_TRY(src_raw_.ensure(st * 12));
    {
      std::vector<size_t> off(n_src);
      for (int i = 0; i < n_src; ++i) off[i] = srcs_[i].off;
      RTR_TRY(upload_xyz(n_src, src, ns, s_stride, off.data(), st, src_raw_.p));
    }
    RTR_TRY(sx_.ensure(st * 4));
    RTR_TRY(sy_.ensure(st * 4));
    RTR_TRY(sz_.ensure(st * 4));
    RTR_TRY(perm_.ensure(st * 4));
    RTR_TRY(bin_id_.ensure(st * 4));
    RTR_TRY(bin_rank_.ensure(st * 4));
    RTR_TRY(bin_count_.ensure((size_t)n_bins_ * 4));
    RTR_TRY(bin_start_.ensure((size_t)n_bins_ * 4));
  }
  if (n_tgt >= 0 && tgt) {
    tgts_.assign(n_tgt, TargetDesc{});
    size_t off = 0;
    max_tgt_ = 0;
    for (int j = 0; j < n_tgt; ++j) {
      TargetDesc& d = tgts_[j];
      d.raw_off = (uint32_t)off;
      d.n = (uint32_t)nt[j];
      off += nt[j];
      if ((int)nt[j] > max_tgt_) max_tgt_ = (int)nt[j];
    }
    tgt_total_ = tt;
    RTR_TRY(tgt_raw_.ensure(tt * 12));
    {
      std::vector<size_t> off(n_tgt);
      for (int j = 0; j < n_tgt; ++j) off[j] = tgts_[j].raw_off;
      RTR_TRY(upload_xyz(n_tgt, tgt, nt, t_stride, off.data(), tt, tgt_raw_.p));
    }
    RTR_TRY(tgt_sorted_.ensure(tt * 16));
    RTR_TRY(cell_id_.ensure(tt * 4));
    RTR_TRY(cell_rank_.ensure(tt * 4));
    normals_valid_ = false;
  }
  grid_dirty_ = true;
  graph_dirty_ = true;
  results_valid_ = false;
  // a problem set that names clouds which no longer exist is dropped (the
  // caller sets problems again); one that still fits is re-laid out
  for (const Problem& pr : probs_)
    if (pr.src >= srcs_.size() || pr.tgt >= tgts_.size()) {
      probs_.clear();
      break;
    }
  if (!probs_.empty()) return upload_problems();
  return layout_targets();
}

int Engine::set_target_normals(const float* const* nrm, const size_t* stride) {
  RTR_HIP(hipSetDevice(device_));
  RTR_HIP(hipStreamSynchronize(stream_));
  RTR_TRY(tgt_nrm_raw_.ensure(tgt_total_ * 12));
  RTR_TRY(tgt_nrm_sorted_.ensure(tgt_total_ * 16));
  RTR_TRY(ensure_pinned(tgt_total_ * 12));
  for (size_t j = 0; j < tgts_.size(); ++j)
    pack_xyz(nrm[j], tgts_[j].n, stride ? stride[j] : 12, static_cast<float*>(pin_) + 3 * (size_t)tgts_[j].raw_off);
  RTR_HIP(hipMemcpyAsync(tgt_nrm_raw_.p, pin_, tgt_total_ * 12, hipMemcpyHostToDevice, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  normals_valid_ = true;
  grid_dirty_ = true;
  graph_dirty_ = true;
  return RTR_OK;
}

int Engine::set_problems(int n, const int* src_id, const int* tgt_id) {
  probs_.assign(n, Problem{});
  for (int p = 0; p < n; ++p) {
    if (src_id[p] < 0 || src_id[p] >= (int)srcs_.size() || tgt_id[p] < 0 ||
        tgt_id[p] >= (int)tgts_.size()) {
      set_error("problem %d references a missing cloud (src %d, tgt %d)", p, src_id[p], tgt_id[p]);
      probs_.clear();
      return RTR_ERR_INVALID_ARG;
    }
    probs_[p].src = (uint32_t)src_id[p];
    probs_[p].tgt = (uint32_t)tgt_id[p];
  }
  graph_dirty_ = true;
  return layout_problems();
}

int Engine::set_params(const rtr_params& p) {
  if (p.max_iterations < 1 || p.max_iterations > 100000 || !(p.max_correspondence_distance > 0) ||
      (p.mode != RTR_MODE_POINT_TO_POINT && p.mode != RTR_MODE_POINT_TO_PLANE) || p.anderson_memory < 0 ||
      p.anderson_memory > kAndersonMax || p.coarse_stride < 0 || p.coarse_stride > 64 || p.coarse_iterations < 0) {
    set_error("invalid parameters (max_iterations %d, max distance %g, mode %d, anderson_memory %d, coarse %d/%d)",
              p.max_iterations, p.max_correspondence_distance, p.mode, p.anderson_memory, p.coarse_stride,
              p.coarse_iterations);
    return RTR_ERR_INVALID_ARG;
  }
  // the grids depend on the correspondence distance, the cell settings and
  // the mode (normals); the iteration settings alone leave them valid
  if (p.max_correspondence_distance != prm_.max_correspondence_distance || p.cell_size != prm_.cell_size ||
      p.cells_per_point != prm_.cells_per_point || p.mode != prm_.mode)
    grid_dirty_ = true;
  prm_ = p;
  graph_dirty_ = true;
  return upload_problems();
}

size_t Engine::source_size(int prob) const {
  if (prob < 0 || prob >= (int)probs_.size()) return 0;
  return srcs_[probs_[prob].src].n;
}

// Descriptors, budgets and the item list: everything that depends on sizes
// and parameters.  Allocates; never called inside an align.
int Engine::upload_problems() {
  RTR_TRY(layout_targets());
  return layout_problems();
}

// Target and source descriptors: cell budgets and the requested cell edge
// (from the parameters and the tuning).  Uploading them resets the grid
// fields the build writes on the device, so the grids are dirty afterwards.
int Engine::layout_targets() {
  RTR_HIP(hipSetDevice(device_));
  RTR_HIP(hipStreamSynchronize(stream_));
  // target cell budgets and the requested cell edge
  const double rmax = prm_.max_correspondence_distance;
  float h_req = prm_.cell_size > 0 ? prm_.cell_size : 0.f;
  if (h_req <= 0 && rmax < 1e6) h_req = (float)(tune_.cell_factor * rmax);
  const float cpp = prm_.cells_per_point > 0 ? prm_.cells_per_point : tune_.cells_per_point;
  uint32_t coff = 0;
  max_cells_ = 0;
  for (auto& d : tgts_) {
    d.cell_off = coff;
    double b = std::ceil((double)cpp * d.n) + 64.0;
    if (b > 64.0e6) b = 64.0e6;
    d.cell_budget = (uint32_t)b;
    d.h_req = h_req;
    coff += d.cell_budget + 1;
    if ((int)d.cell_budget > max_cells_) max_cells_ = (int)d.cell_budget;
  }
  n_cells_ = coff;
  RTR_TRY(cell_count_.ensure((size_t)n_cells_ * 4));
  RTR_TRY(cell_start_.ensure((size_t)n_cells_ * 4));
  RTR_TRY(cell_dist_.ensure((size_t)n_cells_));
  RTR_TRY(cell_dist_tmp_.ensure((size_t)n_cells_));
  size_t scan_words = scan_tmp_words(n_cells_ > n_bins_ ? n_cells_ : n_bins_) + 16;
  RTR_TRY(scan_tmp_.ensure(scan_words * 4));
  RTR_TRY(tgt_desc_.ensure(tgts_.size() * sizeof(TargetDesc)));
  RTR_TRY(src_desc_.ensure(srcs_.size() * sizeof(SourceDesc)));
  RTR_TRY(bbox_tmp_.ensure((tgts_.size() + srcs_.size()) * 8 * sizeof(uint32_t)));
  if (!tgts_.empty())
    RTR_HIP(hipMemcpy(tgt_desc_.p, tgts_.data(), tgts_.size() * sizeof(TargetDesc), hipMemcpyHostToDevice));
  if (!srcs_.empty())
    RTR_HIP(hipMemcpy(src_desc_.p, srcs_.data(), srcs_.size() * sizeof(SourceDesc), hipMemcpyHostToDevice));
  grid_dirty_ = true;
  graph_dirty_ = true;
  results_valid_ = false;
  return RTR_OK;
}

// Problems, items and per-problem state.  Leaves the grids as they are.
int Engine::layout_problems() {
  RTR_HIP(hipSetDevice(device_));
  RTR_HIP(hipStreamSynchronize(stream_));
  for (const Problem& pr : probs_)
    if (pr.src >= srcs_.size() || pr.tgt >= tgts_.size()) {
      set_error("problem references a missing cloud (src %u of %zu, tgt %u of %zu)", pr.src, srcs_.size(), pr.tgt,
                tgts_.size());
      probs_.clear();
      return RTR_ERR_INVALID_ARG;
    }
  // Items, grouped into one queue per XCD.  All problems of one target go to
  // the same XCD (its L2 then holds that target's grid); targets are placed
  // longest-first on the least-loaded XCD.
  items_.clear();
  pt_total_ = 0;
  const uint32_t ip = (uint32_t)tune_.item_pts;
  std::vector<uint64_t> tload(tgts_.size(), 0);
  for (const Problem& pr : probs_) tload[pr.tgt] += (srcs_[pr.src].n + ip - 1) / ip;
  std::vector<int> order(tgts_.size());
  for (size_t j = 0; j < order.size(); ++j) order[j] = (int)j;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return tload[a] > tload[b] || (tload[a] == tload[b] && a < b); });
  std::vector<int> txcd(tgts_.size(), 0);
  uint64_t xload[kXcds] = {0};
  for (int j : order) {
    int best = 0;
    for (int x = 1; x < kXcds; ++x)
      if (xload[x] < xload[best]) best = x;
    txcd[j] = best;
    xload[best] += tload[j];
  }
  for (size_t p = 0; p < probs_.size(); ++p) {
    Problem& pr = probs_[p];
    const uint32_t n = srcs_[pr.src].n;
    pr.n_items = (n + ip - 1) / ip;
    pr.pt_off = (uint32_t)pt_total_;
    pt_total_ += n;
  }
  for (int x = 0; x < kXcds; ++x) {
    xcd_off_[x] = (uint32_t)items_.size();
    for (size_t p = 0; p < probs_.size(); ++p) {
      Problem& pr = probs_[p];
      if (txcd[pr.tgt] != x) continue;
      pr.item_off = (uint32_t)items_.size();
      const uint32_t n = srcs_[pr.src].n;
      for (uint32_t b = 0; b < n; b += ip) items_.push_back(make_uint2((uint32_t)p, b));
    }
  }
  xcd_off_[kXcds] = (uint32_t)items_.size();
  // fine items for the late passes (the same queues, kFinePts points each):
  // the problems still iterating then are few, and smaller items spread each
  // iteration over more workgroups
  for (Problem& pr : probs_) {
    pr.fitem_off = (uint32_t)items_.size();
    pr.n_fitems = 0;
  }
  for (int x = 0; x < kXcds; ++x) {
    fxcd_off_[x] = (uint32_t)items_.size();
    for (size_t p = 0; p < probs_.size() && tune_.fine_from > 0; ++p) {
      Problem& pr = probs_[p];
      if (txcd[pr.tgt] != x) continue;
      pr.fitem_off = (uint32_t)items_.size();
      const uint32_t n = srcs_[pr.src].n;
      for (uint32_t b = 0; b < n; b += (uint32_t)kFinePts) items_.push_back(make_uint2((uint32_t)p, b));
      pr.n_fitems = (n + kFinePts - 1) / kFinePts;
    }
  }
  fxcd_off_[kXcds] = (uint32_t)items_.size();
  const size_t np = probs_.size();
  RTR_TRY(prob_desc_.ensure(np * sizeof(Problem)));
  RTR_TRY(state_.ensure(np * sizeof(ProblemState)));
  RTR_TRY(guesses_.ensure(np * 16 * sizeof(double)));
  RTR_TRY(done_.ensure(np * 4));
  RTR_TRY(items_buf_.ensure(items_.size() * sizeof(uint2)));
  RTR_TRY(partials_.ensure(items_.size() * NSUM_MAX * sizeof(double)));
  RTR_TRY(counters_.ensure(counter_words() * 4));
  RTR_TRY(out_T_.ensure(16 * sizeof(double)));
  RTR_TRY(hist_.ensure(np * (size_t)(prm_.max_iterations + 1) * 12 * sizeof(float)));
  if (timeline_) RTR_TRY(tline_.ensure(np * (size_t)prm_.max_iterations * 2 * sizeof(uint64_t)));
  if (prm_.anderson_memory > 0) RTR_TRY(aa_.ensure(np * (size_t)kAndersonStride * sizeof(double)));
  // search seeds: one slot per problem and source point (pt_off is 32-bit)
  if (pt_total_ > 0 && pt_total_ <= 0xffffffffull) RTR_TRY(nn_seed_.ensure(pt_total_ * 4));
  if (np) RTR_HIP(hipMemcpy(prob_desc_.p, probs_.data(), np * sizeof(Problem), hipMemcpyHostToDevice));
  if (!items_.empty())
    RTR_HIP(hipMemcpy(items_buf_.p, items_.data(), items_.size() * sizeof(uint2), hipMemcpyHostToDevice));
  graph_dirty_ = true;
  results_valid_ = false;
  return RTR_OK;
}

int Engine::build() {
  RTR_HIP(hipSetDevice(device_));
  if (srcs_.empty() || tgts_.empty()) {
    set_error("no clouds set");
    return RTR_ERR_NOT_READY;
  }
  RTR_TRY(enqueue_build());
  RTR_HIP(hipStreamSynchronize(stream_));
  grid_dirty_ = false;
  results_valid_ = false;
  return RTR_OK;
}

DevEngine Engine::dev_view() const {
  DevEngine e;
  memset(&e, 0, sizeof(e));
  e.tgt_raw = tgt_raw_.as<float>();
  e.tgt_nrm_raw = normals_valid_ ? tgt_nrm_raw_.as<float>() : nullptr;
  e.tgts = tgt_desc_.as<TargetDesc>();
  e.n_tgt = (int)tgts_.size();
  e.tgt_sorted = tgt_sorted_.as<float4>();
  e.tgt_nrm_sorted = tgt_nrm_sorted_.p ? tgt_nrm_sorted_.as<float4>() : nullptr;
  e.cell_count = cell_count_.as<uint32_t>();
  e.cell_start = cell_start_.as<uint32_t>();
  e.cell_id = cell_id_.as<uint32_t>();
  e.cell_rank = cell_rank_.as<uint32_t>();
  e.cell_dist = cell_dist_.as<uint8_t>();
  e.cell_dist_tmp = cell_dist_tmp_.as<uint8_t>();
  e.r_max = prm_.max_correspondence_distance < 3.0e38 ? (float)prm_.max_correspondence_distance : 3.0e38f;
  e.n_cells_total = n_cells_;
  e.src_raw = src_raw_.as<float>();
  e.srcs = src_desc_.as<SourceDesc>();
  e.n_src = (int)srcs_.size();
  e.sx = sx_.as<float>();
  e.sy = sy_.as<float>();
  e.sz = sz_.as<float>();
  e.perm = perm_.as<uint32_t>();
  e.bin_count = bin_count_.as<uint32_t>();
  e.bin_start = bin_start_.as<uint32_t>();
  e.bin_id = bin_id_.as<uint32_t>();
  e.bin_rank = bin_rank_.as<uint32_t>();
  e.n_bins_total = n_bins_;
  e.probs = prob_desc_.as<Problem>();
  e.n_prob = (int)probs_.size();
  e.st = state_.as<ProblemState>();
  e.guesses = guesses_.as<double>();
  e.hist = hist_.as<float>();
  e.tline = timeline_ && tline_.bytes >= probs_.size() * (size_t)prm_.max_iterations * 16 ? tline_.as<uint64_t>()
                                                                                          : nullptr;
  e.aa = prm_.anderson_memory > 0 ? aa_.as<double>() : nullptr;
  e.nn_seed = (pt_total_ > 0 && pt_total_ <= 0xffffffffull) ? nn_seed_.as<uint32_t>() : nullptr;
  e.items = items_buf_.as<uint2>();
  e.n_items = (int)items_.size();
  e.item_pts = tune_.item_pts;
  e.use_tile = tune_.use_tile;
  e.variant = variant();
  e.retire_from = tune_.retire > 0.f ? cu_count_ : 0x7fffffff;
  e.retire_active = (uint32_t)((double)tune_.retire * (double)probs_.size());
  e.partials = partials_.as<double>();
  uint32_t* ctr = counters_.as<uint32_t>();
  // the done count, the error flag and every claim counter on cache lines of
  // their own: spinning workgroups poll the flag while others claim
  e.problems_done = ctr;
  e.error = reinterpret_cast<int*>(ctr + 32);
  e.queries_searched = reinterpret_cast<unsigned long long*>(ctr + 64);
  e.claim = ctr + 96;
  e.claim_stride = claim_stride();
  for (int x = 0; x <= kXcds; ++x) e.xcd_off[x] = xcd_off_[x];
  for (int x = 0; x <= kXcds; ++x) e.fxcd_off[x] = fxcd_off_[x];
  e.fine_from = tune_.fine_from > 0 ? tune_.fine_from : 0x7fffffff;
  e.done = done_.as<uint32_t>();
  e.scan_tmp = scan_tmp_.as<uint32_t>();
  e.bbox_tmp = bbox_tmp_.as<uint32_t>();
  e.prm.max_iter = prm_.max_iterations;
  e.prm.min_corr = prm_.min_correspondences > 0 ? prm_.min_correspondences : 3;
  e.prm.mode = prm_.mode;
  {
    const double r2 = prm_.max_correspondence_distance * prm_.max_correspondence_distance;
    float f;
    if (!(r2 < (double)FLT_MAX))
      f = FLT_MAX;
    else {
      f = (float)r2;
      if ((double)f > r2) f = std::nextafter(f, 0.0f);
    }
    e.prm.r2f = f;
  }
  e.prm.eps_T = prm_.transformation_epsilon;
  e.prm.eps_rot = prm_.rotation_epsilon;
  e.prm.eps_fit = prm_.euclidean_fitness_epsilon;
  e.prm.anderson = prm_.anderson_memory;
  e.prm.coarse_stride = prm_.coarse_stride > 1 && prm_.coarse_iterations > 0 ? prm_.coarse_stride : 1;
  e.prm.coarse_iters = prm_.coarse_iterations;
  return e;
}

int Engine::grid_blocks() const {
  int bpc = icp_blocks_per_cu(prm_.mode, tune_.use_tile, variant());
  if (tune_.blocks_per_cu > 0 && tune_.blocks_per_cu < bpc) bpc = tune_.blocks_per_cu;
  long long b = (long long)cu_count_ * bpc;
  if (b > (long long)items_.size()) b = (long long)items_.size();
  if (b < 1) b = 1;
  return (int)b;
}

int Engine::enqueue_build() {
  DevEngine e = dev_view();
  launch_target_grid(e, max_tgt_, max_cells_, stream_);
  launch_source_sort(e, max_src_, (uint32_t)src_total_, stream_);
  RTR_HIP(hipGetLastError());
  return RTR_OK;
}

int Engine::enqueue_align_body() {
  DevEngine e = dev_view();
  launch_fill_u32(counters_.as<uint32_t>(), 0u, counter_words(), stream_);
  launch_icp(e, grid_blocks(), stream_);
  RTR_HIP(hipGetLastError());
  return RTR_OK;
}

int Engine::ensure_built() {
  if (srcs_.empty() || tgts_.empty() || probs_.empty()) {
    set_error("no clouds or problems set");
    return RTR_ERR_NOT_READY;
  }
  if (prm_.mode == RTR_MODE_POINT_TO_PLANE && !normals_valid_) {
    set_error("point-to-plane needs target normals (set or compute them first)");
    return RTR_ERR_NOT_READY;
  }
  if (normals_valid_) RTR_TRY(tgt_nrm_sorted_.ensure(tgt_total_ * 16));
  if (grid_dirty_ && !rebuild_always_) {
    RTR_TRY(enqueue_build());
    grid_dirty_ = false;
  }
  return RTR_OK;
}

int Engine::align(const double* guesses, rtr_result* results) {
  RTR_HIP(hipSetDevice(device_));
  RTR_TRY(ensure_built());
  const size_t np = probs_.size();
  const size_t gbytes = np * 16 * sizeof(double);
  const size_t sbytes = np * sizeof(ProblemState);
  RTR_TRY(ensure_pinned(gbytes + sbytes + 512));
  double* g = static_cast<double*>(pin_);
  for (size_t p = 0; p < np; ++p)
    for (int k = 0; k < 16; ++k) g[16 * p + k] = guesses ? guesses[16 * p + k] : ((k % 5 == 0) ? 1.0 : 0.0);
  RTR_HIP(hipMemcpyAsync(guesses_.p, g, gbytes, hipMemcpyHostToDevice, stream_));
  memset(&timing_, 0, sizeof(timing_));
  if (profiling_) {
    RTR_HIP(hipEventRecord(ev_[0], stream_));
    if (rebuild_always_) RTR_TRY(enqueue_build());
    RTR_HIP(hipEventRecord(ev_[1], stream_));
    RTR_TRY(enqueue_align_body());
    RTR_HIP(hipEventRecord(ev_[2], stream_));
  } else if (use_graph_) {
    if (graph_dirty_ || !graph_exec_) {
      if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
      if (graph_) (void)hipGraphDestroy(graph_);
      graph_exec_ = nullptr;
      graph_ = nullptr;
      RTR_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      int rc = RTR_OK;
      if (rebuild_always_) rc = enqueue_build();
      if (rc == RTR_OK) rc = enqueue_align_body();
      hipGraph_t gr = nullptr;
      hipError_t ce = hipStreamEndCapture(stream_, &gr);
      if (rc != RTR_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (ce != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(ce));
        return RTR_ERR_DEVICE;
      }
      graph_ = gr;
      RTR_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
      graph_dirty_ = false;
    }
    RTR_HIP(hipGraphLaunch(graph_exec_, stream_));
  } else {
    if (rebuild_always_) RTR_TRY(enqueue_build());
    RTR_TRY(enqueue_align_body());
  }
  ProblemState* hs = reinterpret_cast<ProblemState*>(static_cast<char*>(pin_) + gbytes);
  uint32_t* hc = reinterpret_cast<uint32_t*>(static_cast<char*>(pin_) + gbytes + sbytes);
  RTR_HIP(hipMemcpyAsync(hs, state_.p, sbytes, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipMemcpyAsync(hc, counters_.p, 66 * 4, hipMemcpyDeviceToHost, stream_));  // done count .. searches
  if (profiling_) RTR_HIP(hipEventRecord(ev_[3], stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  if (profiling_) {
    RTR_HIP(hipEventElapsedTime(&timing_.ms_grid, ev_[0], ev_[1]));
    RTR_HIP(hipEventElapsedTime(&timing_.ms_icp, ev_[1], ev_[2]));
    RTR_HIP(hipEventElapsedTime(&timing_.ms_finish, ev_[2], ev_[3]));
    RTR_HIP(hipEventElapsedTime(&timing_.ms_total, ev_[0], ev_[3]));
  }
  timing_.icp_launches = 1;
  {
    unsigned long long ns;
    memcpy(&ns, hc + 64, sizeof(ns));
    timing_.queries_searched = (int64_t)ns;
  }
  int64_t it_total = 0;
  for (size_t p = 0; p < np; ++p) {
    it_total += hs[p].iter;
    if (!results) continue;
    rtr_result& r = results[p];
    memcpy(r.T, hs[p].T, sizeof(r.T));
    r.mse = hs[p].mse;
    r.fitness_rmse = hs[p].mse > 0 ? std::sqrt(hs[p].mse) : 0.0;
    r.n_corr = (int64_t)hs[p].n_corr;
    r.iterations = hs[p].iter;
    r.state = hs[p].state;
    r.converged = (hs[p].state >= RTR_CONV_ITERATIONS && hs[p].state <= RTR_CONV_REL_MSE) ? 1 : 0;
    r.coarse_iterations = hs[p].coarse_its;
  }
  timing_.iterations_total = it_total;
  if (hc[32] != 0) {  // the error flag (word 32)
    set_error("work-queue wait exceeded its bound; results are incomplete");
    return RTR_ERR_DEVICE;
  }
  results_valid_ = true;
  return RTR_OK;
}

int Engine::correspondences(int prob, const double T[16], double max_dist, int32_t* idx, float* d2) {
  RTR_HIP(hipSetDevice(device_));
  if (prob < 0 || prob >= (int)probs_.size()) {
    set_error("problem %d out of range", prob);
    return RTR_ERR_INVALID_ARG;
  }
  RTR_TRY(ensure_built());
  if (grid_dirty_ || rebuild_always_) {
    // grids/sorts are produced inside align for always-rebuild engines
    RTR_TRY(enqueue_build());
    grid_dirty_ = false;
  }
  const size_t n = source_size(prob);
  RTR_TRY(out_i32_.ensure(n * 4));
  RTR_TRY(out_f32_.ensure(n * 4));
  RTR_HIP(hipMemcpyAsync(out_T_.p, T, 16 * sizeof(double), hipMemcpyHostToDevice, stream_));
  const double r2 = max_dist * max_dist;
  float f;
  if (!(r2 < (double)FLT_MAX))
    f = FLT_MAX;
  else {
    f = (float)r2;
    if ((double)f > r2) f = std::nextafter(f, 0.0f);
  }
  launch_correspondences(dev_view(), prob, out_T_.as<double>(), f, out_i32_.as<int32_t>(),
                         out_f32_.as<float>(), stream_);
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipMemcpyAsync(idx, out_i32_.p, n * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipMemcpyAsync(d2, out_f32_.p, n * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  return RTR_OK;
}

int Engine::transformed_source(int prob, float* out, size_t stride) {
  RTR_HIP(hipSetDevice(device_));
  const size_t n = source_size(prob);
  RTR_TRY(out_f32_.ensure(n * 12));
  launch_transform_out(dev_view(), prob, out_f32_.as<float>(), stream_);
  RTR_HIP(hipGetLastError());
  RTR_TRY(ensure_pinned(n * 12));
  RTR_HIP(hipMemcpyAsync(pin_, out_f32_.p, n * 12, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  const float* h = static_cast<const float*>(pin_);
  if (stride == 0) stride = 12;
  char* o = reinterpret_cast<char*>(out);
  for (size_t i = 0; i < n; ++i) memcpy(o + i * stride, h + 3 * i, 12);
  return RTR_OK;
}

int Engine::fitness(double max_range, double* out) {
  RTR_HIP(hipSetDevice(device_));
  if (!results_valid_) {
    set_error("fitness needs a completed align of the current clouds and problems");
    return RTR_ERR_NOT_READY;
  }
  RTR_TRY(out_f32_.ensure(pt_total_ * 4));
  launch_fitness_d2(dev_view(), out_f32_.as<float>(), stream_);
  RTR_HIP(hipGetLastError());
  std::vector<float> d2(pt_total_);
  RTR_HIP(hipMemcpyAsync(d2.data(), out_f32_.p, pt_total_ * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  for (size_t p = 0; p < probs_.size(); ++p) {
    const size_t b = probs_[p].pt_off, n = srcs_[probs_[p].src].n;
    double s = 0;
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
      const float d = d2[b + i];
      if (d != d || d == INFINITY) continue;
      if ((double)d <= max_range) {
        s += d;
        ++m;
      }
    }
    out[p] = m ? s / (double)m : DBL_MAX;
  }
  return RTR_OK;
}

int Engine::judge(int k, int max_pts, double max_curv, double* plane_rmse, double* inlier_frac,
                  double* planar_frac) {
  RTR_HIP(hipSetDevice(device_));
  if (!results_valid_) {
    set_error("judge needs a completed align of the current clouds and problems");
    return RTR_ERR_NOT_READY;
  }
  if (k < 3 || k > 32 || max_pts < 1 || !(max_curv > 0.0)) {
    set_error("judge: k must be in [3, 32], max_pts >= 1, max_curvature > 0 (got %d, %d, %g)", k, max_pts,
              max_curv);
    return RTR_ERR_INVALID_ARG;
  }
  const size_t np = probs_.size();
  if (np == 0) return RTR_OK;
  const size_t words = np * kJudgeSplit * 4;
  RTR_TRY(out_f32_.ensure(words * sizeof(double)));
  DevEngine e = dev_view();
  launch_judge(e, k, max_pts, e.prm.r2f, max_curv, reinterpret_cast<double*>(out_f32_.p), stream_);
  RTR_HIP(hipGetLastError());
  RTR_TRY(ensure_pinned(words * sizeof(double)));
  const double* const h = static_cast<const double*>(pin_);
  RTR_HIP(hipMemcpyAsync(pin_, out_f32_.p, words * sizeof(double), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  for (size_t p = 0; p < np; ++p) {
    double s = 0.0, pc = 0.0, c = 0.0, m = 0.0;
    for (int b = 0; b < kJudgeSplit; ++b) {
      const double* o = &h[4 * (p * kJudgeSplit + b)];
      s += o[0];
      pc += o[1];
      c += o[2];
      m += o[3];
    }
    if (plane_rmse) plane_rmse[p] = pc > 0 ? std::sqrt(s / pc) : DBL_MAX;
    if (inlier_frac) inlier_frac[p] = m > 0 ? c / m : 0.0;
    if (planar_frac) planar_frac[p] = c > 0 ? pc / c : 0.0;
  }
  return RTR_OK;
}

int Engine::compute_target_normals(int k, const float vp[3]) {
  RTR_HIP(hipSetDevice(device_));
  if (tgts_.empty()) {
    set_error("no target cloud");
    return RTR_ERR_NOT_READY;
  }
  if (k < 3 || k > 32) {
    set_error("k must be in [3, 32] (got %d)", k);
    return RTR_ERR_INVALID_ARG;
  }
  RTR_TRY(tgt_nrm_raw_.ensure(tgt_total_ * 12));
  RTR_TRY(tgt_nrm_sorted_.ensure(tgt_total_ * 16));
  // grids over the targets (normals are found in the same cell-sorted layout)
  normals_valid_ = false;
  RTR_TRY(enqueue_build());
  DevEngine e = dev_view();
  e.tgt_nrm_raw = tgt_nrm_raw_.as<float>();
  launch_target_normals(e, k, vp[0], vp[1], vp[2], max_tgt_, stream_);
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipStreamSynchronize(stream_));
  normals_valid_ = true;
  grid_dirty_ = true;
  graph_dirty_ = true;
  return RTR_OK;
}

int Engine::target_normals(int tgt, float* out) {
  RTR_HIP(hipSetDevice(device_));
  if (tgt < 0 || tgt >= (int)tgts_.size()) {
    set_error("target %d out of range", tgt);
    return RTR_ERR_INVALID_ARG;
  }
  if (!normals_valid_) {
    set_error("no target normals have been set or computed");
    return RTR_ERR_NOT_READY;
  }
  RTR_HIP(hipMemcpyAsync(out, tgt_nrm_raw_.as<float>() + 3 * (size_t)tgts_[tgt].raw_off,
                         (size_t)tgts_[tgt].n * 12, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  return RTR_OK;
}

// Float rows (3x4) of the transform each iteration of problem `prob` queried
// with: entry i for iteration i, entry `iterations` the final transform.
int Engine::history(int prob, float* rows12, int cap, int* n_out) {
  RTR_HIP(hipSetDevice(device_));
  if (prob < 0 || prob >= (int)probs_.size()) {
    set_error("problem %d out of range", prob);
    return RTR_ERR_INVALID_ARG;
  }
  ProblemState st;
  RTR_HIP(hipMemcpyAsync(&st, state_.as<ProblemState>() + prob, sizeof(st), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  const int n = st.iter + 1;
  *n_out = n;
  if (n > cap) {
    set_error("history of %d entries, buffer holds %d", n, cap);
    return RTR_ERR_CAPACITY;
  }
  const size_t off = (size_t)prob * (prm_.max_iterations + 1) * 12;
  RTR_HIP(hipMemcpyAsync(rows12, hist_.as<float>() + off, (size_t)n * 12 * sizeof(float), hipMemcpyDeviceToHost,
                         stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  return RTR_OK;
}

// Wall-clock stamps of problem `prob`'s solves in the last align (timeline
// on): 2 per iteration (the solve's start and end), in ticks of tick_ns.
int Engine::timeline(int prob, uint64_t* stamps, int cap, int* n_out, double* tick_ns) {
  RTR_HIP(hipSetDevice(device_));
  if (prob < 0 || prob >= (int)probs_.size()) {
    set_error("problem %d out of range", prob);
    return RTR_ERR_INVALID_ARG;
  }
  if (!timeline_ || tline_.bytes < probs_.size() * (size_t)prm_.max_iterations * 16) {
    set_error("the timeline is off (rtr_batch_set_timeline), or no align has run with it on");
    return RTR_ERR_NOT_READY;
  }
  ProblemState st;
  RTR_HIP(hipMemcpyAsync(&st, state_.as<ProblemState>() + prob, sizeof(st), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  const int n = 2 * st.iter;
  *n_out = n;
  if (n > cap) {
    set_error("timeline of %d stamps, buffer holds %d", n, cap);
    return RTR_ERR_CAPACITY;
  }
  int khz = 0;
  RTR_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
  *tick_ns = khz > 0 ? 1e6 / (double)khz : 0.0;
  RTR_HIP(hipMemcpyAsync(stamps, tline_.as<uint64_t>() + (size_t)prob * prm_.max_iterations * 2,
                         (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  return RTR_OK;
}

// The correspondences of the last iteration the latest align ran for `prob`,
// as its search found them: each searched source point leaves the slot of its
// nearest target point (the next iteration's seed).  The distance is taken
// again here, in the kernel's float arithmetic, from the transform rows that
// iteration queried with.
int Engine::last_correspondences(int prob, int32_t* idx, float* d2) {
  RTR_HIP(hipSetDevice(device_));
  if (prob < 0 || prob >= (int)probs_.size()) {
    set_error("problem %d out of range", prob);
    return RTR_ERR_INVALID_ARG;
  }
  if (!results_valid_) {
    set_error("last correspondences need a completed align of the current clouds and problems");
    return RTR_ERR_NOT_READY;
  }
  if (tune_.use_tile || !nn_seed_.p || pt_total_ > 0xffffffffull) {
    set_error("last correspondences are recorded by the global-walk search only (use_tile 0)");
    return RTR_ERR_NOT_READY;
  }
  ProblemState st;
  RTR_HIP(hipMemcpyAsync(&st, state_.as<ProblemState>() + prob, sizeof(st), hipMemcpyDeviceToHost, stream_));
  const Problem pr = probs_[prob];
  TargetDesc td;
  RTR_HIP(hipMemcpyAsync(&td, tgt_desc_.as<TargetDesc>() + pr.tgt, sizeof(td), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  if (st.iter < 1) {
    set_error("problem %d ran no iteration", prob);
    return RTR_ERR_NOT_READY;
  }
  const SourceDesc& sd = srcs_[pr.src];
  const size_t n = sd.n;
  std::vector<uint32_t> slot(n), perm(n);
  std::vector<float> x(n), y(n), z(n), rows(12);
  std::vector<float4> tp(td.n_valid);
  RTR_HIP(hipMemcpyAsync(slot.data(), nn_seed_.as<uint32_t>() + pr.pt_off, n * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipMemcpyAsync(perm.data(), perm_.as<uint32_t>() + sd.off, n * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipMemcpyAsync(x.data(), sx_.as<float>() + sd.off, n * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipMemcpyAsync(y.data(), sy_.as<float>() + sd.off, n * 4, hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipMemcpyAsync(z.data(), sz_.as<float>() + sd.off, n * 4, hipMemcpyDeviceToHost, stream_));
  if (td.n_valid)
    RTR_HIP(hipMemcpyAsync(tp.data(), tgt_sorted_.as<float4>() + td.raw_off, (size_t)td.n_valid * sizeof(float4),
                           hipMemcpyDeviceToHost, stream_));
  const size_t hoff = ((size_t)prob * (prm_.max_iterations + 1) + (size_t)(st.iter - 1)) * 12;
  RTR_HIP(hipMemcpyAsync(rows.data(), hist_.as<float>() + hoff, 12 * sizeof(float), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  const float r2f = dev_view().prm.r2f;
  for (size_t j = 0; j < n; ++j) {
    const uint32_t o = perm[j];
    idx[o] = -1;
    d2[o] = INFINITY;
    const uint32_t s = slot[j] - td.raw_off;
    if (s >= td.n_valid) continue;
    float qx, qy, qz;
    apply_f32(rows.data(), x[j], y[j], z[j], qx, qy, qz);
    const float4 t = tp[s];
    const float d = d2_f32(qx, qy, qz, t.x, t.y, t.z);
    if (!(d <= r2f)) continue;
    int32_t i;
    memcpy(&i, &t.w, 4);
    idx[o] = i;
    d2[o] = d;
  }
  return RTR_OK;
}

int Engine::grid_info(int tgt, float* cell, int dims[3], float anchor[3]) {
  RTR_HIP(hipSetDevice(device_));
  if (tgt < 0 || tgt >= (int)tgts_.size()) {
    set_error("target %d out of range", tgt);
    return RTR_ERR_INVALID_ARG;
  }
  TargetDesc d;
  RTR_HIP(hipMemcpyAsync(&d, tgt_desc_.as<TargetDesc>() + tgt, sizeof(d), hipMemcpyDeviceToHost, stream_));
  RTR_HIP(hipStreamSynchronize(stream_));
  if (cell) *cell = d.h;
  if (dims) {
    dims[0] = d.dx;
    dims[1] = d.dy;
    dims[2] = d.dz;
  }
  if (anchor) memcpy(anchor, d.anchor, sizeof(d.anchor));
  return RTR_OK;
}

}  // namespace rtr
