// Learned-init registration of many pairs per call, behind the C ABI
// (rtr_learned_*, declared in include/rtr.h): the pose network proposes an
// initial pose per pair, one batched ICP launch refines every proposal, the
// registration judge scores the refined poses, and the pairs it rejects run
// plain multi-start ICP, whose best start replaces them.  This is the
// reference's coarse stage followed by ICP (keypoints, TDF yaw search and
// consensus, RealTimeRobot.cpp:45-105, matching.h:122-222, function.h:35-109;
// then pcl::IterativeClosestPoint, function.h:111-123), rebuilt as one
// batched pipeline.
//
// Host code only, on the library's own C ABI: the network (rtr_prop_*), the
// refine engine and the fallback engine (rtr_batch_*).  The refine engine's
// grid build runs on a host thread while the network runs, each on its own
// HIP stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rtr.h"
#include "guard.hpp"
#include "internal.h"

using rtr::set_error;

namespace {

using clock_type = std::chrono::steady_clock;

double ms_between(clock_type::time_point a, clock_type::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// splitmix64 (Steele, Lea & Flood 2014): the library's fixed pseudo-random
// stream for its default rotation sets
uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return ((double)(splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// n rotations spread over SO(3) (row-major 3x3 each): the identity, then
// greedy farthest-point picks by geodesic angle 2 acos|q.p| from a fixed pool
// of 4096 unit quaternions, uniform on SO(3) (normalised Gaussian 4-vectors,
// Box-Muller over splitmix64).
void default_rotations(int n, std::vector<double>& R) {
  const int pool = 4096;
  std::vector<double> q(4 * (size_t)pool);
  uint64_t s = 20260101ull;
  for (int i = 0; i < pool; ++i) {
    double g[4];
    for (int k = 0; k < 4; k += 2) {
      const double u1 = uniform01(s), u2 = uniform01(s);
      const double r = std::sqrt(-2.0 * std::log(u1));
      g[k] = r * std::cos(6.283185307179586 * u2);
      g[k + 1] = r * std::sin(6.283185307179586 * u2);
    }
    const double nn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3]);
    for (int k = 0; k < 4; ++k) q[4 * (size_t)i + k] = g[k] / nn;
  }
  std::vector<double> chosen = {1.0, 0.0, 0.0, 0.0};
  std::vector<double> dist(pool);
  auto geo = [&](const double* a, const double* b) {
    const double d = std::fabs(a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
    return 2.0 * std::acos(std::min(d, 1.0));
  };
  for (int i = 0; i < pool; ++i) dist[i] = geo(&q[4 * (size_t)i], chosen.data());
  for (int k = 1; k < n; ++k) {
    int best = 0;
    for (int i = 1; i < pool; ++i)
      if (dist[i] > dist[best]) best = i;
    const double* p = &q[4 * (size_t)best];
    chosen.insert(chosen.end(), p, p + 4);
    for (int i = 0; i < pool; ++i) dist[i] = std::min(dist[i], geo(&q[4 * (size_t)i], p));
  }
  R.assign(9 * (size_t)n, 0.0);
  for (int k = 0; k < n; ++k) {
    const double w = chosen[4 * k], x = chosen[4 * k + 1], y = chosen[4 * k + 2], z = chosen[4 * k + 3];
    double* M = &R[9 * (size_t)k];
    M[0] = 1 - 2 * (y * y + z * z), M[1] = 2 * (x * y - w * z), M[2] = 2 * (x * z + w * y);
    M[3] = 2 * (x * y + w * z), M[4] = 1 - 2 * (x * x + z * z), M[5] = 2 * (y * z - w * x);
    M[6] = 2 * (x * z - w * y), M[7] = 2 * (y * z + w * x), M[8] = 1 - 2 * (x * x + y * y);
  }
}

// FP64 mean of a packed cloud (Neumaier-compensated sums)
void centroid(const float* p, size_t n, double c[3]) {
  for (int a = 0; a < 3; ++a) {
    double s = 0.0, comp = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double v = p[3 * i + a], t = s + v;
      comp += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    c[a] = n ? (s + comp) / (double)n : 0.0;
  }
}

// Best start: most correspondences (a failed start counts none), then the
// lowest MSE among the starts within rel_corr of that count (first on ties).
int select_best(const rtr_result* r, int n, double rel_corr = 0.98) {
  double mx = 0.0;
  for (int i = 0; i < n; ++i) mx = std::max(mx, r[i].state != RTR_FAIL_NO_CORRESPONDENCES ? (double)r[i].n_corr : 0.0);
  if (mx <= 0.0) return 0;
  int best = -1;
  for (int i = 0; i < n; ++i) {
    const double c = r[i].state != RTR_FAIL_NO_CORRESPONDENCES ? (double)r[i].n_corr : 0.0;
    if (c >= rel_corr * mx && (best < 0 || r[i].mse < r[best].mse)) best = i;
  }
  return best;
}

}  // namespace

struct rtr_learned {
  int device = 0;
  int max_pairs = 0;
  rtr_learned_config cfg;
  rtr_params plain;                 // the caller's ICP parameters (the fallback's iteration)
  rtr_params refine;                // the same with the refine's acceleration and coarse-to-fine stage
  rtr_prop* prop = nullptr;
  int prop_cap = 0;                 // pairs x hypotheses the proposer holds
  rtr_batch* reg = nullptr;         // refine: one problem per pair, its graph replayed every call
  rtr_batch* fb = nullptr;          // fallback: the rejected pairs x the starts
  std::vector<float> weights;
  std::vector<double> hyp_rots, start_rots;
  // the caller's clouds (shared, as PCL shares its input clouds): the
  // fallback reads the rejected pairs' clouds from them
  int n = 0;
  std::vector<const float*> src, tgt;
  std::vector<size_t> ns, nt;
  std::vector<double> cs, ct;
  // last call
  std::vector<double> poses, rmse, frac, pfrac;
  std::vector<int> rejected;
  rtr_learned_stats stats;
  bool refine_params_active = true;

  ~rtr_learned() {
    if (prop) rtr_prop_destroy(prop);
    if (reg) rtr_batch_destroy(reg);
    if (fb) rtr_batch_destroy(fb);
  }

  int ensure_proposer() {
    const int k = std::max(1, cfg.hypotheses);
    if (prop && prop_cap >= max_pairs * k) return RTR_OK;
    if (prop) rtr_prop_destroy(prop);
    prop = nullptr;
    int rc = rtr_prop_create(&prop, device, max_pairs * k);
    if (rc) return rc;
    prop_cap = max_pairs * k;
    if (!weights.empty()) RTR_TRY(rtr_prop_set_weights(prop, weights.data()));
    if (n > 0) RTR_TRY(rtr_prop_set_clouds(prop, n, src.data(), ns.data(), tgt.data(), nt.data()));
    return RTR_OK;
  }

  int use_refine_params(bool on) {
    if (on == refine_params_active) return RTR_OK;
    RTR_TRY(rtr_batch_set_params(reg, on ? &refine : &plain));
    refine_params_active = on;
    return RTR_OK;
  }

  // refine from `poses` (n x 16), judge, fallback; results (n), inits (n x 16,
  // nullable), rejected flags (n, nullable)
  int refine_judge_fallback(rtr_result* results, double* inits, int* rejected_out, clock_type::time_point t1) {
    RTR_TRY(use_refine_params(true));
    RTR_TRY(rtr_batch_align(reg, poses.data(), results));
    const auto t2 = clock_type::now();
    rmse.assign(n, 0.0);
    frac.assign(n, 0.0);
    pfrac.assign(n, 0.0);
    RTR_TRY(rtr_batch_judge(reg, cfg.judge_k, cfg.judge_points, cfg.max_curvature, rmse.data(), frac.data(),
                            pfrac.data()));
    const auto t3 = clock_type::now();
    rejected.clear();
    for (int i = 0; i < n; ++i)
      if (rmse[i] > cfg.rmse_threshold || frac[i] < cfg.min_inlier_frac) rejected.push_back(i);
    if (inits) std::memcpy(inits, poses.data(), sizeof(double) * 16 * (size_t)n);
    if (rejected_out)
      for (int i = 0; i < n; ++i) rejected_out[i] = 0;
    const int nf = (int)rejected.size(), K = cfg.n_starts;
    if (nf > 0) {
      // the fallback engine takes the rejected pairs' clouds and the refine
      // engine's cell settings
      float cf = 0.f, cpp = 0.f, fcf = 0.f, fcpp = 0.f;
      RTR_TRY(rtr_batch_get_tuning(reg, &cf, &cpp, nullptr, nullptr, nullptr));
      RTR_TRY(rtr_batch_get_tuning(fb, &fcf, &fcpp, nullptr, nullptr, nullptr));
      if (cf != fcf || cpp != fcpp) RTR_TRY(rtr_batch_set_tuning(fb, cf, cpp, 0, -1, -1));
      std::vector<const float*> fs(nf), ft(nf);
      std::vector<size_t> fns(nf), fnt(nf);
      for (int k = 0; k < nf; ++k) {
        const int i = rejected[k];
        fs[k] = src[i];
        ft[k] = tgt[i];
        fns[k] = ns[i];
        fnt[k] = nt[i];
      }
      RTR_TRY(rtr_batch_set_clouds(fb, nf, fs.data(), fns.data(), nf, ft.data(), fnt.data()));
      std::vector<int> ids((size_t)nf * K);
      for (int k = 0; k < nf; ++k)
        for (int j = 0; j < K; ++j) ids[(size_t)k * K + j] = k;
      RTR_TRY(rtr_batch_set_problems(fb, nf * K, ids.data(), ids.data()));
      // starts: the source turned about its centroid by each start rotation,
      // centroid moved onto the target's: T = [R | ct - R cs]
      std::vector<double> g((size_t)nf * K * 16, 0.0);
      for (int k = 0; k < nf; ++k) {
        const int i = rejected[k];
        const double* c_s = &cs[3 * (size_t)i];
        const double* c_t = &ct[3 * (size_t)i];
        for (int j = 0; j < K; ++j) {
          const double* R = &start_rots[9 * (size_t)j];
          double* T = &g[16 * ((size_t)k * K + j)];
          for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) T[c * 4 + r] = R[3 * r + c];
            double Rc = R[3 * r] * c_s[0];
            Rc = Rc + R[3 * r + 1] * c_s[1];
            Rc = Rc + R[3 * r + 2] * c_s[2];
            T[12 + r] = c_t[r] - Rc;
          }
          T[15] = 1.0;
        }
      }
      std::vector<rtr_result> r8((size_t)nf * K);
      RTR_TRY(rtr_batch_align(fb, g.data(), r8.data()));
      for (int k = 0; k < nf; ++k) {
        const int i = rejected[k];
        const int b = select_best(&r8[(size_t)k * K], K);
        results[i] = r8[(size_t)k * K + b];
        if (inits) std::memcpy(inits + 16 * (size_t)i, &g[16 * ((size_t)k * K + b)], sizeof(double) * 16);
        if (rejected_out) rejected_out[i] = 1;
      }
    }
    const auto t4 = clock_type::now();
    stats.n_pairs = n;
    stats.n_fallback = nf;
    stats.ms_refine = (float)ms_between(t1, t2);
    stats.ms_judge = (float)ms_between(t2, t3);
    stats.ms_fallback = (float)ms_between(t3, t4);
    return RTR_OK;
  }
};

extern "C" {

int rtr_learned_config_default(rtr_learned_config* c) {
  if (!c) return RTR_ERR_INVALID_ARG;
  std::memset(c, 0, sizeof(*c));
  c->hypotheses = 8;
  c->n_starts = 8;
  c->tau = 0.1f;
  c->rmse_threshold = 2e-3;
  c->min_inlier_frac = 0.8;
  c->judge_k = 20;
  c->judge_points = 1024;
  c->max_curvature = 0.02;
  c->refine_anderson = 5;
  c->refine_coarse_stride = 4;
  c->refine_coarse_iterations = 25;
  c->fallback_anderson = 0;
  c->fallback_item_pts = 256;
  c->fallback_blocks_per_cu = 1;
  return RTR_OK;
}

int rtr_learned_create(rtr_learned** out, int device, int max_pairs, const rtr_params* icp,
                       const rtr_learned_config* cfg) {
  RTR_GUARD_BEGIN
  if (!out || max_pairs < 1) {
    set_error("rtr_learned_create: null handle pointer or max_pairs < 1");
    return RTR_ERR_INVALID_ARG;
  }
  *out = nullptr;
  rtr_learned_config c;
  if (cfg)
    c = *cfg;
  else
    rtr_learned_config_default(&c);
  if (c.hypotheses < 1 || c.hypotheses > 64 || c.n_starts < 1 || c.n_starts > 64 || !(c.tau > 0.f) ||
      c.refine_anderson < 0 || c.refine_anderson > 8 || c.fallback_anderson < 0 || c.fallback_anderson > 8 ||
      c.fallback_item_pts < 0 || c.fallback_item_pts % 256 != 0) {
    set_error("rtr_learned_create: configuration out of range (hypotheses and starts 1..64, tau > 0, Anderson "
              "memories 0..8, fallback item size a multiple of 256)");
    return RTR_ERR_INVALID_ARG;
  }
  rtr_learned* h = new rtr_learned();
  h->device = device;
  h->max_pairs = max_pairs;
  h->cfg = c;
  if (icp)
    h->plain = *icp;
  else
    rtr_params_default(&h->plain);
  h->plain.anderson_memory = 0;
  h->plain.coarse_stride = 0;
  h->plain.coarse_iterations = 0;
  h->refine = h->plain;
  h->refine.anderson_memory = c.refine_anderson;
  h->refine.coarse_stride = c.refine_coarse_stride;
  h->refine.coarse_iterations = c.refine_coarse_iterations;
  rtr_params fbp = h->plain;
  fbp.anderson_memory = c.fallback_anderson;
  default_rotations(c.hypotheses, h->hyp_rots);
  default_rotations(c.n_starts, h->start_rots);
  std::memset(&h->stats, 0, sizeof(h->stats));
  int rc = h->ensure_proposer();
  if (!rc) rc = rtr_batch_create(&h->reg, device);
  if (!rc) rc = rtr_batch_set_params(h->reg, &h->refine);
  if (!rc) rc = rtr_batch_set_rebuild(h->reg, 0);  // one grid build per call, beside the proposal
  if (!rc) rc = rtr_batch_create(&h->fb, device);
  if (!rc) rc = rtr_batch_set_params(h->fb, &fbp);
  // the fallback launch (a few pairs x the starts) is latency-bound: small
  // work items, few workgroups per CU, launched directly (its layout changes
  // every call, so a captured graph would be replayed once)
  if (!rc) rc = rtr_batch_set_tuning(h->fb, 0.f, 0.f, c.fallback_item_pts, c.fallback_blocks_per_cu, -1);
  if (!rc) rc = rtr_batch_set_graph(h->fb, 0);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_learned_destroy(rtr_learned* h) {
  RTR_GUARD_BEGIN
  delete h;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_learned_set_weights(rtr_learned* h, const float* w) {
  RTR_GUARD_BEGIN
  if (!h || !w) return RTR_ERR_INVALID_ARG;
  size_t np = 0;
  RTR_TRY(rtr_prop_num_params(&np));
  h->weights.assign(w, w + np);
  return rtr_prop_set_weights(h->prop, w);
  RTR_GUARD_END
}

int rtr_learned_load_weights(rtr_learned* h, const char* path) {
  RTR_GUARD_BEGIN
  if (!h || !path) return RTR_ERR_INVALID_ARG;
  size_t np = 0;
  RTR_TRY(rtr_prop_num_params(&np));
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_error("cannot open %s", path);
    return RTR_ERR_IO;
  }
  std::vector<float> w(np + 1);
  const size_t got = std::fread(w.data(), sizeof(float), np + 1, f);
  std::fclose(f);
  if (got != np) {
    set_error("%s holds %zu%s float32 values, the network has %zu parameters", path, got,
              got > np ? " or more" : "", np);
    return RTR_ERR_FORMAT;
  }
  return rtr_learned_set_weights(h, w.data());
  RTR_GUARD_END
}

int rtr_learned_set_rotations(rtr_learned* h, const double* hyp, int n_hyp, const double* starts, int n_starts) {
  RTR_GUARD_BEGIN
  if (!h || n_hyp < 1 || n_hyp > 64 || n_starts < 1 || n_starts > 64) {
    set_error("rtr_learned_set_rotations: 1..64 hypotheses and starts");
    return RTR_ERR_INVALID_ARG;
  }
  h->cfg.hypotheses = n_hyp;
  h->cfg.n_starts = n_starts;
  if (hyp)
    h->hyp_rots.assign(hyp, hyp + 9 * (size_t)n_hyp);
  else
    default_rotations(n_hyp, h->hyp_rots);
  if (starts)
    h->start_rots.assign(starts, starts + 9 * (size_t)n_starts);
  else
    default_rotations(n_starts, h->start_rots);
  return h->ensure_proposer();
  RTR_GUARD_END
}

int rtr_learned_get_rotations(rtr_learned* h, double* hyp, double* starts) {
  if (!h) return RTR_ERR_INVALID_ARG;
  if (hyp) std::memcpy(hyp, h->hyp_rots.data(), sizeof(double) * h->hyp_rots.size());
  if (starts) std::memcpy(starts, h->start_rots.data(), sizeof(double) * h->start_rots.size());
  return RTR_OK;
}

int rtr_learned_set_clouds(rtr_learned* h, int n, const float* const* src, const size_t* ns,
                           const float* const* tgt, const size_t* nt, const double* src_centroids,
                           const double* tgt_centroids) {
  RTR_GUARD_BEGIN
  if (!h || n < 1 || !src || !ns || !tgt || !nt) return RTR_ERR_INVALID_ARG;
  if (n > h->max_pairs) {
    set_error("rtr_learned_set_clouds: %d pairs, the handle was created for at most %d", n, h->max_pairs);
    return RTR_ERR_CAPACITY;
  }
  for (int i = 0; i < n; ++i)
    if (!src[i] || !tgt[i] || ns[i] == 0 || nt[i] == 0) {
      set_error("rtr_learned_set_clouds: pair %d has an empty or null cloud", i);
      return RTR_ERR_EMPTY_INPUT;
    }
  h->n = 0;
  // one upload from host memory (the refine engine's); the network takes
  // its copy of the clouds on the device
  RTR_TRY(rtr_batch_set_clouds(h->reg, n, src, ns, n, tgt, nt));
  {
    const float *dsrc = nullptr, *dtgt = nullptr;
    RTR_TRY(rtr_batch_device_clouds(h->reg, &dsrc, &dtgt));
    RTR_TRY(rtr_prop_set_device_clouds(h->prop, n, dsrc, ns, dtgt, nt));
  }
  std::vector<int> ids(n);
  for (int i = 0; i < n; ++i) ids[i] = i;
  RTR_TRY(rtr_batch_set_problems(h->reg, n, ids.data(), ids.data()));
  h->src.assign(src, src + n);
  h->tgt.assign(tgt, tgt + n);
  h->ns.assign(ns, ns + n);
  h->nt.assign(nt, nt + n);
  h->cs.resize(3 * (size_t)n);
  h->ct.resize(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    if (src_centroids)
      std::memcpy(&h->cs[3 * (size_t)i], src_centroids + 3 * (size_t)i, 3 * sizeof(double));
    else
      centroid(src[i], ns[i], &h->cs[3 * (size_t)i]);
    if (tgt_centroids)
      std::memcpy(&h->ct[3 * (size_t)i], tgt_centroids + 3 * (size_t)i, 3 * sizeof(double));
    else
      centroid(tgt[i], nt[i], &h->ct[3 * (size_t)i]);
  }
  h->n = n;
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_learned_align(rtr_learned* h, rtr_result* results, double* inits, int* rejected) {
  RTR_GUARD_BEGIN
  if (!h || !results) return RTR_ERR_INVALID_ARG;
  if (h->n < 1) {
    set_error("rtr_learned_align: no clouds set");
    return RTR_ERR_NOT_READY;
  }
  if (h->weights.empty()) {
    set_error("rtr_learned_align: no network weights set (rtr_learned_set_weights / rtr_learned_load_weights)");
    return RTR_ERR_NOT_READY;
  }
  const auto t0 = clock_type::now();
  // the refine engine's grids are built on its stream while the network runs
  // on the proposer's
  int rc_build = RTR_OK;
  std::string build_err;
  std::thread build([&] {
    rc_build = rtr_batch_build(h->reg);
    if (rc_build) build_err = rtr_last_error();
  });
  h->poses.assign(16 * (size_t)h->n, 0.0);
  int rc;
  if (h->cfg.hypotheses > 1)
    rc = rtr_prop_propose_multi(h->prop, h->cfg.hypotheses, h->hyp_rots.data(), h->cfg.tau, nullptr, nullptr,
                                h->poses.data());
  else
    rc = rtr_prop_propose(h->prop, h->poses.data());
  std::string prop_err = rc ? rtr_last_error() : "";
  const auto tb = clock_type::now();
  build.join();
  const auto t1 = clock_type::now();
  if (rc) {
    set_error("%s", prop_err.c_str());
    return rc;
  }
  if (rc_build) {
    set_error("%s", build_err.c_str());
    return rc_build;
  }
  float dev_ms = 0.f;
  rtr_prop_last_ms(h->prop, &dev_ms);
  RTR_TRY(h->refine_judge_fallback(results, inits, rejected, t1));
  h->stats.ms_propose = (float)ms_between(t0, tb);
  h->stats.ms_build_wait = (float)ms_between(tb, t1);
  h->stats.ms_propose_device = dev_ms;
  h->stats.ms_total = (float)ms_between(t0, clock_type::now());
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_learned_align_from(rtr_learned* h, const double* proposals, rtr_result* results, double* inits,
                           int* rejected) {
  RTR_GUARD_BEGIN
  if (!h || !proposals || !results) return RTR_ERR_INVALID_ARG;
  if (h->n < 1) {
    set_error("rtr_learned_align_from: no clouds set");
    return RTR_ERR_NOT_READY;
  }
  const auto t0 = clock_type::now();
  RTR_TRY(rtr_batch_build(h->reg));
  h->poses.assign(proposals, proposals + 16 * (size_t)h->n);
  const auto t1 = clock_type::now();
  RTR_TRY(h->refine_judge_fallback(results, inits, rejected, t1));
  h->stats.ms_propose = 0.f;
  h->stats.ms_build_wait = (float)ms_between(t0, t1);
  h->stats.ms_propose_device = 0.f;
  h->stats.ms_total = (float)ms_between(t0, clock_type::now());
  return RTR_OK;
  RTR_GUARD_END
}

int rtr_learned_get_stats(rtr_learned* h, rtr_learned_stats* s) {
  if (!h || !s) return RTR_ERR_INVALID_ARG;
  *s = h->stats;
  return RTR_OK;
}

int rtr_learned_get_proposals(rtr_learned* h, double* poses) {
  if (!h || !poses) return RTR_ERR_INVALID_ARG;
  if (h->poses.size() != 16 * (size_t)h->n || h->n < 1) {
    set_error("no proposals yet");
    return RTR_ERR_NOT_READY;
  }
  std::memcpy(poses, h->poses.data(), sizeof(double) * h->poses.size());
  return RTR_OK;
}

int rtr_learned_get_judge(rtr_learned* h, double* plane_rmse, double* inlier_frac, double* planar_frac) {
  if (!h) return RTR_ERR_INVALID_ARG;
  if (h->rmse.size() != (size_t)h->n || h->n < 1) {
    set_error("no judged call yet");
    return RTR_ERR_NOT_READY;
  }
  if (plane_rmse) std::memcpy(plane_rmse, h->rmse.data(), sizeof(double) * h->n);
  if (inlier_frac) std::memcpy(inlier_frac, h->frac.data(), sizeof(double) * h->n);
  if (planar_frac) std::memcpy(planar_frac, h->pfrac.data(), sizeof(double) * h->n);
  return RTR_OK;
}

int rtr_learned_use_plain_params(rtr_learned* h, int plain) {
  RTR_GUARD_BEGIN
  if (!h) return RTR_ERR_INVALID_ARG;
  return h->use_refine_params(plain == 0);
  RTR_GUARD_END
}

int rtr_learned_batches(rtr_learned* h, rtr_batch** refine, rtr_batch** fallback, rtr_prop** proposer) {
  if (!h) return RTR_ERR_INVALID_ARG;
  if (refine) *refine = h->reg;
  if (fallback) *fallback = h->fb;
  if (proposer) *proposer = h->prop;
  return RTR_OK;
}

}  // extern "C"
