// Persistent batched ICP over a work queue, plus the per-point output
// kernels (correspondences, transformed cloud, fitness) and k-NN normals.
//
// Work queue.  The source points of every problem are split into items of
// item_pts consecutive (spatially sorted) points.  Workgroups loop over passes
// (one pass = one ICP iteration of every unfinished problem); in each pass they
// claim items from that pass's atomic counter.  An item of problem p in pass k
// waits (bounded) until p has completed k iterations, computes its points'
// correspondences and FP64 moments (exact anchor-relative differences, one
// rounding per product; wave totals from DPP row sums), and the workgroup that
// finishes p's last item of the pass adds the item moments in FP64, solves
// the increment, applies the convergence rules and publishes iteration k+1.
// Problems therefore advance independently; no grid-wide barrier is needed,
// and every wait is bounded (a timeout sets e.error and the waiters bail out).
// Each pass has one queue per XCD (problems of one target share an XCD, whose
// L2 then holds that target's grid); a workgroup claims from its own XCD's
// queue and steals from the others when it is empty, and a claim that lands
// on a finished problem skips that problem's remaining items of the pass.
//
// Two search variants: the throughput variant (4 gathers in flight per lane,
// 3 waves/SIMD) classifies an item's queries first and walks them from a
// compacted list grouped by distance-field class; the latency variant (8
// gathers in flight, 2 waves/SIMD) walks each 64-query round directly.  With
// Anderson acceleration on, the solving workgroup also extrapolates the next
// iterate (geom.hpp).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <type_traits>

#include "engine_dev.hpp"
#include "geom.hpp"
#include "wave.hpp"

namespace rtr {

constexpr uint32_t kFinishedBit = 1u << 30;
constexpr int kMaxSpins = 1 << 22;

__device__ __forceinline__ uint32_t load_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ GridView grid_of(const DevEngine& e, const TargetDesc& td) {
  GridView g;
  g.pts = e.tgt_sorted;
  g.cell_start = e.cell_start + td.cell_off;
  g.dx = td.dx;
  g.dy = td.dy;
  g.dz = td.dz;
  g.n_pts = td.n_valid;
  g.pt_lo = td.raw_off;
  g.ox = td.ox;
  g.oy = td.oy;
  g.oz = td.oz;
  g.h = td.h;
  g.inv_h = td.inv_h;
  g.margin = td.margin;
  return g;
}

// Work items of problem pr in a pass: the coarse ones (item_pts points each)
// before e.fine_from, the fine ones (kFinePts) from then on.  The items a
// problem completes up to and including pass k (its done count then).
__device__ __forceinline__ bool fine_pass(const DevEngine& e, int pass) { return pass >= e.fine_from; }
__device__ __forceinline__ uint32_t items_of(const Problem& pr, bool fine) { return fine ? pr.n_fitems : pr.n_items; }
__device__ __forceinline__ uint32_t first_item_of(const Problem& pr, bool fine) {
  return fine ? pr.fitem_off : pr.item_off;
}
__device__ __forceinline__ uint32_t done_through(const DevEngine& e, const Problem& pr, int pass) {
  if (pass < e.fine_from) return (uint32_t)(pass + 1) * pr.n_items;
  return (uint32_t)e.fine_from * pr.n_items + (uint32_t)(pass + 1 - e.fine_from) * pr.n_fitems;
}

__global__ void k_icp_init(DevEngine e) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= e.n_prob) return;
  ProblemState& S = e.st[p];
  for (int k = 0; k < 16; ++k) S.T[k] = e.guesses ? e.guesses[16 * (size_t)p + k] : ((k % 5 == 0) ? 1.0 : 0.0);
  S.prev_mse = DBL_MAX;
  S.mse = 0.0;
  S.n_corr = 0.0;
  S.iter = 0;
  S.state = RTR_NOT_CONVERGED;
  S.finished = 0;
  S.stage = e.prm.coarse_stride > 1 ? 1 : 0;
  S.coarse_its = 0;
  e.done[p] = 0;
  if (e.hist) tf_rows_f32(S.T, e.hist + (size_t)p * (e.prm.max_iter + 1) * 12);
  if (e.aa) anderson_reset(e.aa + (size_t)p * kAndersonStride);
  uint32_t* word = (uint32_t*)&S.gen;
  if (e.probs[p].n_items == 0) {
    S.state = RTR_FAIL_NO_CORRESPONDENCES;
    S.finished = 1;
    *word = kFinishedBit;
    atomicAdd(e.problems_done, 1u);
  } else {
    *word = 0u;
  }
}

// ---------------------------------------------------------------------------
// LDS-staged candidate tiles
//
// Every wave takes 64 consecutive (spatially sorted) source points.  The
// union of their 3x3x3 cell blocks is a small box of cells; its x-rows are
// contiguous ranges of the cell-sorted target array, so the wave copies them
// into LDS with coalesced 16-byte loads, together with the LDS offset of every
// cell of the box.  Each lane then scans its own nine rows from LDS.  Lanes
// whose answer is not proven inside the block continue over outer rings
// straight from global memory; a wave whose box does not fit the tile walks
// its blocks from global memory too.  Either way the candidate set, and so
// the result, is exactly that of the plain ring walk.
// ---------------------------------------------------------------------------
constexpr int kTilePts = 512;    // staged points per wave (8 KB)
constexpr int kTileCells = 640;  // staged cell offsets per wave
constexpr int kTileRows = 64;    // staged x-rows per wave (one per lane)

struct WaveTile {
  float4 pts[kTilePts];
  uint16_t cell[kTileCells];     // LDS offset of each box cell, plus each row's end
  uint32_t row_a[kTileRows];     // first cell-sorted position of each row
  uint32_t row_off[kTileRows];   // LDS offset of each row
};

// Placeholder with the same members for the kernel variant without tiles
// (its LDS then only holds the reductions, which raises occupancy).
struct WaveTileNone {
  float4 pts[1];
  uint16_t cell[2];
  uint32_t row_a[1];
  uint32_t row_off[1];
};

// Queries of an item are classified, then walked from a compacted list in
// chunks of at most kChunk (the throughput search variant).
constexpr int kChunk = 1024;
constexpr int kQueryClasses = 5;  // distance field 0 (own cell occupied), 1 (block), 2, 3, 4+ (rings)

template <int NS, class TileT>
struct IcpShared {
  TileT tile[kWaves];
  uint8_t qcls[kChunk];   // per query of the chunk: 0 no correspondence possible, 1 block walk, 2.. rings only
  uint8_t qd[kChunk];     // its cell's distance-field value
  uint16_t qlist[kChunk]; // the chunk's queries to walk by class, each class in query order
  uint32_t qseed[kChunk]; // its search seed: the last iteration's nearest-neighbour slot (or ~0)
  uint32_t ccount[kWaves][kQueryClasses];  // queries per class in each wave's quarter of the chunk
  int nlist;
  int lnext;              // the walk list's next unclaimed block of 64 entries
  double red[kWaves][NS];
  double dred[kBlock / NSUM_MAX][NSUM_MAX];
  double ws[kAndersonWs];
  double Tinc[16];
  double Tn[16];
  double Ta[16];
  double Tq[16];
  double aa[kAndersonStride];
  float m[12];
  float anchor[3];
  uint32_t n_search;      // queries this workgroup searched
  int stride;             // this iteration's source stride (coarse stage: > 1)
  int stage_reset;        // the solve moved the problem to all source points
  int item;
  int run;
  int stop;
  int last;
  int fin;
};

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  *total = __shfl(inc, 63, 64);
  return inc - v;
}

// "No point yet": above every real key, with a squared distance of +inf.
constexpr uint64_t kNoKey = 0x7f800000ffffffffull;


// Nearest-neighbour state: packed (d2, index) key and where the winner lives
// (an LDS tile index, or a cell-sorted slot with the top bit set).
template <int D>
struct GlobalVis {
  const float4* pts;
  float qx, qy, qz;
  uint64_t key;
  uint32_t where;
  __device__ __forceinline__ void init(const float4* p, float x, float y, float z) {
    pts = p;
    qx = x;
    qy = y;
    qz = z;
    key = kNoKey;
    where = 0u;
  }
  __device__ __forceinline__ void take(const float4& t, uint32_t j) {
    const uint64_t k = nn_key(d2_f32(qx, qy, qz, t.x, t.y, t.z), __float_as_int(t.w));
    if (k < key) {
      key = k;
      where = j | 0x80000000u;
    }
  }
  // D (4 or 8) independent loads in flight per step: the gathers are
  // latency-bound, and the deeper variant needs the registers of a kernel
  // that runs 2 waves per SIMD
  // (a walking pointer: one address per group, the loads at immediate offsets)
  __device__ __forceinline__ void scan(uint32_t b, uint32_t e) {
    uint32_t j = b;
    const float4* q = pts + b;
    if (D == 8) {
      for (; j + 8 <= e; j += 8, q += 8) {
        const float4 t0 = q[0], t1 = q[1], t2 = q[2], t3 = q[3];
        const float4 t4 = q[4], t5 = q[5], t6 = q[6], t7 = q[7];
        take(t0, j);
        take(t1, j + 1);
        take(t2, j + 2);
        take(t3, j + 3);
        take(t4, j + 4);
        take(t5, j + 5);
        take(t6, j + 6);
        take(t7, j + 7);
      }
    }
    for (; j + 4 <= e; j += 4, q += 4) {
      const float4 t0 = q[0], t1 = q[1], t2 = q[2], t3 = q[3];
      take(t0, j);
      take(t1, j + 1);
      take(t2, j + 2);
      take(t3, j + 3);
    }
    for (; j < e; ++j, ++q) take(*q, j);
  }
  __device__ __forceinline__ float bound() const { return key_d2(key); }
  // Start from target point s (cell-sorted slot): its distance is an upper
  // bound on the nearest one, and the walk's pruning (strict comparisons on
  // the packed key) still finds the exact winner, ties to the lowest index.
  __device__ __forceinline__ void seed(uint32_t s) {
    const float4 t = pts[s];
    key = nn_key(d2_f32(qx, qy, qz, t.x, t.y, t.z), __float_as_int(t.w));
    where = s | 0x80000000u;
  }
};

// The 3x3x3 block straight from global memory.  Each x-row of three cells is
// first cut to the cells whose box lies within the bound known before the walk
// (a seed's distance, else r2f): a row is dropped whole, or loses its end
// cells.  The row bounds of what is left are loaded together before any row is
// scanned; rows are then scanned own row first, each skipped when its box lies
// beyond the best so far (or beyond r2f).
template <int D>
__device__ __forceinline__ void walk_block_global(const GridView& g, float qx, float qy, float qz, int cx, int cy,
                                                  int cz, float r2f, GlobalVis<D>& vis) {
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < g.dx - 1 ? cx + 1 : g.dx - 1;
  // separable squared gaps: the x extent of the whole row and of its end
  // cells, the y and z offsets of the side rows
  const float ax = slab_gap(qx, g.ox, g.h, g.margin, x0, x1);
  const float axl = slab_gap(qx, g.ox, g.h, g.margin, x0, x0);
  const float axr = slab_gap(qx, g.ox, g.h, g.margin, x1, x1);
  const float ay0 = slab_gap(qy, g.oy, g.h, g.margin, cy - 1, cy - 1);
  const float ay1 = slab_gap(qy, g.oy, g.h, g.margin, cy + 1, cy + 1);
  const float az0 = slab_gap(qz, g.oz, g.h, g.margin, cz - 1, cz - 1);
  const float az1 = slab_gap(qz, g.oz, g.h, g.margin, cz + 1, cz + 1);
  const float gx2 = ax * ax, gxl2 = axl * axl, gxr2 = axr * axr;
  const float gy2[3] = {ay0 * ay0, 0.f, ay1 * ay1};
  const float gz2[3] = {az0 * az0, 0.f, az1 * az1};
  const float b0 = vis.bound() < r2f ? vis.bound() : r2f;
  uint32_t rb[9], re[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int dy = kBlockRowOrder[k][0], dz = kBlockRowOrder[k][1];
    const int y = cy + dy, z = cz + dz;
    rb[k] = re[k] = 0u;
    if (z >= 0 && z < g.dz && y >= 0 && y < g.dy) {
      if (!(lb2_of(gx2, gy2[dy + 1], gz2[dz + 1]) > b0)) {
        // the own cell is the row's nearest (the query lies in its slab, or
        // beyond the grid's edge next to it), so what is kept is contiguous
        const int xa = lb2_of(gxl2, gy2[dy + 1], gz2[dz + 1]) > b0 ? cx : x0;
        const int xb = lb2_of(gxr2, gy2[dy + 1], gz2[dz + 1]) > b0 ? cx : x1;
        const uint32_t row = ((uint32_t)z * g.dy + y) * g.dx;
        rb[k] = g.cell_start[row + xa];
        re[k] = g.cell_start[row + xb + 1];
      }
    }
  }
  vis.scan(rb[0], re[0]);
#pragma unroll
  for (int k = 1; k < 9; ++k) {
    if (rb[k] == re[k]) continue;
    const float lb2 = lb2_of(gx2, gy2[kBlockRowOrder[k][0] + 1], gz2[kBlockRowOrder[k][1] + 1]);
    if (row_skip(lb2, vis.bound(), r2f)) continue;
    vis.scan(rb[k], re[k]);
  }
}

template <class TileT>
__device__ __forceinline__ int tile_row_of(const TileT& tl, int nrows, uint32_t t) {
  int lo = 0, hi = nrows - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tl.row_off[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int MODE, int NS, class TileT>
__device__ void solve_problem(const DevEngine& e, uint32_t p, const Problem& pr,
                              const TargetDesc& td, int pass, IcpShared<NS, TileT>& sh) {
  const int tid = threadIdx.x;
  const int v = tid % NSUM_MAX, grp = tid / NSUM_MAX;
  // the pose and the acceleration state are staged in LDS by the whole block
  // (one load per lane), so the single-thread FP64 work below reads no
  // global memory
  double* gaa = e.aa ? e.aa + (size_t)p * kAndersonStride : nullptr;
  if (tid < 16) sh.Tq[tid] = e.st[p].T[tid];
  if (gaa && tid < kAndersonStride) sh.aa[tid] = gaa[tid];
  double acc = 0.0;
  const bool fine = fine_pass(e, pass);
  const uint32_t i0 = first_item_of(pr, fine), ni = items_of(pr, fine);
  if (v < NS) {
    for (uint32_t i = grp; i < ni; i += kBlock / NSUM_MAX) {
      acc += e.partials[(size_t)(i0 + i) * NSUM_MAX + v];
    }
  }
  sh.dred[grp][v] = acc;
  __syncthreads();
  if (tid < NSUM_MAX) {
    double t = 0.0;
    for (int g2 = 0; g2 < kBlock / NSUM_MAX; ++g2) t += sh.dred[g2][tid];
    sh.dred[0][tid] = t;
  }
  __syncthreads();
  const uint64_t t_solve = e.tline ? wall_clock64() : 0ull;
  if (tid == 0) {
    sh.stage_reset = 0;
    // all FP64 work of the solve reads and writes LDS (sh.dred, sh.ws, sh.Tinc),
    // so it adds no VGPR pressure to the search loop of this kernel
    const double* vals = sh.dred[0];
    ProblemState& S = e.st[p];
    const double n = vals[SUM_N];
    bool fin = false;
    S.n_corr = n;
    if (n < (double)e.prm.min_corr) {
      S.state = RTR_FAIL_NO_CORRESPONDENCES;
      S.mse = n > 0 ? vals[SUM_D2] / n : 0.0;
      fin = true;
    } else {
      double mse;
      solve_increment(MODE, vals, td.anchor, sh.Tinc, &mse, sh.ws);
      mat4_mul_to(sh.Tinc, sh.Tq, sh.Tn);
      S.iter += 1;
      S.mse = mse;
      double prev = S.prev_mse;
      int stt;
      if (S.stage) {
        // coarse stage: only the iteration limit ends the registration; an
        // increment under 1 mm and 1 mrad, the coarse limit or half of the
        // iteration limit moves it to all source points, with the MSE
        // history and the acceleration restarted (the iteration map changes)
        stt = S.iter >= e.prm.max_iter ? RTR_CONV_ITERATIONS : RTR_NOT_CONVERGED;
        const double cos_angle = 0.5 * (sh.Tinc[0] + sh.Tinc[5] + sh.Tinc[10] - 1.0);
        const double tsq = sh.Tinc[12] * sh.Tinc[12] + sh.Tinc[13] * sh.Tinc[13] + sh.Tinc[14] * sh.Tinc[14];
        if ((cos_angle >= 1.0 - 5e-7 && tsq <= 1e-6) || S.iter >= e.prm.coarse_iters ||
            2 * S.iter >= e.prm.max_iter) {
          S.stage = 0;
          S.coarse_its = S.iter;
          prev = DBL_MAX;
          sh.stage_reset = 1;
        }
      } else {
        stt = convergence(e.prm, S.iter, sh.Tinc, mse, &prev);
      }
      S.prev_mse = prev;
      S.state = stt;
      fin = (stt != RTR_NOT_CONVERGED);
      const double* Tnext = sh.Tn;
      if (!fin && gaa) {
        if (sh.stage_reset) {
          // first full-resolution iterate: the plain step, a fresh history
          anderson_reset(sh.aa);
        } else {
          // the next iterate is extrapolated from the recent ICP steps; iterates
          // are compared by the truncated energy (points without a
          // correspondence count r_max^2), which the plain iteration never raises
          const uint32_t ns = e.srcs[pr.src].n, st = S.stage ? (uint32_t)e.prm.coarse_stride : 1u;
          const double N = (double)((ns + st - 1u) / st);
          const double energy = (vals[SUM_D2] + (N - n) * (double)e.prm.r2f) / N;
          anderson_pose_step(sh.aa, e.prm.anderson, sh.Tq, sh.Tn, energy, sh.Ta, sh.ws);
          Tnext = sh.Ta;
        }
      }
      for (int k = 0; k < 16; ++k) S.T[k] = Tnext[k];
      if (e.hist) tf_rows_f32(Tnext, e.hist + ((size_t)p * (e.prm.max_iter + 1) + S.iter) * 12);
    }
    if (fin) S.finished = 1;
    if (fin && S.stage) S.coarse_its = S.iter;  // ended on the subsample (no correspondences)
    sh.fin = fin ? 1 : 0;
    if (e.tline && S.iter >= 1 && S.iter <= e.prm.max_iter) {
      uint64_t* ts = e.tline + ((size_t)p * e.prm.max_iter + (S.iter - 1)) * 2;
      ts[0] = t_solve;
      ts[1] = wall_clock64();
    }
  }
  __syncthreads();
  if (gaa) {
    if (tid < kAndersonStride) gaa[tid] = sh.aa[tid];
    __threadfence();
    __syncthreads();
  }
  if (tid == 0) {
    const bool fin = sh.fin != 0;
    __threadfence();
    store_release((uint32_t*)&e.st[p].gen, (uint32_t)(pass + 1) | (fin ? kFinishedBit : 0u));
    if (fin) atomicAdd(e.problems_done, 1u);
  }
}

// One correspondence (query q, target point c at cell-sorted slot, float d2)
// into the lane's FP64 moment sums.
template <int MODE>
__device__ __forceinline__ void add_pair(const DevEngine& e, float qx, float qy, float qz, const float4& c,
                                         uint32_t slot, float d2, double ax, double ay, double az, double* acc) {
  float nx = 0.f, ny = 0.f, nz = 0.f;
  if (MODE == RTR_MODE_POINT_TO_PLANE) {
    const float4 nn = e.tgt_nrm_sorted[slot];
    nx = nn.x;
    ny = nn.y;
    nz = nn.z;
  }
  accumulate_pair<MODE>(qx, qy, qz, c.x, c.y, c.z, nx, ny, nz, d2, ax, ay, az, acc);
}


// The wave's moment sums (DPP row sums, then the four rows) into its LDS
// slot, added to what earlier chunks of the item left there when `add`.
template <int NS, class Sh>
__device__ __forceinline__ void wave_totals(const double* acc, Sh& sh, int w, int lane, bool add) {
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const double s = wave_sum_f64(acc[k]);
    if (lane == 0) sh.red[w][k] = add ? sh.red[w][k] + s : s;
  }
}

// A copy of a wave-uniform value in scalar registers: loads of descriptors
// through a uniform address come back in vector registers, and the pointers
// derived from them would otherwise hold vector registers through the walk.
template <class T>
__device__ __forceinline__ T scalar_copy(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  T out;
  const uint32_t* s = reinterpret_cast<const uint32_t*>(&v);
  uint32_t* d = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
  return out;
}

template <int MODE, int NS, bool TILE, class TileT, int DEPTH>
__device__ void process_item(const DevEngine& e, int it, int pass, IcpShared<NS, TileT>& sh) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  TileT& tl = sh.tile[w];
  it = __builtin_amdgcn_readfirstlane(it);
  const uint2 itm = scalar_copy(e.items[it]);
  const uint32_t p = itm.x;
  const Problem pr = scalar_copy(e.probs[p]);
  const SourceDesc sd = scalar_copy(e.srcs[pr.src]);
  const TargetDesc td = scalar_copy(e.tgts[pr.tgt]);
  const GridView g = grid_of(e, td);
  const float r2f = e.prm.r2f;
  // the transform rows and the anchor are the same for every lane: keep them
  // in scalar registers
  float m[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) m[k] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__float_as_int(sh.m[k])));
  const double ax = (double)__builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__float_as_int(sh.anchor[0])));
  const double ay = (double)__builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__float_as_int(sh.anchor[1])));
  const double az = (double)__builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__float_as_int(sh.anchor[2])));
  uint32_t n_search = 0u;
  const uint32_t end = min(itm.y + (fine_pass(e, pass) ? (uint32_t)kFinePts : (uint32_t)e.item_pts), sd.n);
  constexpr bool kCompact = !TILE && DEPTH == 4;
  // Search seeds (global walks only): every searched query stores its
  // nearest neighbour's slot, and the next iteration starts from it.  The
  // transform moves little between iterations, so the seed's distance is
  // close to the nearest one: it trims the block's rows before their bounds
  // are loaded and stops the outer walk early.  Any slot of this target gives
  // a valid bound, so a seed is only range-checked, and the first iteration
  // of a launch (slots from an earlier launch) starts without one.
  uint32_t* const seeds = !TILE && e.nn_seed ? e.nn_seed + pr.pt_off : nullptr;
  const int stride = sh.stride;  // coarse stage: every stride-th source point
  const bool seeded = seeds && pass > 0;
  if (kCompact) {
    // Compacted walk (the throughput variant).  Queries that cannot have a
    // correspondence (distance field beyond r_max, NaN) take no lane, and the
    // others are grouped by their cell's distance-field value (own cell
    // occupied; a neighbour occupied; the outer ring the walk starts on: 2, 3,
    // 4 and beyond), so a wave's 64 lanes carry similar work.  The list order is fixed (class,
    // then query order), so the results are deterministic.
    for (uint32_t chunk = itm.y; chunk < end; chunk += kChunk) {
      const uint32_t cn = min(end - chunk, (uint32_t)kChunk);
      for (uint32_t o = tid; o < cn; o += kBlock) {
        const uint32_t sj = sd.off + chunk + o;
        float qx, qy, qz;
        int cx, cy, cz;
        apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
        uint8_t cls = 0, dv = 0;
        if ((stride == 1 || (chunk + o) % (uint32_t)stride == 0u) && query_cell(g, qx, qy, qz, cx, cy, cz)) {
          const int D = e.cell_dist ? (int)e.cell_dist[td.cell_off + ((uint32_t)cz * g.dy + cy) * g.dx + cx] : 0;
          if (!(e.cell_dist && D >= td.skip_ring)) {
            cls = (uint8_t)(D + 1 >= kQueryClasses ? kQueryClasses : D + 1);
            dv = (uint8_t)D;
          }
        }
        sh.qcls[o] = cls;
        sh.qd[o] = dv;
        sh.qseed[o] = seeded && cls ? seeds[chunk + o] : ~0u;
      }
      __syncthreads();
      {
        // The walk list: class by class, each class in query order.  Wave w
        // lists the classes of its quarter of the chunk (kSegs runs of 64
        // queries): counts per class first, then each wave writes from the
        // offset of its class in the earlier classes and earlier waves.
        constexpr int kSegs = kChunk / 64 / kWaves;
        const uint64_t below = (1ull << lane) - 1ull;
        uint8_t c[kSegs];
#pragma unroll
        for (int sg = 0; sg < kSegs; ++sg) {
          const uint32_t o = (uint32_t)((w * kSegs + sg) * 64 + lane);
          c[sg] = o < cn ? sh.qcls[o] : (uint8_t)0;
        }
#pragma unroll
        for (int cls = 1; cls <= kQueryClasses; ++cls) {
          uint32_t n = 0;
#pragma unroll
          for (int sg = 0; sg < kSegs; ++sg) n += (uint32_t)__popcll(__ballot(c[sg] == cls));
          if (lane == 0) sh.ccount[w][cls - 1] = n;
        }
        __syncthreads();
        uint32_t nl = 0;  // the list's length so far (classes below cls, all waves)
#pragma unroll
        for (int cls = 1; cls <= kQueryClasses; ++cls) {
          uint32_t off = nl;
          for (int v = 0; v < kWaves; ++v) {
            const uint32_t n = sh.ccount[v][cls - 1];
            off += v < w ? n : 0u;
            nl += n;
          }
#pragma unroll
          for (int sg = 0; sg < kSegs; ++sg) {
            const bool f = c[sg] == cls;
            const uint64_t bal = __ballot(f);
            if (f) sh.qlist[off + (uint32_t)__popcll(bal & below)] = (uint16_t)((w * kSegs + sg) * 64 + lane);
            off += (uint32_t)__popcll(bal);
          }
        }
        if (tid == 0) {
          sh.nlist = (int)nl;
          sh.lnext = kBlock;
        }
      }
      __syncthreads();
      const int nl = sh.nlist;
      // Blocks of 64 list entries go to the waves as they finish (an LDS
      // counter), so a wave that drew cheap queries takes the next block
      // instead of idling at the chunk's closing barrier; every query's
      // result lands in its own LDS slot, whichever wave walks it.
      for (int base = w * 64; base < nl;) {
        const int li = base + lane;
        if (li < nl) {
          const uint32_t o = sh.qlist[li];
          const uint32_t sj = sd.off + chunk + o;
          float qx, qy, qz;
          int cx, cy, cz;
          apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
          query_cell(g, qx, qy, qz, cx, cy, cz);
          const int D = sh.qd[o];
          ++n_search;
          GlobalVis<DEPTH> vis;
          vis.init(g.pts, qx, qy, qz);
          const uint32_t s0 = sh.qseed[o];
          if (s0 - g.pt_lo < g.n_pts) vis.seed(s0);
          if (D < 2) walk_block_global(g, qx, qy, qz, cx, cy, cz, r2f, vis);
          walk_far(g, qx, qy, qz, cx, cy, cz, r2f, vis, D >= 2 ? D - 1 : 1);
          // a seed that is still the nearest neighbour is not written again
          // (most queries of a converging problem keep theirs)
          if (seeds && vis.key != kNoKey && (vis.where & 0x7fffffffu) != s0) seeds[chunk + o] = vis.where & 0x7fffffffu;
          // the query's correspondence (its seed slot is read above by this
          // lane only, so the slot can hold the result)
          sh.qseed[o] = key_d2(vis.key) <= r2f ? (vis.where & 0x7fffffffu) : ~0u;
        }
        int nb = 0;
        if (lane == 0) nb = atomicAdd(&sh.lnext, 64);
        base = __builtin_amdgcn_readfirstlane(nb);
      }
      __syncthreads();
      // Moment phase: the walks above keep no moment sums live (their
      // registers go to the gathers in flight); the chunk's correspondences
      // are re-read from LDS in query order, each query's transform and
      // squared distance recomputed exactly as the walk computed them.
      double acc[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) acc[k] = 0.0;
      for (uint32_t o = tid; o < cn; o += kBlock) {
        const uint32_t slot = sh.qseed[o];
        if (slot == ~0u) continue;
        const uint32_t sj = sd.off + chunk + o;
        float qx, qy, qz;
        apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
        const float4 c = g.pts[slot];
        add_pair<MODE>(e, qx, qy, qz, c, slot, d2_f32(qx, qy, qz, c.x, c.y, c.z), ax, ay, az, acc);
      }
      wave_totals<NS>(acc, sh, w, lane, chunk != itm.y);
      __syncthreads();  // the next chunk rewrites the lists
    }
  }
  double acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k] = 0.0;
  for (uint32_t base = itm.y + (uint32_t)w * 64u; !kCompact && base < end; base += kBlock) {
    const uint32_t j = base + lane;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    int cx = 0, cy = 0, cz = 0;
    bool act = false;
    uint32_t s0 = ~0u;
    if (j < end) {
      const uint32_t sj = sd.off + j;
      apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
      act = (stride == 1 || j % (uint32_t)stride == 0u) && query_cell(g, qx, qy, qz, cx, cy, cz);
      if (seeded) s0 = seeds[j];
    }
    // distance field: cells within D - 1 rings of the query's cell are empty
    bool blk = act;  // the lane walks its 3x3x3 block
    int k0 = 1;      // box radius searched (or known empty) before the outer walk
    if (act && e.cell_dist) {
      const int D = e.cell_dist[td.cell_off + ((uint32_t)cz * g.dy + cy) * g.dx + cx];
      if (D >= td.skip_ring) {
        act = false;  // every target point is farther than r_max
        blk = false;
      } else if (D >= 2) {
        blk = false;
        k0 = D - 1;
      }
    }
    if (act) ++n_search;
    GlobalVis<DEPTH> vis;
    vis.init(g.pts, qx, qy, qz);
    if (act && s0 - g.pt_lo < g.n_pts) vis.seed(s0);
    bool staged = false;
    int X0 = 0, Y0 = 0, Z0 = 0, NY = 1, W1 = 1, nrows = 0;
    if (TILE && e.use_tile && __ballot(blk) != 0ull) {
      X0 = max(wave_min_i32(blk ? cx : 0x7fffffff) - 1, 0);
      Y0 = max(wave_min_i32(blk ? cy : 0x7fffffff) - 1, 0);
      Z0 = max(wave_min_i32(blk ? cz : 0x7fffffff) - 1, 0);
      const int X1 = min(wave_max_i32(blk ? cx : -1) + 1, g.dx - 1);
      const int Y1 = min(wave_max_i32(blk ? cy : -1) + 1, g.dy - 1);
      const int Z1 = min(wave_max_i32(blk ? cz : -1) + 1, g.dz - 1);
      NY = Y1 - Y0 + 1;
      W1 = X1 - X0 + 2;
      nrows = NY * (Z1 - Z0 + 1);
      if (TILE && e.use_tile && nrows <= kTileRows && nrows * W1 <= kTileCells) {
        uint32_t a = 0u, len = 0u;
        if (lane < nrows) {
          const int y = Y0 + lane % NY, z = Z0 + lane / NY;
          const uint32_t row = ((uint32_t)z * g.dy + y) * g.dx;
          a = g.cell_start[row + X0];
          len = g.cell_start[row + X1 + 1] - a;
        }
        uint32_t total;
        const uint32_t off = wave_excl_scan(len, &total);
        if (total <= (uint32_t)kTilePts) {
          staged = true;
          if (lane < nrows) {
            tl.row_a[lane] = a;
            tl.row_off[lane] = off;
          }
          __builtin_amdgcn_wave_barrier();
          for (int t = lane; t < nrows * W1; t += 64) {
            const int r = t / W1, i = t - r * W1;
            const int y = Y0 + r % NY, z = Z0 + r / NY;
            const uint32_t row = ((uint32_t)z * g.dy + y) * g.dx;
            tl.cell[t] = (uint16_t)(g.cell_start[row + X0 + i] - tl.row_a[r] + tl.row_off[r]);
          }
          for (uint32_t t = lane; t < total; t += 64) {
            const int r = tile_row_of(tl, nrows, t);
            tl.pts[t] = g.pts[tl.row_a[r] + (t - tl.row_off[r])];
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
    if (act) {
      if (!blk) {
        // block known empty: only the outer rings
      } else if (staged) {
        const int xlo = max(cx - 1, 0) - X0;
        const int xhi = min(cx + 1, g.dx - 1) - X0 + 1;
        for (int dz = -1; dz <= 1; ++dz) {
          const int z = cz + dz;
          if (z < 0 || z >= g.dz) continue;
          for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.dy) continue;
            const int r = (z - Z0) * NY + (y - Y0);
            const uint32_t b0 = tl.cell[r * W1 + xlo], b1 = tl.cell[r * W1 + xhi];
            uint32_t t = b0;
            for (; t + 2 <= b1; t += 2) {
              const float4 c0 = tl.pts[t];
              const float4 c1 = tl.pts[t + 1];
              const uint64_t k0v = nn_key(d2_f32(qx, qy, qz, c0.x, c0.y, c0.z), __float_as_int(c0.w));
              const uint64_t k1v = nn_key(d2_f32(qx, qy, qz, c1.x, c1.y, c1.z), __float_as_int(c1.w));
              if (k0v < vis.key) {
                vis.key = k0v;
                vis.where = t;
              }
              if (k1v < vis.key) {
                vis.key = k1v;
                vis.where = t + 1;
              }
            }
            if (t < b1) {
              const float4 c = tl.pts[t];
              const uint64_t kv = nn_key(d2_f32(qx, qy, qz, c.x, c.y, c.z), __float_as_int(c.w));
              if (kv < vis.key) {
                vis.key = kv;
                vis.where = t;
              }
            }
          }
        }
      } else {
        walk_block_global(g, qx, qy, qz, cx, cy, cz, r2f, vis);
      }
      walk_far(g, qx, qy, qz, cx, cy, cz, r2f, vis, k0);
      if (seeds && vis.key != kNoKey && (vis.where & 0x7fffffffu) != s0) seeds[j] = vis.where & 0x7fffffffu;
      if (key_d2(vis.key) <= r2f) {
        float4 c;
        uint32_t slot = 0u;
        if (vis.where & 0x80000000u) {
          slot = vis.where & 0x7fffffffu;
          c = g.pts[slot];
        } else {
          c = tl.pts[vis.where];
          if (MODE == RTR_MODE_POINT_TO_PLANE) {
            const int r = tile_row_of(tl, nrows, vis.where);
            slot = tl.row_a[r] + (vis.where - tl.row_off[r]);
          }
        }
        add_pair<MODE>(e, qx, qy, qz, c, slot, key_d2(vis.key), ax, ay, az, acc);
      }
    }
    // the next round rewrites this wave's tile
    __builtin_amdgcn_wave_barrier();
  }
  // search statistics: per wave, then the workgroup's LDS totals
  n_search = wave_sum_u32(n_search);
  if (lane == 0) atomicAdd(&sh.n_search, n_search);
  // wave totals (DPP row sums), then the four waves in a fixed order
  if (!kCompact) wave_totals<NS>(acc, sh, w, lane, false);
  __syncthreads();
  if (tid < NS) {
    double s = sh.red[0][tid];
    for (int ww = 1; ww < kWaves; ++ww) s += sh.red[ww][tid];
    e.partials[(size_t)it * NSUM_MAX + tid] = s;
    __threadfence();
  }
  __syncthreads();
  if (tid == 0) {
    const uint32_t prev = atomicAdd(&e.done[p], 1u);
    sh.last = (prev + 1u == done_through(e, pr, pass));
  }
  __syncthreads();
  if (sh.last) {
    __threadfence();
    solve_problem<MODE, NS, TileT>(e, p, pr, td, pass, sh);
  }
}

template <int MODE, bool TILE, int DEPTH, int WAVES>
// Point-to-point with 4-deep gathers asks for at least 3 waves per SIMD
// (<= 168 VGPRs): the gathers of the ring walk are latency-bound, so
// resident waves are what hides them; a variant asks for 4 (<= 128 VGPRs,
// its few spills outside the walks) for launches tuned to 4 workgroups per
// CU.  The 8-deep variant keeps more gathers in flight per wave and runs 2
// waves per SIMD; it serves launches tuned to 2 workgroups per CU, where the
// third wave would not be resident anyway.  Point-to-plane carries 29 FP64
// sums, runs 2 waves per SIMD and is always 8-deep.
__global__ __launch_bounds__(kBlock, WAVES) void k_icp(DevEngine e) {
  constexpr int NS = MODE == RTR_MODE_POINT_TO_POINT ? SUM_P2P_COUNT : SUM_P2L_COUNT;
  using TileT = typename std::conditional<TILE, WaveTile, WaveTileNone>::type;
  __shared__ IcpShared<NS, TileT> sh;
  const int tid = threadIdx.x;
  if (tid == 0) sh.n_search = 0u;
  for (int pass = 0; pass < e.prm.max_iter; ++pass) {
    // Workgroups beyond one per CU leave once few problems are unfinished:
    // they raise throughput while many problems iterate, but idle ones
    // polling the work queue stretch the last problems' iterations.  They
    // leave between passes, holding no claimed item, and one workgroup per CU
    // stays, so every remaining item is still claimed.
    if (tid == 0) {
      const uint32_t done = load_acquire(e.problems_done);
      sh.stop = (done >= (uint32_t)e.n_prob) || (load_acquire((const uint32_t*)e.error) != 0u) ||
                ((int)blockIdx.x >= e.retire_from && done + e.retire_active >= (uint32_t)e.n_prob);
    }
    __syncthreads();
    if (sh.stop) break;
    for (;;) {
      if (tid == 0) {
        // own XCD's queue first (L2 locality), then steal from the others
        int it = e.n_items;
        const int x0 = (int)(blockIdx.x % kXcds);
        const bool fine = fine_pass(e, pass);
        const uint32_t* qoff = fine ? e.fxcd_off : e.xcd_off;
        for (int t = 0; t < kXcds; ++t) {
          const int x = (x0 + t) % kXcds;
          const uint32_t cnt = qoff[x + 1] - qoff[x];
          uint32_t* ctr = &e.claim[(size_t)(pass * kXcds + x) * e.claim_stride];
          if (cnt == 0u || __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt) continue;
          const uint32_t c = atomicAdd(ctr, 1u);
          if (c < cnt) {
            it = (int)(qoff[x] + c);
            const uint32_t p = e.items[it].x;
            if (load_acquire((const uint32_t*)&e.st[p].gen) & kFinishedBit) {
              // a finished problem: its remaining items of this pass (contiguous
              // in this queue) are skipped in one step, and the claim goes on
              const Problem& pr = e.probs[p];
              atomicMax(ctr, first_item_of(pr, fine) + items_of(pr, fine) - qoff[x]);
              it = e.n_items;
              t = -1;  // rescan the queues from the own XCD's
              continue;
            }
            break;
          }
        }
        int run = 0;
        if (it < e.n_items) {
          const uint32_t p = e.items[it].x;
          const uint32_t* word = (const uint32_t*)&e.st[p].gen;
          uint32_t wv = load_acquire(word);
          int spins = 0;
          // bounded wait with backoff; the error flag (on its own cache line)
          // is polled every 256 rounds
          while (!(wv & kFinishedBit) && (int)wv < pass) {
            ++spins;
            if (spins > kMaxSpins || ((spins & 255) == 0 && load_acquire((const uint32_t*)e.error) != 0u)) {
              atomicOr(e.error, 1);
              break;
            }
            if (spins < 64)
              __builtin_amdgcn_s_sleep(1);
            else
              __builtin_amdgcn_s_sleep(8);
            wv = load_acquire(word);
          }
          if (!(wv & kFinishedBit) && (int)wv == pass) {
            run = 1;
            tf_rows_f32(e.st[p].T, sh.m);
            sh.stride = e.st[p].stage ? e.prm.coarse_stride : 1;
            const TargetDesc& td = e.tgts[e.probs[p].tgt];
            sh.anchor[0] = td.anchor[0];
            sh.anchor[1] = td.anchor[1];
            sh.anchor[2] = td.anchor[2];
          }
        }
        sh.item = it;
        sh.run = run;
      }
      __syncthreads();
      const int it = sh.item;
      if (it >= e.n_items) break;
      if (sh.run) process_item<MODE, NS, TILE, TileT, DEPTH>(e, it, pass, sh);
      __syncthreads();
    }
  }
  __syncthreads();
  if (tid == 0) atomicAdd(e.queries_searched, (unsigned long long)sh.n_search);
}

template <int MODE, bool TILE, int D, int W>
static int blocks_per_cu_of() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp<MODE, TILE, D, W>, kBlock, 0) != hipSuccess || nb < 1)
    nb = 1;
  return nb;
}

int icp_blocks_per_cu(int mode, int use_tile, int variant) {
  if (mode == RTR_MODE_POINT_TO_PLANE)
    return use_tile ? blocks_per_cu_of<RTR_MODE_POINT_TO_PLANE, true, 8, 2>()
                    : blocks_per_cu_of<RTR_MODE_POINT_TO_PLANE, false, 8, 2>();
  if (variant == 1)
    return use_tile ? blocks_per_cu_of<RTR_MODE_POINT_TO_POINT, true, 8, 2>()
                    : blocks_per_cu_of<RTR_MODE_POINT_TO_POINT, false, 8, 2>();
  // (the LDS tiles leave room for 3 workgroups per CU: no 4-wave tile variant)
  if (variant == 2 && !use_tile) return blocks_per_cu_of<RTR_MODE_POINT_TO_POINT, false, 4, 4>();
  return use_tile ? blocks_per_cu_of<RTR_MODE_POINT_TO_POINT, true, 4, 3>()
                  : blocks_per_cu_of<RTR_MODE_POINT_TO_POINT, false, 4, 3>();
}

void launch_icp(const DevEngine& e, int grid_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_icp_init, dim3((e.n_prob + kBlock - 1) / kBlock), dim3(kBlock), 0, s, e);
  const dim3 g(grid_blocks), b(kBlock);
  if (e.prm.mode == RTR_MODE_POINT_TO_PLANE) {
    if (e.use_tile)
      hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_PLANE, true, 8, 2>), g, b, 0, s, e);
    else
      hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_PLANE, false, 8, 2>), g, b, 0, s, e);
  } else if (e.variant == 1) {
    if (e.use_tile)
      hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_POINT, true, 8, 2>), g, b, 0, s, e);
    else
      hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_POINT, false, 8, 2>), g, b, 0, s, e);
  } else if (e.variant == 2 && !e.use_tile) {
    hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_POINT, false, 4, 4>), g, b, 0, s, e);
  } else {
    if (e.use_tile)
      hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_POINT, true, 4, 3>), g, b, 0, s, e);
    else
      hipLaunchKernelGGL((k_icp<RTR_MODE_POINT_TO_POINT, false, 4, 3>), g, b, 0, s, e);
  }
}

// ---------------------------------------------------------------------------
// per-point outputs
// ---------------------------------------------------------------------------

__global__ void k_corr(DevEngine e, int prob, const double* T, float r2f, int32_t* idx, float* d2) {
  const Problem pr = e.probs[prob];
  const SourceDesc sd = e.srcs[pr.src];
  const TargetDesc td = e.tgts[pr.tgt];
  const GridView g = grid_of(e, td);
  float m[12];
  tf_rows_f32(T ? T : e.st[prob].T, m);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < sd.n; j += gridDim.x * blockDim.x) {
    const uint32_t sj = sd.off + j;
    float qx, qy, qz;
    apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
    const NNBest b = nn_search(g, qx, qy, qz, r2f);
    const uint32_t o = e.perm[sj];
    if (b.key != ~0ull) {
      idx[o] = (int32_t)(uint32_t)(b.key & 0xffffffffull);
      d2[o] = key_d2(b.key);
    } else {
      idx[o] = -1;
      d2[o] = __builtin_inff();
    }
  }
}

void launch_correspondences(const DevEngine& e, int prob, const double* T, float r2f, int32_t* idx,
                            float* d2, hipStream_t s) {
  hipLaunchKernelGGL(k_corr, dim3(256), dim3(kBlock), 0, s, e, prob, T, r2f, idx, d2);
}

__global__ void k_transform_out(DevEngine e, int prob, float* out) {
  const Problem pr = e.probs[prob];
  const SourceDesc sd = e.srcs[pr.src];
  float m[12];
  tf_rows_f32(e.st[prob].T, m);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < sd.n; j += gridDim.x * blockDim.x) {
    const uint32_t sj = sd.off + j;
    float qx, qy, qz;
    apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
    const size_t o = e.perm[sj];
    out[3 * o] = qx;
    out[3 * o + 1] = qy;
    out[3 * o + 2] = qz;
  }
}

void launch_transform_out(const DevEngine& e, int prob, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_transform_out, dim3(256), dim3(kBlock), 0, s, e, prob, out);
}

__global__ void k_fitness_d2(DevEngine e, float* d2_all) {
  const int prob = blockIdx.y;
  const Problem pr = e.probs[prob];
  const SourceDesc sd = e.srcs[pr.src];
  const TargetDesc td = e.tgts[pr.tgt];
  const GridView g = grid_of(e, td);
  float m[12];
  tf_rows_f32(e.st[prob].T, m);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < sd.n; j += gridDim.x * blockDim.x) {
    const uint32_t sj = sd.off + j;
    float qx, qy, qz;
    apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
    const NNBest b = nn_search(g, qx, qy, qz, FLT_MAX);
    d2_all[pr.pt_off + j] = b.key != ~0ull ? key_d2(b.key) : __builtin_inff();
  }
}

void launch_fitness_d2(const DevEngine& e, float* d2_all, hipStream_t s) {
  if (e.n_prob == 0) return;
  hipLaunchKernelGGL(k_fitness_d2, dim3(64, e.n_prob), dim3(kBlock), 0, s, e, d2_all);
}

// ---------------------------------------------------------------------------
// k-NN PCA normals of the targets
// ---------------------------------------------------------------------------

constexpr int kNrmBlock = 128;
constexpr int kMaxK = 32;

// Sorted insertion of the candidates into a k-long list.  Keys are unique
// (their low word is the point's index), so the final list does not depend on
// the order in which the candidates arrive: the scan loads four candidates
// before inserting them one by one, keeping four loads in flight, and keeps
// the list's last key in a register.  Initialise with cnt = 0.
struct KnnVisitor {
  const float4* pts;
  float qx, qy, qz;
  int k, cnt;
  uint64_t kth;    // key[k - 1] once cnt == k
  uint64_t* key;   // LDS, stride kNrmBlock
  uint32_t* slot;  // LDS, stride kNrmBlock; nullptr keeps keys only (their low word is the point's original index)
  __device__ void insert(const float4& t, uint32_t j) {
    union {
      float f;
      int32_t i;
    } w;
    w.f = t.w;
    const uint64_t kk = nn_key(d2_f32(qx, qy, qz, t.x, t.y, t.z), w.i);
    if (cnt == k && kk >= kth) return;
    int pos = cnt < k ? cnt++ : k - 1;
    while (pos > 0 && kk < key[(pos - 1) * kNrmBlock]) {
      key[pos * kNrmBlock] = key[(pos - 1) * kNrmBlock];
      if (slot) slot[pos * kNrmBlock] = slot[(pos - 1) * kNrmBlock];
      --pos;
    }
    key[pos * kNrmBlock] = kk;
    if (slot) slot[pos * kNrmBlock] = j;
    if (cnt == k) kth = key[(k - 1) * kNrmBlock];
  }
  __device__ void scan(uint32_t b, uint32_t e) {
    uint32_t j = b;
    for (; j + 4 <= e; j += 4) {
      float4 t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = pts[j + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) insert(t[u], j + u);
    }
    for (; j < e; ++j) insert(pts[j], j);
  }
  __device__ float bound() const { return cnt < k ? 3.4e38f : key_d2(kth); }
  // the judge's list interface: reset, size, and the keys in ascending order
  __device__ void reset() { cnt = 0; }
  __device__ int size() const { return cnt; }
  __device__ uint64_t first() const { return key[0]; }
  __device__ uint64_t last() const { return key[(k - 1) * kNrmBlock]; }
  template <class F>
  __device__ void for_each(F f) const {
#pragma unroll 4
    for (int i = 0; i < cnt; ++i) f(key[i * kNrmBlock]);
  }
};

// The K nearest keys in registers, ascending, empty slots holding ~0 (no
// real key reaches it: a finite or infinite float d2 has high word below
// 0xffffffff).  A candidate below the last key goes through K compare-swap
// steps, which leave the K smallest keys in order: no serial LDS shifts, whose
// latency, taken by every lane of a wave at the pace of its longest shift, set
// the walk's time.  Keys are unique, so the list does not depend on the order
// of the candidates.
template <int K>
struct KnnRegList {
  const float4* pts;
  float qx, qy, qz;
  uint64_t key[K];
  __device__ void reset() {
#pragma unroll
    for (int u = 0; u < K; ++u) key[u] = ~0ull;
  }
  __device__ void insert(const float4& t) {
    union {
      float f;
      int32_t i;
    } w;
    w.f = t.w;
    uint64_t kk = nn_key(d2_f32(qx, qy, qz, t.x, t.y, t.z), w.i);
    if (kk >= key[K - 1]) return;
    // two halves: a candidate not below the first half's last key leaves
    // that half as it is, and a wave skips the half when none of its lanes
    // reaches it (four quarters spill)
    constexpr int H = K / 2;
    if (kk < key[H - 1]) {
#pragma unroll
      for (int u = 0; u < H; ++u) {
        const uint64_t a = key[u];
        const bool lt = kk < a;
        key[u] = lt ? kk : a;
        kk = lt ? a : kk;
      }
    }
#pragma unroll
    for (int u = H; u < K; ++u) {
      const uint64_t a = key[u];
      const bool lt = kk < a;
      key[u] = lt ? kk : a;
      kk = lt ? a : kk;
    }
  }
  __device__ void scan(uint32_t b, uint32_t e) {
    uint32_t j = b;
    for (; j + 4 <= e; j += 4) {
      float4 t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = pts[j + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) insert(t[u]);
    }
    for (; j < e; ++j) insert(pts[j]);
  }
  __device__ float bound() const { return key[K - 1] == ~0ull ? 3.4e38f : key_d2(key[K - 1]); }
  __device__ int size() const {
    int c = 0;
#pragma unroll
    for (int u = 0; u < K; ++u) c += key[u] != ~0ull;
    return c;
  }
  __device__ uint64_t first() const { return key[0]; }
  __device__ uint64_t last() const { return key[K - 1]; }
  template <class F>
  __device__ void for_each(F f) const {
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (key[u] != ~0ull) f(key[u]);
  }
};

__device__ void jacobi3(double* a, double* V) {
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    const double off = fabs(a[1]) + fabs(a[2]) + fabs(a[5]);
    const double dg = fabs(a[0]) + fabs(a[4]) + fabs(a[8]);
    if (off == 0.0 || off <= 1e-18 * dg) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        const double apq = a[p * 3 + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * 4] - a[p * 4]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k * 3 + p], akq = a[k * 3 + q];
          a[k * 3 + p] = c * akp - s * akq;
          a[k * 3 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p * 3 + k], aqk = a[q * 3 + k];
          a[p * 3 + k] = c * apk - s * aqk;
          a[q * 3 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
          V[k * 3 + p] = c * vkp - s * vkq;
          V[k * 3 + q] = s * vkp + c * vkq;
        }
      }
  }
}

// One thread per cell-sorted target point: exact k nearest (ordered by
// (d2, index), the point itself included), FP64 two-pass covariance, the
// eigenvector of the smallest eigenvalue, flipped towards the viewpoint
// (pcl::NormalEstimation + flipNormalTowardsViewpoint semantics).
__global__ __launch_bounds__(kNrmBlock) void k_target_normals(DevEngine e, int k, float vpx,
                                                              float vpy, float vpz,
                                                              float* nrm_raw) {
  // k-NN lists in dynamic LDS sized for k (as in k_judge)
  extern __shared__ uint64_t s_dyn[];
  uint64_t* const s_key = s_dyn;
  uint32_t* const s_slot = reinterpret_cast<uint32_t*>(s_dyn + (size_t)k * kNrmBlock);
  const TargetDesc td = e.tgts[blockIdx.y];
  const GridView g = grid_of(e, td);
  const uint32_t base = e.cell_start[td.cell_off];
  for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < td.n_valid; j0 += gridDim.x * blockDim.x) {
    const uint32_t j = j0 + threadIdx.x;
    if (j >= td.n_valid) continue;
    const float4 q = e.tgt_sorted[base + j];
    KnnVisitor v;
    v.pts = e.tgt_sorted;
    v.qx = q.x;
    v.qy = q.y;
    v.qz = q.z;
    v.k = k;
    v.cnt = 0;
    v.key = s_key + threadIdx.x;
    v.slot = s_slot + threadIdx.x;
    ring_walk(g, q.x, q.y, q.z, FLT_MAX, v);
    double nv[3] = {__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
    if (v.cnt >= 3) {
      double mean[3] = {0, 0, 0};
#pragma unroll 4
      for (int i = 0; i < v.cnt; ++i) {
        const float4 p = e.tgt_sorted[v.slot[i * kNrmBlock]];
        mean[0] += p.x;
        mean[1] += p.y;
        mean[2] += p.z;
      }
      for (int a = 0; a < 3; ++a) mean[a] /= v.cnt;
      double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
      for (int i = 0; i < v.cnt; ++i) {
        const float4 p = e.tgt_sorted[v.slot[i * kNrmBlock]];
        const double d[3] = {p.x - mean[0], p.y - mean[1], p.z - mean[2]};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) C[r * 3 + c] += d[r] * d[c];
      }
      for (int i = 0; i < 9; ++i) C[i] /= v.cnt;
      double V[9];
      jacobi3(C, V);
      int mi = 0;
      if (C[4] < C[mi * 4]) mi = 1;
      if (C[8] < C[mi * 4]) mi = 2;
      nv[0] = V[mi];
      nv[1] = V[3 + mi];
      nv[2] = V[6 + mi];
      const double vx = (double)vpx - q.x, vy = (double)vpy - q.y, vz = (double)vpz - q.z;
      if (vx * nv[0] + vy * nv[1] + vz * nv[2] < 0) {
        nv[0] = -nv[0];
        nv[1] = -nv[1];
        nv[2] = -nv[2];
      }
    }
    union {
      float f;
      int32_t i;
    } w;
    w.f = q.w;
    float* o = nrm_raw + 3 * ((size_t)td.raw_off + (uint32_t)w.i);
    o[0] = (float)nv[0];
    o[1] = (float)nv[1];
    o[2] = (float)nv[2];
    if (e.tgt_nrm_sorted) e.tgt_nrm_sorted[base + j] = make_float4((float)nv[0], (float)nv[1], (float)nv[2], 0.f);
  }
}

void launch_target_normals(const DevEngine& e, int k, float vpx, float vpy, float vpz,
                           int max_tgt_pts, hipStream_t s) {
  if (e.n_tgt == 0) return;
  if (k > kMaxK) k = kMaxK;
  unsigned bx = (unsigned)((max_tgt_pts + kNrmBlock - 1) / kNrmBlock);
  if (bx < 1) bx = 1;
  if (bx > 1024) bx = 1024;
  const size_t lds = (size_t)k * kNrmBlock * (sizeof(uint64_t) + sizeof(uint32_t));
  hipLaunchKernelGGL(k_target_normals, dim3(bx, e.n_tgt), dim3(kNrmBlock), lds, s, e, k, vpx, vpy,
                     vpz, const_cast<float*>(e.tgt_nrm_raw));
}


// ---------------------------------------------------------------------------
// registration judge (plane-distance RMSE of the inliers)
// ---------------------------------------------------------------------------

// gridDim.y workgroups per problem.  A strided sample of up to max_pts source points
// (in the stored order) is moved by the problem's final pose (float, the same
// rounding as the ICP queries).  A point is an inlier when its nearest target
// point lies within the correspondence radius (d2 <= r2f).  For an inlier the
// k nearest target points give the local plane (PCA through their centroid,
// FP64); when that neighbourhood is planar (surface variation
// l0 / (l0 + l1 + l2) <= max_curv, PCL's curvature) the inlier's squared
// distance to the plane is summed.  Edges and corners (several faces in one
// neighbourhood) are left out of the sum: there the fitted plane is not the
// surface.  Unlike the point-to-point MSE, this residual does not grow with
// the target's sampling spacing, so one threshold serves sparse and dense
// targets.
// The sample of problem p is split over gridDim.y workgroups in consecutive
// runs (the stored order is spatially sorted, so the lanes of a wave hold
// nearby points and walk alike); workgroup s writes its partial sums to
// out[4 (p * gridDim.y + s) + {0, 1, 2, 3}] = {sum of squared plane
// distances, planar inliers, inliers, sampled points}, which the host adds
// in fixed order.
// k = 20 (the default) keeps each k-NN list in registers (KnnRegList);
// other k keep theirs in dynamic LDS sized for k, holding keys only (k *
// kNrmBlock keys).  Either way a neighbour's coordinates are read back from
// the raw target cloud by the original index in its key's low word.  Three
// waves per SIMD: the register list costs 40 VGPRs, and holding the kernel
// to 168 VGPRs (a few spilled) measured 0.99 ms on 256 pairs x 1024 points
// against 1.17 ms at two waves and 1.19 ms at four (more spills).
template <int K>
__global__ __launch_bounds__(kNrmBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_judge(
    DevEngine e, int k, int max_pts, float r2f, double max_curv, double* out) {
  extern __shared__ uint64_t s_dyn[];
  uint64_t* const s_key = s_dyn;
  const int prob = blockIdx.x;
  const Problem pr = e.probs[prob];
  const SourceDesc sd = e.srcs[pr.src];
  const TargetDesc td = e.tgts[pr.tgt];
  const GridView g = grid_of(e, td);
  float m[12];
  tf_rows_f32(e.st[prob].T, m);
  const uint32_t n = sd.n;
  const uint32_t ms = n < (uint32_t)max_pts ? n : (uint32_t)max_pts;
  const int kk = K ? K : k;
  double ssum = 0.0, cnt = 0.0, pcnt = 0.0;
  const uint32_t split = gridDim.y;
  const uint32_t per = (ms + split - 1) / split;
  const uint32_t j0 = blockIdx.y * per, j1 = min(ms, j0 + per);
  for (uint32_t j = j0 + threadIdx.x; j < j1; j += kNrmBlock) {
    const uint32_t sj = sd.off + (uint32_t)(((uint64_t)j * n) / ms);
    float qx, qy, qz;
    apply_f32(m, e.sx[sj], e.sy[sj], e.sz[sj], qx, qy, qz);
    typename std::conditional<K == 0, KnnVisitor, KnnRegList<(K > 0 ? K : 1)>>::type v;
    v.pts = e.tgt_sorted;
    v.qx = qx;
    v.qy = qy;
    v.qz = qz;
    if constexpr (K == 0) {
      v.k = k;
      v.key = s_key + threadIdx.x;
      v.slot = nullptr;
    }
    v.reset();
    // A query whose cell's distance-field value reaches skip_ring has no
    // target point within the correspondence radius: an outlier, no walk.
    // Every other query takes one unbounded walk, whose list is the exact
    // k-NN list and whose nearest key decides the inlier test.
    {
      int cx, cy, cz;
      if (e.cell_dist && query_cell(g, qx, qy, qz, cx, cy, cz) &&
          (int)e.cell_dist[td.cell_off + ((uint32_t)cz * g.dy + cy) * g.dx + cx] >= td.skip_ring)
        continue;
    }
    ring_walk(g, qx, qy, qz, FLT_MAX, v);
    if (v.size() == 0 || !(key_d2(v.first()) <= r2f)) continue;
    cnt += 1.0;
    const int nn = v.size();
    if (nn < 3) continue;
    const float* const raw = e.tgt_raw + 3 * (size_t)td.raw_off;
    double mean[3] = {0, 0, 0};
    v.for_each([&](uint64_t key) {
      const float* p = raw + 3 * (size_t)(uint32_t)key;
      mean[0] += p[0];
      mean[1] += p[1];
      mean[2] += p[2];
    });
    for (int a = 0; a < 3; ++a) mean[a] /= nn;
    double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    v.for_each([&](uint64_t key) {
      const float* p = raw + 3 * (size_t)(uint32_t)key;
      const double d[3] = {p[0] - mean[0], p[1] - mean[1], p[2] - mean[2]};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) C[r * 3 + c] += d[r] * d[c];
    });
    double V[9];
    jacobi3(C, V);
    int mi = 0;
    if (C[4] < C[mi * 4]) mi = 1;
    if (C[8] < C[mi * 4]) mi = 2;
    const double lsum = C[0] + C[4] + C[8];
    if (!(lsum > 0.0) || C[mi * 4] > max_curv * lsum) continue;
    const double dist = V[mi] * (qx - mean[0]) + V[3 + mi] * (qy - mean[1]) + V[6 + mi] * (qz - mean[2]);
    ssum += dist * dist;
    pcnt += 1.0;
  }
  // fixed-order block reduction: wave totals (DPP row sums), then the waves
  ssum = wave_sum_f64(ssum);
  cnt = wave_sum_f64(cnt);
  pcnt = wave_sum_f64(pcnt);
  // the partials reuse the start of the dynamic LDS (for LDS lists, k >= 3
  // lists hold far more than 3 doubles per wave) once every thread is done
  // with its list: no static LDS
  double(*s_red)[kNrmBlock / 64] = reinterpret_cast<double(*)[kNrmBlock / 64]>(s_dyn);
  __syncthreads();
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[0][w] = ssum;
    s_red[1][w] = pcnt;
    s_red[2][w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, pc = 0.0, c = 0.0;
    for (int i = 0; i < kNrmBlock / 64; ++i) {
      a += s_red[0][i];
      pc += s_red[1][i];
      c += s_red[2][i];
    }
    double* o = out + 4 * ((size_t)prob * split + blockIdx.y);
    o[0] = a;
    o[1] = pc;
    o[2] = c;
    o[3] = blockIdx.y == 0 ? (double)ms : 0.0;
  }
}

void launch_judge(const DevEngine& e, int k, int max_pts, float r2f, double max_curv, double* out,
                  hipStream_t s) {
  if (e.n_prob == 0) return;
  if (k > kMaxK) k = kMaxK;
  const size_t red = 3 * (kNrmBlock / 64) * sizeof(double);
  if (k == 20) {
    hipLaunchKernelGGL(k_judge<20>, dim3(e.n_prob, kJudgeSplit), dim3(kNrmBlock), red, s, e, k, max_pts, r2f,
                       max_curv, out);
    return;
  }
  const size_t lds = (size_t)k * kNrmBlock * sizeof(uint64_t);
  hipLaunchKernelGGL(k_judge<0>, dim3(e.n_prob, kJudgeSplit), dim3(kNrmBlock), lds > red ? lds : red, s, e, k,
                     max_pts, r2f, max_curv, out);
}

}  // namespace rtr
