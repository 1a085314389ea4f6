// Grid construction on the device: per-target uniform grids (cell-sorted
// float4 points + cell start offsets) and per-source spatial (Morton) sort,
// batched over all clouds of an engine.  Counting sort: count -> flat
// exclusive scan over every cloud's cell range -> scatter.
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "geom.hpp"
#include "wave.hpp"

namespace rtr {

// ---------------------------------------------------------------------------
// fill and scan
// ---------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  size_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, v, n);
}

constexpr int kScanPer = 8;                    // elements per thread
constexpr int kScanTile = kBlock * kScanPer;   // 2048 per block

size_t scan_tmp_words(uint32_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// exclusive block scan of one value per thread; returns prefix, total in *tot
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tot) {
  __shared__ uint32_t wsum[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int i = 0; i < kWaves; ++i) {
    if (i < w) base += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *tot = all;
  return base + inc - v;
}

__global__ void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* tile_sums) {
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void k_scan_tiles(uint32_t* tile_sums, uint32_t ntiles) {
  uint32_t carry = 0;
  for (uint32_t b = 0; b < ntiles; b += kBlock) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < ntiles ? tile_sums[i] : 0;
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, &tot);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

__global__ void k_scan_apply(const uint32_t* in, uint32_t* out, uint32_t n,
                             const uint32_t* tile_sums) {
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan(s, &tot) + tile_sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp,
                           hipStream_t s) {
  if (n == 0) return;
  uint32_t ntiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kBlock), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, tmp, ntiles);
  hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kBlock), 0, s, in, out, n, tmp);
}

// ---------------------------------------------------------------------------
// block reductions
// ---------------------------------------------------------------------------

// (wave reductions: DPP row operations, wave.hpp)


// Bounding boxes of many clouds at once: each cloud is split over many
// workgroups, whose wave minima / maxima (as order-preserving integer keys)
// and finite-point counts meet in 8 words per cloud through atomics.  Min and
// max do not depend on the order they are taken in, so the box is the same
// as a serial pass.
__device__ __forceinline__ uint32_t ord_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_val(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// words per cloud: lo key x3, hi key x3, finite count, unused
__global__ void k_bbox_init(uint32_t* bb, uint32_t n_words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_words) bb[i] = (i & 7u) < 3u ? 0xffffffffu : 0u;
}

template <bool TGT>
__global__ void k_cloud_bbox(DevEngine e) {
  const uint32_t c = blockIdx.y;
  const float* xyz = TGT ? e.tgt_raw + 3 * (size_t)e.tgts[c].raw_off : e.src_raw + 3 * (size_t)e.srcs[c].off;
  const uint32_t n = TGT ? e.tgts[c].n : e.srcs[c].n;
  uint32_t* bb = e.bbox_tmp + 8 * (size_t)(TGT ? c : e.n_tgt + c);
  float l[3] = {3.0e38f, 3.0e38f, 3.0e38f}, h[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
  uint32_t cnt = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) continue;
    l[0] = fminf(l[0], x);
    l[1] = fminf(l[1], y);
    l[2] = fminf(l[2], z);
    h[0] = fmaxf(h[0], x);
    h[1] = fmaxf(h[1], y);
    h[2] = fmaxf(h[2], z);
    ++cnt;
  }
  for (int a = 0; a < 3; ++a) {
    l[a] = wave_min_f32(l[a]);
    h[a] = wave_max_f32(h[a]);
  }
  cnt = wave_sum_u32(cnt);
  // the workgroup's waves combine in LDS first: one set of atomics per
  // workgroup instead of per wave (every workgroup of a cloud updates the
  // same seven words, so these atomics serialise)
  __shared__ float s_l[kBlock / 64][3], s_h[kBlock / 64][3];
  __shared__ uint32_t s_c[kBlock / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int a = 0; a < 3; ++a) {
      s_l[w][a] = l[a];
      s_h[w][a] = h[a];
    }
    s_c[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 1; v < kBlock / 64; ++v) {
      for (int a = 0; a < 3; ++a) {
        l[a] = fminf(l[a], s_l[v][a]);
        h[a] = fmaxf(h[a], s_h[v][a]);
      }
      cnt += s_c[v];
    }
    if (cnt) {
      for (int a = 0; a < 3; ++a) {
        atomicMin(&bb[a], ord_key(l[a]));
        atomicMax(&bb[3 + a], ord_key(h[a]));
      }
      atomicAdd(&bb[6], cnt);
    }
  }
}

__device__ __forceinline__ void bbox_of(const uint32_t* bb, float* lo, float* hi, uint32_t* cnt) {
  *cnt = bb[6];
  for (int a = 0; a < 3; ++a) {
    lo[a] = *cnt ? ord_val(bb[a]) : 3.0e38f;
    hi[a] = *cnt ? ord_val(bb[3 + a]) : -3.0e38f;
  }
}

// ---------------------------------------------------------------------------
// target grids
// ---------------------------------------------------------------------------

// Choose the cell edge: the smallest h >= h_req (growing by 5%) whose grid
// over the bbox fits the cell budget.  dims use the same float expression as
// the cell assignment, so every point lands inside the grid.
__device__ void choose_grid(TargetDesc& t, const float* lo, const float* hi, uint32_t cnt) {
  t.n_valid = cnt;
  if (cnt == 0) {
    t.dx = t.dy = t.dz = 1;
    t.ox = t.oy = t.oz = 0.f;
    t.h = t.h_req > 0.f ? t.h_req : 1.f;
    t.inv_h = 1.0f / t.h;
    t.margin = 1e-3f * t.h;
    t.anchor[0] = t.anchor[1] = t.anchor[2] = 0.f;
    return;
  }
  const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  float h = t.h_req > 0.f ? t.h_req : 1e-3f;
  const double vol = (double)ex * ey * ez;
  const double hb = cbrt(vol / (double)t.cell_budget);
  if ((float)hb > h) h = (float)hb;
  int dx = 1, dy = 1, dz = 1;
  float inv = 1.0f / h;
  bool fits = false;
  if (isfinite(ex) && isfinite(ey) && isfinite(ez)) {
    for (int it = 0; it < 400; ++it) {
      inv = 1.0f / h;
      const float fx = floorf((hi[0] - lo[0]) * inv), fy = floorf((hi[1] - lo[1]) * inv),
                  fz = floorf((hi[2] - lo[2]) * inv);
      if (fx < 1.0e7f && fy < 1.0e7f && fz < 1.0e7f) {
        dx = (int)fx + 1;
        dy = (int)fy + 1;
        dz = (int)fz + 1;
        if ((long long)dx * dy * dz <= (long long)t.cell_budget) {
          fits = true;
          break;
        }
      }
      h *= 1.05f;
    }
  }
  if (!fits) {
    // a single cell holds everything (cell_of clamps every index to 0): slow, exact
    dx = dy = dz = 1;
    h = fmaxf(fmaxf(ex, ey), ez);
    if (!(h > 0.f) || !isfinite(h)) h = 1.0f;
    inv = 1.0f / h;
  }
  t.dx = dx;
  t.dy = dy;
  t.dz = dz;
  t.ox = lo[0];
  t.oy = lo[1];
  t.oz = lo[2];
  t.h = h;
  t.inv_h = inv;
  float mag = fmaxf(fmaxf(fabsf(lo[0]), fabsf(hi[0])), fmaxf(fmaxf(fabsf(lo[1]), fabsf(hi[1])),
                                                             fmaxf(fabsf(lo[2]), fabsf(hi[2]))));
  t.margin = 1e-3f * h + 4e-6f * (mag + fmaxf(ex, fmaxf(ey, ez)));
  t.anchor[0] = lo[0] + 0.5f * ex;
  t.anchor[1] = lo[1] + 0.5f * ey;
  t.anchor[2] = lo[2] + 0.5f * ez;
}

// Distance-field parameters: a query whose cell is at Chebyshev distance >= R
// from every occupied cell is at least (R - 1) * h - margin from every target
// point; skip_ring is the smallest R for which that exceeds r_max.
__device__ void choose_dist_cap(TargetDesc& t, float r_max) {
  dist_caps(r_max, t.h, t.margin, &t.skip_ring, &t.dt_cap);
}

// one thread per target, after k_cloud_bbox<true>
__global__ void k_tgt_choose(DevEngine e) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= e.n_tgt) return;
  TargetDesc& t = e.tgts[j];
  float lo[3], hi[3];
  uint32_t cnt;
  bbox_of(e.bbox_tmp + 8 * (size_t)j, lo, hi, &cnt);
  choose_grid(t, lo, hi, cnt);
  choose_dist_cap(t, e.r_max);
}

__device__ __forceinline__ bool cell_of(const TargetDesc& t, float x, float y, float z,
                                        uint32_t* c) {
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) return false;
  int ix = clampi((int)floorf((x - t.ox) * t.inv_h), 0, t.dx - 1);
  int iy = clampi((int)floorf((y - t.oy) * t.inv_h), 0, t.dy - 1);
  int iz = clampi((int)floorf((z - t.oz) * t.inv_h), 0, t.dz - 1);
  *c = ((uint32_t)iz * t.dy + iy) * t.dx + ix;
  return true;
}

__global__ void k_tgt_count(DevEngine e) {
  const TargetDesc& t = e.tgts[blockIdx.y];
  const float* xyz = e.tgt_raw + 3 * (size_t)t.raw_off;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x) {
    uint32_t c;
    if (cell_of(t, xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], &c)) {
      e.cell_id[t.raw_off + i] = c;
      e.cell_rank[t.raw_off + i] = atomicAdd(&e.cell_count[t.cell_off + c], 1u);
    } else {
      e.cell_id[t.raw_off + i] = 0xffffffffu;
    }
  }
}

__global__ void k_tgt_scatter(DevEngine e) {
  const TargetDesc& t = e.tgts[blockIdx.y];
  const float* xyz = e.tgt_raw + 3 * (size_t)t.raw_off;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x) {
    const uint32_t c = e.cell_id[t.raw_off + i];
    if (c == 0xffffffffu) continue;
    const uint32_t slot = t.cell_off + c;
    // the count pass handed every point its slot within the cell
    const uint32_t pos = e.cell_start[slot] + e.cell_rank[t.raw_off + i];
    union {
      int32_t i;
      float f;
    } w;
    w.i = (int32_t)i;
    e.tgt_sorted[pos] =
        make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], w.f);
    if (e.tgt_nrm_raw && e.tgt_nrm_sorted) {
      const float* nr = e.tgt_nrm_raw + 3 * ((size_t)t.raw_off + i);
      e.tgt_nrm_sorted[pos] = make_float4(nr[0], nr[1], nr[2], 0.f);
    }
  }
}

// Separable Chebyshev distance transform over each target's cells, capped at
// dt_cap + 1: pass x from the occupancy (cell_start differences), then y, then
// z, each a min over |d| <= cap of max(|d|, previous pass at offset d).  The
// offsets go outward by |d| and stop once |d| reaches the best value so far:
// no farther offset can give less than its |d|.
template <int AXIS>
__global__ void k_tgt_dist(DevEngine e, const uint8_t* src, uint8_t* dst) {
  const TargetDesc& t = e.tgts[blockIdx.y];
  const uint32_t ncell = (uint32_t)t.dx * t.dy * t.dz;
  const int K = t.dt_cap;
  const int len = AXIS == 0 ? t.dx : (AXIS == 1 ? t.dy : t.dz);
  const int stride = AXIS == 0 ? 1 : (AXIS == 1 ? t.dx : t.dx * t.dy);
  const uint32_t* cs = e.cell_start + t.cell_off;
  auto val = [&](uint32_t cc) -> int {
    if (AXIS == 0) return cs[cc + 1] > cs[cc] ? 0 : K + 1;
    return src[t.cell_off + cc];
  };
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
    const int pos = AXIS == 0 ? (int)(c % t.dx) : (AXIS == 1 ? (int)((c / t.dx) % t.dy) : (int)(c / ((uint32_t)t.dx * t.dy)));
    int best = val(c);
    const int lo = min(K, pos), hi = min(K, len - 1 - pos);
    for (int a = 1; a < best && (a <= lo || a <= hi); ++a) {
      int v = K + 1;
      if (a <= lo) v = val(c - (uint32_t)(a * stride));
      if (a <= hi) {
        const int w = val(c + (uint32_t)(a * stride));
        v = w < v ? w : v;
      }
      v = v > a ? v : a;
      best = v < best ? v : best;
    }
    dst[t.cell_off + c] = (uint8_t)best;
  }
}

// workgroups per cloud of a bounding-box pass: about 16 points per thread
static unsigned bbox_blocks_for(int max_pts) {
  unsigned b = (unsigned)((max_pts + 16 * kBlock - 1) / (16 * kBlock));
  return b < 1 ? 1 : (b > 128 ? 128 : b);
}

static unsigned blocks_for(int max_pts) {
  unsigned b = (unsigned)((max_pts + kBlock - 1) / kBlock);
  if (b < 1) b = 1;
  if (b > 256) b = 256;
  return b;
}

void launch_target_grid(const DevEngine& e, int max_tgt_pts, int max_cells, hipStream_t s) {
  if (e.n_tgt == 0) return;
  launch_fill_u32(e.cell_count, 0u, e.n_cells_total, s);
  const uint32_t bw = 8u * (uint32_t)e.n_tgt;
  hipLaunchKernelGGL(k_bbox_init, dim3((bw + kBlock - 1) / kBlock), dim3(kBlock), 0, s, e.bbox_tmp, bw);
  dim3 g(blocks_for(max_tgt_pts), e.n_tgt);
  hipLaunchKernelGGL(k_cloud_bbox<true>, dim3(bbox_blocks_for(max_tgt_pts), e.n_tgt), dim3(kBlock), 0, s, e);
  hipLaunchKernelGGL(k_tgt_choose, dim3((e.n_tgt + 63) / 64), dim3(64), 0, s, e);
  hipLaunchKernelGGL(k_tgt_count, g, dim3(kBlock), 0, s, e);
  launch_exclusive_scan(e.cell_count, e.cell_start, e.n_cells_total, e.scan_tmp, s);
  hipLaunchKernelGGL(k_tgt_scatter, g, dim3(kBlock), 0, s, e);
  if (e.cell_dist) {
    dim3 gd(blocks_for(max_cells), e.n_tgt);
    hipLaunchKernelGGL(k_tgt_dist<0>, gd, dim3(kBlock), 0, s, e, (const uint8_t*)nullptr, e.cell_dist);
    hipLaunchKernelGGL(k_tgt_dist<1>, gd, dim3(kBlock), 0, s, e, (const uint8_t*)e.cell_dist, e.cell_dist_tmp);
    hipLaunchKernelGGL(k_tgt_dist<2>, gd, dim3(kBlock), 0, s, e, (const uint8_t*)e.cell_dist_tmp, e.cell_dist);
  }
}

// ---------------------------------------------------------------------------
// source spatial sort (Morton order of a coarse grid in the source frame)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  // 5-bit value -> bits at positions 0, 3, 6, 9, 12
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 5; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

// one thread per source, after k_cloud_bbox<false>
__global__ void k_src_choose(DevEngine e) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= e.n_src) return;
  SourceDesc& s = e.srcs[j];
  float lo[3], hi[3];
  uint32_t cnt;
  bbox_of(e.bbox_tmp + 8 * (size_t)(e.n_tgt + j), lo, hi, &cnt);
  {
    // bin_budget = P^3 with P a power of two <= 32 (chosen by the host)
    int P = 1;
    while (P * P * P < (int)s.bin_budget) P <<= 1;
    s.dx = s.dy = s.dz = P;
    float ext = cnt ? fmaxf(hi[0] - lo[0], fmaxf(hi[1] - lo[1], hi[2] - lo[2])) : 0.f;
    s.ox = cnt ? lo[0] : 0.f;
    s.oy = cnt ? lo[1] : 0.f;
    s.oz = cnt ? lo[2] : 0.f;
    s.inv_h = ext > 0.f ? (float)P / ext : 0.f;
  }
}

__device__ __forceinline__ uint32_t bin_of(const SourceDesc& s, float x, float y, float z) {
  const uint32_t last = s.bin_budget - 1;
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) return last;
  const int P = s.dx;
  int ix = clampi((int)floorf((x - s.ox) * s.inv_h), 0, P - 1);
  int iy = clampi((int)floorf((y - s.oy) * s.inv_h), 0, P - 1);
  int iz = clampi((int)floorf((z - s.oz) * s.inv_h), 0, P - 1);
  uint32_t m = spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
  return m > last ? last : m;
}

__global__ void k_src_count(DevEngine e) {
  const SourceDesc& s = e.srcs[blockIdx.y];
  const float* xyz = e.src_raw + 3 * (size_t)s.off;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    uint32_t b = bin_of(s, xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]);
    e.bin_id[s.off + i] = b;
    e.bin_rank[s.off + i] = atomicAdd(&e.bin_count[s.bin_off + b], 1u);
  }
}

__global__ void k_src_scatter(DevEngine e) {
  const SourceDesc& s = e.srcs[blockIdx.y];
  const float* xyz = e.src_raw + 3 * (size_t)s.off;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    const uint32_t slot = s.bin_off + e.bin_id[s.off + i];
    // bin_start is global over all sources; every source point is placed, so
    // source s occupies exactly [s.off, s.off + s.n) of the sorted arrays
    const uint32_t pos = e.bin_start[slot] + e.bin_rank[s.off + i];
    e.sx[pos] = xyz[3 * (size_t)i];
    e.sy[pos] = xyz[3 * (size_t)i + 1];
    e.sz[pos] = xyz[3 * (size_t)i + 2];
    e.perm[pos] = i;
  }
}

// The scatter fills each bin in atomic order.  One wave per bin then orders
// the bin's points by (Morton code of their position inside the bin at 16
// sub-cells per axis, original index): the layout, and with it the FP32
// summation order of the moments, is identical from run to run, and
// consecutive points stay spatially compact below the bin size.  Ranks are
// found by counting (O(n^2 / 64) per bin, n ~ tens); a bin larger than the
// wave's LDS tile (1024 points) is ordered by original index alone by one
// lane.  Only the keys are staged (8 KB per wave, 20 waves per CU): a point is
// read back from the unsorted input by its original index when it is written
// to its rank.
constexpr int kBinCap = 1024;

__device__ __forceinline__ uint32_t spread3_4(uint32_t v) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

// the source owning bin b (bins of consecutive sources are consecutive)
__device__ __forceinline__ int source_of_bin(const DevEngine& e, uint32_t b) {
  int lo = 0, hi = e.n_src - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (e.srcs[mid].bin_off <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// (Morton code of the point inside its bin at 16 sub-cells per axis, original index)
__device__ __forceinline__ uint64_t fine_key(const SourceDesc& sd, float x, float y, float z, uint32_t id) {
  uint32_t fine = 0xfffu;
  if (isfinite(x) && isfinite(y) && isfinite(z)) {
    const int P = sd.dx;
    const float f[3] = {(x - sd.ox) * sd.inv_h, (y - sd.oy) * sd.inv_h, (z - sd.oz) * sd.inv_h};
    uint32_t sub[3];
    for (int a = 0; a < 3; ++a) {
      const int c = clampi((int)floorf(f[a]), 0, P - 1);
      sub[a] = (uint32_t)clampi((int)floorf((f[a] - (float)c) * 16.0f), 0, 15);
    }
    fine = spread3_4(sub[0]) | (spread3_4(sub[1]) << 1) | (spread3_4(sub[2]) << 2);
  }
  return ((uint64_t)fine << 32) | id;
}

// One wave orders one bin.
__device__ __forceinline__ void order_bin(const DevEngine& e, uint32_t b, uint32_t total, int lane,
                                          uint64_t* s_key) {
  const uint32_t beg = e.bin_start[b];
  const uint32_t end = (b + 1 < e.n_bins_total) ? e.bin_start[b + 1] : total;
  const uint32_t n = end - beg;
  if (n <= 1) return;
  if (n > (uint32_t)kBinCap) {
    if (lane == 0) {
      for (uint32_t i = beg + 1; i < end; ++i) {
        const uint32_t key = e.perm[i];
        const float x = e.sx[i], y = e.sy[i], z = e.sz[i];
        uint32_t j = i;
        while (j > beg && e.perm[j - 1] > key) {
          e.perm[j] = e.perm[j - 1];
          e.sx[j] = e.sx[j - 1];
          e.sy[j] = e.sy[j - 1];
          e.sz[j] = e.sz[j - 1];
          --j;
        }
        e.perm[j] = key;
        e.sx[j] = x;
        e.sy[j] = y;
        e.sz[j] = z;
      }
    }
    return;
  }
  const SourceDesc sd = e.srcs[source_of_bin(e, b)];
  for (uint32_t i = lane; i < n; i += 64) s_key[i] = fine_key(sd, e.sx[beg + i], e.sy[beg + i], e.sz[beg + i], e.perm[beg + i]);
  __builtin_amdgcn_wave_barrier();
  const float* raw = e.src_raw + 3 * (size_t)sd.off;
  for (uint32_t i = lane; i < n; i += 64) {
    const uint64_t k = s_key[i];
    uint32_t r = 0;
    for (uint32_t j = 0; j < n; ++j) r += (s_key[j] < k) ? 1u : 0u;
    const uint32_t id = (uint32_t)k;
    e.sx[beg + r] = raw[3 * (size_t)id];
    e.sy[beg + r] = raw[3 * (size_t)id + 1];
    e.sz[beg + r] = raw[3 * (size_t)id + 2];
    e.perm[beg + r] = id;
  }
  __builtin_amdgcn_wave_barrier();  // the LDS slots are reused by the wave's next bin
}

// One wave orders the 64 consecutive bins b0 .. b0 + nb - 1 of one source
// together, their points at positions [gbeg, gbeg + N) with N <= kBinCap:
// each point is ranked against the keys of its own bin only, so the result is
// the per-bin order, but the wave pays one load-rank-store chain per group
// instead of one per bin.  st = this lane's bin start, relative to gbeg
// (N for lanes past the group).
__device__ __forceinline__ void order_group(const DevEngine& e, const SourceDesc& sd, uint32_t gbeg, uint32_t N,
                                            uint32_t st, int lane, uint64_t* s_key, uint32_t* s_start,
                                            uint8_t* s_bin) {
  s_start[lane] = st;
  if (lane == 0) s_start[64] = N;
  __builtin_amdgcn_wave_barrier();
  for (uint32_t p = lane; p < N; p += 64) {
    // the bin holding position p: the last k with s_start[k] <= p (k < 64 as p < N)
    int k = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
      if (s_start[k + step] <= p) k += step;
    s_bin[p] = (uint8_t)k;
    s_key[p] = fine_key(sd, e.sx[gbeg + p], e.sy[gbeg + p], e.sz[gbeg + p], e.perm[gbeg + p]);
  }
  __builtin_amdgcn_wave_barrier();
  const float* raw = e.src_raw + 3 * (size_t)sd.off;
  for (uint32_t p = lane; p < N; p += 64) {
    const int k = s_bin[p];
    const uint32_t lo = s_start[k], hi = s_start[k + 1];
    const uint64_t key = s_key[p];
    uint32_t r = 0;
    for (uint32_t j = lo; j < hi; ++j) r += (s_key[j] < key) ? 1u : 0u;
    const uint32_t id = (uint32_t)key;
    const uint32_t pos = gbeg + lo + r;
    e.sx[pos] = raw[3 * (size_t)id];
    e.sy[pos] = raw[3 * (size_t)id + 1];
    e.sz[pos] = raw[3 * (size_t)id + 2];
    e.perm[pos] = id;
  }
  __builtin_amdgcn_wave_barrier();  // the LDS arrays are reused by the wave's next group
}

constexpr uint32_t kOrderBlocks = 16384;

// One wave per workgroup (8 KB of LDS each).  Waves stride over groups of 64
// bins: the lanes read the 64 bins' sizes at once, and the wave then orders
// the bins of more than one point one after another (most bins of a sparse
// surface hold none or one, so they cost one coalesced read).
__global__ __launch_bounds__(64) void k_src_order_bins(DevEngine e, uint32_t total) {
  __shared__ uint64_t s_key[kBinCap];
  __shared__ uint32_t s_start[65];
  __shared__ uint8_t s_bin[kBinCap];
  const int lane = threadIdx.x & 63;
  for (uint32_t b0 = blockIdx.x * 64u; b0 < e.n_bins_total; b0 += gridDim.x * 64u) {
    const uint32_t b = b0 + (uint32_t)lane;
    const uint32_t nb = min(64u, e.n_bins_total - b0);
    uint32_t st = 0u, n = 0u;
    if (b < e.n_bins_total) {
      st = e.bin_start[b];
      n = ((b + 1 < e.n_bins_total) ? e.bin_start[b + 1] : total) - st;
    }
    uint64_t todo = __ballot(n > 1u);
    if (!todo) continue;
    // the whole group at once when it fits the LDS tile and one source owns it
    const uint32_t gbeg = __shfl(st, 0);
    const uint32_t gend = __shfl(st + n, (int)nb - 1);
    const int si = source_of_bin(e, b0);
    const SourceDesc& sd = e.srcs[si];
    if (gend - gbeg <= (uint32_t)kBinCap && b0 + nb <= sd.bin_off + sd.bin_budget) {
      order_group(e, sd, gbeg, gend - gbeg, (uint32_t)lane < nb ? st - gbeg : gend - gbeg, lane, s_key, s_start,
                  s_bin);
      continue;
    }
    while (todo) {
      const uint32_t l = (uint32_t)__builtin_ctzll(todo);
      todo &= todo - 1ull;
      order_bin(e, b0 + l, total, lane, s_key);
    }
  }
}

void launch_source_sort(const DevEngine& e, int max_src_pts, uint32_t src_total, hipStream_t s) {
  if (e.n_src == 0) return;
  launch_fill_u32(e.bin_count, 0u, e.n_bins_total, s);
  const uint32_t bw = 8u * (uint32_t)e.n_src;
  hipLaunchKernelGGL(k_bbox_init, dim3((bw + kBlock - 1) / kBlock), dim3(kBlock), 0, s, e.bbox_tmp + 8 * (size_t)e.n_tgt,
                     bw);
  hipLaunchKernelGGL(k_cloud_bbox<false>, dim3(bbox_blocks_for(max_src_pts), e.n_src), dim3(kBlock), 0, s, e);
  hipLaunchKernelGGL(k_src_choose, dim3((e.n_src + 63) / 64), dim3(64), 0, s, e);
  dim3 g(blocks_for(max_src_pts), e.n_src);
  hipLaunchKernelGGL(k_src_count, g, dim3(kBlock), 0, s, e);
  launch_exclusive_scan(e.bin_count, e.bin_start, e.n_bins_total, e.scan_tmp, s);
  hipLaunchKernelGGL(k_src_scatter, g, dim3(kBlock), 0, s, e);
  const uint32_t ob = (e.n_bins_total + 63) / 64;
  hipLaunchKernelGGL(k_src_order_bins, dim3(ob < kOrderBlocks ? ob : kOrderBlocks), dim3(64), 0, s, e, src_total);
}

}  // namespace rtr
