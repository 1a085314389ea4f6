// Learned pose proposer: farthest-point sampling, a shared point MLP
// 3-64-128-256 with max-pooling per cloud, and a head 512-256-9 (6D rotation +
// translation) over the concatenated source/target features.  Every matrix
// product runs in one hand-written BF16 MFMA kernel (v_mfma_f32_32x32x8bf16_1k,
// FP32 accumulation); forward and backward are compositions of that kernel
// and small elementwise kernels.  Trained with Adam on poses that the GPU ICP
// converges to.
//
// The reference's counterpart is its hand-crafted coarse alignment (keypoints,
// plane areas, TDF yaw search and exhaustive RANSAC, RealTimeRobot.cpp:45-105,
// function.h:35-109): both produce the initial pose that ICP refines.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rtr.h"
#include "engine.hpp"
#include "internal.h"
#include "wave.hpp"

namespace rtr {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (NaN kept quiet)
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  const uint32_t r = 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)((u + r) >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

constexpr int kP = 1024;      // points per cloud after FPS
constexpr int kC0 = 16;       // input channels (xyz padded to the MFMA K granularity)
constexpr int kC1 = 64, kC2 = 128, kC3 = 256, kH = 256, kOut = 9;
constexpr int kFpsMax = 8192; // stride pre-subsample before FPS
constexpr int kMaxHyp = 64;   // hypothesis rotations per proposal call

// parameter layout (FP32 master and BF16 copy share offsets)
struct ParamLayout {
  size_t W1, b1, W2, b2, W3, b3, W4, b4, W5, b5, total;
  ParamLayout() {
    size_t o = 0;
    W1 = o; o += (size_t)kC0 * kC1;
    b1 = o; o += kC1;
    W2 = o; o += (size_t)kC1 * kC2;
    b2 = o; o += kC2;
    W3 = o; o += (size_t)kC2 * kC3;
    b3 = o; o += kC3;
    W4 = o; o += (size_t)2 * kC3 * kH;
    b4 = o; o += kH;
    W5 = o; o += (size_t)kH * kOut;
    b5 = o; o += kOut;
    total = o;
  }
};

// ---------------------------------------------------------------------------
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N] (+ bias[N]) (ReLU)
//   A(i,k) = ta ? A[k*lda + i] : A[i*lda + k]   (bf16)
//   B(k,j) = tb ? B[j*ldb + k] : B[k*ldb + j]   (bf16)
// Output FP32 (atomicAdd when split-K) or BF16.  64 x 64 tile per workgroup
// (2 x 2 waves of 32 x 32 MFMA tiles), K staged through LDS 32 at a time.
// ---------------------------------------------------------------------------
struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  void* C;
  const float* bias;
  int M, N, K, lda, ldb, ldc;
  int ta, tb, relu, out_bf16, atomic_out, k_split;
};

constexpr int kTM = 64, kTN = 64, kTK = 32;

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  // rows padded by 4 elements (8 B): 8-byte aligned bf16x4 operand reads
  __shared__ __align__(16) uint16_t sA[kTM][kTK + 4];
  __shared__ __align__(16) uint16_t sBt[kTN][kTK + 4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
  // split-K range
  const int kchunk = ((g.K + g.k_split - 1) / g.k_split + kTK - 1) / kTK * kTK;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(g.K, kb + kchunk);
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += kTK) {
    // stage A (64 x 32) and B (32 x 64, stored transposed): 2048 elements
    // each, 8 per thread; consecutive threads walk the operand's contiguous
    // dimension
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + e * 256;
      int r, c;  // A tile row r (M), column c (K)
      if (g.ta) {
        r = idx % kTM;
        c = idx / kTM;
      } else {
        r = idx / kTK;
        c = idx % kTK;
      }
      const int gi = m0 + r, gk = k0 + c;
      uint16_t v = 0;
      if (gi < g.M && gk < ke) v = g.ta ? g.A[(size_t)gk * g.lda + gi] : g.A[(size_t)gi * g.lda + gk];
      sA[r][c] = v;
      int rk, cn;  // B tile row rk (K), column cn (N)
      if (g.tb) {
        rk = idx % kTK;
        cn = idx / kTK;
      } else {
        rk = idx / kTN;
        cn = idx % kTN;
      }
      const int gk2 = k0 + rk, gj = n0 + cn;
      uint16_t u = 0;
      if (gk2 < ke && gj < g.N) u = g.tb ? g.B[(size_t)gj * g.ldb + gk2] : g.B[(size_t)gk2 * g.ldb + gj];
      sBt[cn][rk] = u;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 8) {
      const int kq = kk + (lane >> 5) * 4;
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(&sA[wm * 32 + (lane & 31)][kq]);
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(&sBt[wn * 32 + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: lane l, register j -> row (j/4)*8 + (l/32)*4 + j%4, column l%32
  const int col = n0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int row = m0 + wm * 32 + (j >> 2) * 8 + (lane >> 5) * 4 + (j & 3);
    if (row >= g.M || col >= g.N) continue;
    float v = acc[j];
    if (g.bias && blockIdx.z == 0) v += g.bias[col];
    if (g.relu) v = v > 0.f ? v : 0.f;
    const size_t o = (size_t)row * g.ldc + col;
    if (g.out_bf16)
      reinterpret_cast<uint16_t*>(g.C)[o] = f2bf(v);
    else if (g.atomic_out)
      atomicAdd(reinterpret_cast<float*>(g.C) + o, v);
    else
      reinterpret_cast<float*>(g.C)[o] = v;
  }
}

static void gemm(hipStream_t s, const uint16_t* A, int lda, int ta, const uint16_t* B, int ldb, int tb, void* C,
                 int ldc, int M, int N, int K, const float* bias, int relu, int out_bf16, int k_split = 1) {
  GemmArgs g;
  g.A = A;
  g.B = B;
  g.C = C;
  g.bias = bias;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.ta = ta;
  g.tb = tb;
  g.relu = relu;
  g.out_bf16 = out_bf16;
  g.atomic_out = k_split > 1;
  g.k_split = k_split;
  dim3 grid((M + kTM - 1) / kTM, (N + kTN - 1) / kTN, k_split);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, g);
}

// ---------------------------------------------------------------------------
// Fused point MLP + max-pool (inference)
//
// Rows of a cloud go through 16 -> 64 -> 128 -> 256 with bias + ReLU +
// BF16 rounding after every layer and only the per-channel maxima leave the
// chip.  Each layer is the same sequence of v_mfma_f32_32x32x8bf16_1k steps
// (k ascending) as k_gemm_bf16 and rounds where the layer-by-layer path
// stores BF16, so the pooled features are bit-identical to it.  The
// transposed weights sit in LDS for the workgroup's lifetime; each wave owns
// a 32-row activation tile in LDS; workgroups stride over clouds.
// ---------------------------------------------------------------------------
constexpr int kFWaves = 8;
constexpr int kFThreads = kFWaves * 64;
constexpr int kFPad = 4;  // bf16 elements of row padding (8-byte aligned, spreads LDS banks)

struct FusedShared {
  uint16_t w1t[kC1][kC0 + kFPad];
  uint16_t w2t[kC2][kC1 + kFPad];
  uint16_t w3t[kC3][kC2 + kFPad];
  uint16_t act[kFWaves][32][kC2 + kFPad];
  float b1[kC1], b2[kC2], b3[kC3];
  uint32_t cmax[kC3];
};

__device__ __forceinline__ bf16x4 lds4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

// BF16 of ReLU(v) by the hardware conversion (v_cvt_pk_bf16_f32, round to
// nearest even): equal to f2bf for every non-NaN float, and the ReLU maps NaN
// to 0 first
__device__ __forceinline__ uint16_t bf16_relu(float v) {
  return __builtin_bit_cast(uint16_t, (__bf16)(v > 0.f ? v : 0.f));
}

// The hardware conversion against f2bf over every non-NaN float: 16 bit
// patterns per thread, 2^20 workgroups of 256 cover all 2^32.
__global__ __launch_bounds__(256) void k_bf16_check(unsigned long long* bad) {
  const uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16ull;
  unsigned long long b = 0;
  for (int i = 0; i < 16; ++i) {
    const uint32_t u = (uint32_t)(base + (uint64_t)i);
    if ((u & 0x7fffffffu) > 0x7f800000u) continue;
    const float f = __uint_as_float(u);
    b += __builtin_bit_cast(uint16_t, (__bf16)f) != f2bf(f);
  }
  if (b) atomicAdd(bad, b);
}

// acc row of register j for this lane (32x32 MFMA output layout)
__device__ __forceinline__ int mfma_row(int lane, int j) { return (j >> 2) * 8 + (lane >> 5) * 4 + (j & 3); }

__global__ __launch_bounds__(kFThreads) void k_mlp_pool(const uint16_t* x0, int n_clouds, const uint16_t* wb,
                                                       const float* w, ParamLayout L, float* G) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  FusedShared& sh = *reinterpret_cast<FusedShared*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // transposed weights (B operands read 4 consecutive k per lane) and biases
  for (int i = tid; i < kC0 * kC1; i += kFThreads) sh.w1t[i % kC1][i / kC1] = wb[L.W1 + i];
  for (int i = tid; i < kC1 * kC2; i += kFThreads) sh.w2t[i % kC2][i / kC2] = wb[L.W2 + i];
  for (int i = tid; i < kC2 * kC3; i += kFThreads) sh.w3t[i % kC3][i / kC3] = wb[L.W3 + i];
  for (int i = tid; i < kC1; i += kFThreads) sh.b1[i] = w[L.b1 + i];
  for (int i = tid; i < kC2; i += kFThreads) sh.b2[i] = w[L.b2 + i];
  for (int i = tid; i < kC3; i += kFThreads) sh.b3[i] = w[L.b3 + i];
  uint16_t(*act)[kC2 + kFPad] = sh.act[wv];
  const int lr = lane & 31, kh = (lane >> 5) * 4;
  for (int c = blockIdx.x; c < n_clouds; c += gridDim.x) {
    for (int i = tid; i < kC3; i += kFThreads) sh.cmax[i] = 0u;
    __syncthreads();
    float vmax[kC3 / 32];
#pragma unroll
    for (int t = 0; t < kC3 / 32; ++t) vmax[t] = 0.f;
    for (int r0 = wv * 32; r0 < kP; r0 += kFWaves * 32) {
      const uint16_t* xr = x0 + ((size_t)c * kP + r0 + lr) * kC0;
      // layer 1: K = 16 (two k-steps), N = 64 (two tiles)
      const bf16x4 a0 = *reinterpret_cast<const bf16x4*>(xr + kh);
      const bf16x4 a1 = *reinterpret_cast<const bf16x4*>(xr + 8 + kh);
#pragma unroll
      for (int nt = 0; nt < kC1 / 32; ++nt) {
        f32x16 acc;
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = 0.f;
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a0, lds4(&sh.w1t[nt * 32 + lr][kh]), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, lds4(&sh.w1t[nt * 32 + lr][8 + kh]), acc, 0, 0, 0);
        const int col = nt * 32 + lr;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const float v = acc[j] + sh.b1[col];
          act[mfma_row(lane, j)][col] = bf16_relu(v);
        }
      }
      __builtin_amdgcn_wave_barrier();
      // layer 2: K = 64 (eight k-steps), N = 128 (four tiles)
      {
        f32x16 acc[kC2 / 32];
#pragma unroll
        for (int nt = 0; nt < kC2 / 32; ++nt)
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[nt][j] = 0.f;
#pragma unroll
        for (int ks = 0; ks < kC1 / 8; ++ks) {
          const bf16x4 a = lds4(&act[lr][ks * 8 + kh]);
#pragma unroll
          for (int nt = 0; nt < kC2 / 32; ++nt)
            acc[nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, lds4(&sh.w2t[nt * 32 + lr][ks * 8 + kh]), acc[nt], 0,
                                                               0, 0);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int nt = 0; nt < kC2 / 32; ++nt) {
          const int col = nt * 32 + lr;
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            const float v = acc[nt][j] + sh.b2[col];
            act[mfma_row(lane, j)][col] = bf16_relu(v);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      // layer 3: K = 128 (sixteen k-steps), N = 256 in two halves of four tiles
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        f32x16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[t][j] = 0.f;
#pragma unroll 4
        for (int ks = 0; ks < kC2 / 8; ++ks) {
          const bf16x4 a = lds4(&act[lr][ks * 8 + kh]);
#pragma unroll
          for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, lds4(&sh.w3t[(half * 4 + t) * 32 + lr][ks * 8 + kh]),
                                                              acc[t], 0, 0, 0);
        }
        // the pooled value is the BF16 rounding of the largest ReLU output:
        // rounding is monotone, so rounding the maximum once equals the
        // maximum of the rounded outputs; fmaxf(m, v) with m >= 0 is the
        // ReLU (a NaN v leaves m, as ReLU maps NaN to 0)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int col = (half * 4 + t) * 32 + lr;
          const float b = sh.b3[col];
          float m = vmax[half * 4 + t];
#pragma unroll
          for (int j = 0; j < 16; ++j) m = fmaxf(m, acc[t][j] + b);
          vmax[half * 4 + t] = m;
        }
      }
      __builtin_amdgcn_wave_barrier();  // act is rewritten by the next rows
    }
    // lanes l and l + 32 hold different rows of the same channels
#pragma unroll
    for (int t = 0; t < kC3 / 32; ++t) {
      const float o = __shfl_xor(vmax[t], 32, 64);
      const float mx = fmaxf(vmax[t], o);
      const float m = bf2f(f2bf(mx > 0.f ? mx : 0.f));  // a -0 maximum becomes +0, as the ReLU gives
      if (lane < 32) atomicMax(&sh.cmax[t * 32 + lane], __float_as_uint(m));  // m >= 0: bits order as floats
    }
    __syncthreads();
    for (int i = tid; i < kC3; i += kFThreads) G[(size_t)c * kC3 + i] = __uint_as_float(sh.cmax[i]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// farthest-point sampling
// ---------------------------------------------------------------------------

// One workgroup per cloud.  The cloud is first subsampled with a fixed stride
// to at most kFpsMax points; its centroid is removed; then kP points are
// chosen greedily, starting at subsample index 0, each time the point with
// the largest distance to the chosen set (lowest index on ties).
//
// The points are regrouped spatially before the selection: a counting sort
// by 12-bit Morton cell (16^3 cells over the subsample's bounding box) puts
// them in cell order, and each run of 64 consecutive points is a group,
// compact in space, held in one register slot of one wave's 64 lanes.  A group is updated
// with the newly chosen point q only if q can lower one of its distances:
// with b the group's bounding box and M its largest distance, the update is
// skipped when the rounded squared box distance from q is >= M.  Rounding is
// monotone, so that bound is <= the rounded squared distance from q of every
// point in the box: a skipped update would have left every distance as it
// was, and the selections are exactly those of the plain O(kP * m) loop.
// Late in the selection few groups are near q, so most of the distance work
// is skipped.  Each group keeps its largest (distance, lowest index) as a
// 64-bit key in lane k; the wave's best key and its coordinates go through
// LDS to the other waves (one barrier per selection).  Where the sort puts a
// point decides only which work is skipped, never a selection.
constexpr int kFpsThreads = 512;
constexpr int kFpsPer = kFpsMax / kFpsThreads;   // register slots (groups) per lane
constexpr int kFpsWaves = kFpsThreads / 64;
constexpr int kFpsCells = 4096;                  // 16^3 Morton cells
static_assert(kFpsPer == 16 && kFpsWaves == 8 && kFpsMax == 8192, "the group layout assumes 8 waves x 16 slots");
static_assert(kFpsCells * 4 == kFpsMax * 2, "the inverse order (u16) fills the histogram's space");

// A group's distances lowered by q (the group is one register slot of all
// 64 lanes); returns the group's largest key (0: no point), the same in every
// lane.  key = (distance bits + 1) << 32 | (8191 - index) << 16 | slot.  The
// largest distance comes from a float wave max (fmaxf returns one of its
// operands, so it equals the winner's distance bit for bit); a ballot finds
// the lane holding it, and only when several lanes hold it (equal distances)
// does a second reduction pick the lowest subsample index among them.  The
// index comes from the inverse order in LDS (inv[j], j the sorted position;
// jw = j of lane 0 in this slot).
__device__ __forceinline__ uint64_t fps_group_update(float px, float py, float pz, float& dist, const uint16_t* inv,
                                                     int jw, int lane, int k, float qx, float qy, float qz,
                                                     uint64_t old) {
  // an empty asm keeps the compiler from computing every slot's distances
  // ahead of the branch that selects the slot: they are pure and would be
  // hoisted or speculated
  asm volatile("" : "+v"(qx), "+v"(qy), "+v"(qz));
  const float dx = px - qx, dy = py - qy, dz = pz - qz;
  float d = dx * dx;
  d = d + dy * dy;
  d = d + dz * dz;
  const float nd = fminf(dist, d);  // padding slots hold -1 and keep it
  const uint64_t lowered = __ballot(nd < dist);
  dist = nd;
  // distances only fall: while the lane holding the group's key keeps its
  // distance, it still holds the largest one, and among equal largest
  // distances still the smallest index, so the key stands
  if (!((lowered >> (old & 63)) & 1ull)) return old;
  const float gmax = wave_max_f32(dist);
  if (!(gmax >= 0.f)) return 0ull;  // padding only
  const uint64_t hit = __ballot(dist == gmax);
  int lb, idx;
  if (__builtin_popcountll(hit) == 1) {
    lb = __builtin_ctzll(hit);
    idx = inv[jw + lb];
  } else {
    const int v = wave_min_i32(dist == gmax ? ((int)inv[jw + lane] << 6) | lane : 0x7fffffff);
    idx = v >> 6;
    lb = v & 63;
  }
  return ((uint64_t)(__float_as_uint(gmax) + 1u) << 32) | ((uint32_t)(kFpsMax - 1 - idx) << 16) |
         (uint32_t)(k * 64 + lb);
}

// cell coordinate in [0, 15] of t = (v - lo) * 16 / extent (NaN -> 0)
__device__ __forceinline__ uint32_t fps_cell(float t) { return t >= 15.f ? 15u : (t >= 1.f ? (uint32_t)t : 0u); }

__global__ __launch_bounds__(kFpsThreads, 4) void k_fps(const float* raw, const uint32_t* off, const uint32_t* cnt,
                                                     float* pts_out /* cloud * kP * 3 */,
                                                     uint16_t* x0 /* rows of kC0 */, float* centroid /* cloud * 3 */,
                                                     int32_t* sel_out /* cloud*kP */) {
  const int c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t n = cnt[c];
  const float* p = raw + 3 * (size_t)off[c];
  const int m = (int)min((uint32_t)kFpsMax, n);
  __shared__ float s_q[3];
  // per wave: its best key and that point's coordinates; double buffered by
  // selection parity, so one barrier per selection suffices
  __shared__ unsigned long long red_k[2][kFpsWaves];
  __shared__ float red_p[2][kFpsWaves][3];
  __shared__ double cen[kFpsWaves][3];
  __shared__ float s_box[kFpsWaves][6];
  __shared__ uint32_t s_wsum[kFpsWaves];
  __shared__ uint32_t s_hist[kFpsCells];  // the cell histogram, then the inverse order
  __shared__ float s_sel[kP][3];    // chosen points, centroid removed
  __shared__ int s_cur[kP];         // their subsample indices
  uint32_t* hist = s_hist;
  uint16_t* inv = reinterpret_cast<uint16_t*>(s_hist);  // sorted position -> subsample index

  // ---- subsample i = tid + k * kFpsThreads: centroid, bounding box, cells
  uint32_t pos[kFpsPer];
  {
    float px[kFpsPer], py[kFpsPer], pz[kFpsPer];
    double ax = 0, ay = 0, az = 0;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int k = 0; k < kFpsPer; ++k) {
      const int i = tid + k * kFpsThreads;
      px[k] = py[k] = pz[k] = 0.f;
      if (i < m) {
        const size_t src = (size_t)((uint64_t)i * n / (uint64_t)m);
        px[k] = p[3 * src];
        py[k] = p[3 * src + 1];
        pz[k] = p[3 * src + 2];
        ax += px[k];
        ay += py[k];
        az += pz[k];
        lo[0] = fminf(lo[0], px[k]);
        lo[1] = fminf(lo[1], py[k]);
        lo[2] = fminf(lo[2], pz[k]);
        hi[0] = fmaxf(hi[0], px[k]);
        hi[1] = fmaxf(hi[1], py[k]);
        hi[2] = fmaxf(hi[2], pz[k]);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ax += __shfl_xor(ax, o, 64);
      ay += __shfl_xor(ay, o, 64);
      az += __shfl_xor(az, o, 64);
    }
    const float b0 = wave_min_f32(lo[0]), b1 = wave_min_f32(lo[1]), b2 = wave_min_f32(lo[2]);
    const float b3 = wave_max_f32(hi[0]), b4 = wave_max_f32(hi[1]), b5 = wave_max_f32(hi[2]);
    if (lane == 0) {
      cen[w][0] = ax;
      cen[w][1] = ay;
      cen[w][2] = az;
      s_box[w][0] = b0;
      s_box[w][1] = b1;
      s_box[w][2] = b2;
      s_box[w][3] = b3;
      s_box[w][4] = b4;
      s_box[w][5] = b5;
    }
    if (tid == 0) {  // subsample index 0 starts the selection
      s_q[0] = px[0];
      s_q[1] = py[0];
      s_q[2] = pz[0];
    }
    for (int i = tid; i < kFpsCells; i += kFpsThreads) hist[i] = 0u;
    __syncthreads();
    float blo[3], sc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float l = s_box[0][a], h = s_box[0][3 + a];
      for (int v = 1; v < kFpsWaves; ++v) {
        l = fminf(l, s_box[v][a]);
        h = fmaxf(h, s_box[v][3 + a]);
      }
      blo[a] = l;
      sc[a] = h > l ? 16.f / (h - l) : 0.f;
    }
    // cell of every point and its rank within the cell (counting sort)
#pragma unroll
    for (int k = 0; k < kFpsPer; ++k) {
      const int i = tid + k * kFpsThreads;
      pos[k] = 0u;
      if (i < m) {
        const uint32_t cx = fps_cell((px[k] - blo[0]) * sc[0]);
        const uint32_t cy = fps_cell((py[k] - blo[1]) * sc[1]);
        const uint32_t cz = fps_cell((pz[k] - blo[2]) * sc[2]);
        uint32_t code = 0u;
#pragma unroll
        for (int b = 0; b < 4; ++b)
          code |= (((cx >> b) & 1u) << (3 * b)) | (((cy >> b) & 1u) << (3 * b + 1)) | (((cz >> b) & 1u) << (3 * b + 2));
        pos[k] = (code << 16) | atomicAdd(&hist[code], 1u);
      }
    }
  }
  double sx = 0, sy = 0, sz = 0;
  for (int k = 0; k < kFpsWaves; ++k) {
    sx += cen[k][0];
    sy += cen[k][1];
    sz += cen[k][2];
  }
  const float cx = m > 0 ? (float)(sx / m) : 0.f;
  const float cy = m > 0 ? (float)(sy / m) : 0.f;
  const float cz = m > 0 ? (float)(sz / m) : 0.f;
  if (tid == 0) {
    centroid[3 * c] = cx;
    centroid[3 * c + 1] = cy;
    centroid[3 * c + 2] = cz;
  }
  __syncthreads();
  // exclusive scan of the histogram: 8 cells per thread, then the waves
  {
    uint32_t v[kFpsCells / kFpsThreads], t = 0;
#pragma unroll
    for (int j = 0; j < kFpsCells / kFpsThreads; ++j) {
      v[j] = t;
      t += hist[tid * (kFpsCells / kFpsThreads) + j];
    }
    uint32_t inc = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t base = inc - t;
    for (int v2 = 0; v2 < w; ++v2) base += s_wsum[v2];
#pragma unroll
    for (int j = 0; j < kFpsCells / kFpsThreads; ++j) hist[tid * (kFpsCells / kFpsThreads) + j] = base + v[j];
    __syncthreads();
  }
  // sorted position: valid points fill [0, m) in cell order; padding slot i
  // (i >= m) stays at i
#pragma unroll
  for (int k = 0; k < kFpsPer; ++k) {
    const int i = tid + k * kFpsThreads;
    pos[k] = i < m ? hist[pos[k] >> 16] + (pos[k] & 0xffffu) : (uint32_t)i;
  }
  __syncthreads();  // the histogram is read: its space holds the inverse order next
#pragma unroll
  for (int k = 0; k < kFpsPer; ++k) inv[pos[k]] = (uint16_t)(tid + k * kFpsThreads);
  __syncthreads();
  // ---- the group layout: slot k of thread tid holds sorted position
  // j = tid + 512 k, its point loaded again from the cloud
  float px[kFpsPer], py[kFpsPer], pz[kFpsPer], dist[kFpsPer];
  // group g = 8 k + w is slot k of wave w: sorted position j = tid + 512 k.
  // Consecutive groups (neighbours in space) sit in different waves, so the
  // few groups that one pick reaches are updated by several waves at once.
#pragma unroll
  for (int k = 0; k < kFpsPer; ++k) {
    const int j = tid + k * kFpsThreads;
    const int idx = inv[j];
    px[k] = py[k] = pz[k] = 0.f;
    dist[k] = -1.f;
    if (idx < m) {
      const size_t src = (size_t)((uint64_t)idx * n / (uint64_t)m);
      px[k] = p[3 * src];
      py[k] = p[3 * src + 1];
      pz[k] = p[3 * src + 2];
      dist[k] = 3.0e38f;
    }
  }
  // ---- groups: bounding box and largest key, group k in lane k
  float g_lo[3] = {0.f, 0.f, 0.f}, g_hi[3] = {0.f, 0.f, 0.f};
  float gbest = -1.f;
  uint64_t gkey = 0ull;
#pragma unroll
  for (int k = 0; k < kFpsPer; ++k) {
    const bool v = dist[k] >= 0.f;
    const float l0 = wave_min_f32(v ? px[k] : INFINITY), l1 = wave_min_f32(v ? py[k] : INFINITY);
    const float l2 = wave_min_f32(v ? pz[k] : INFINITY);
    const float h0 = wave_max_f32(v ? px[k] : -INFINITY), h1 = wave_max_f32(v ? py[k] : -INFINITY);
    const float h2 = wave_max_f32(v ? pz[k] : -INFINITY);
    const uint32_t klo = ((uint32_t)(kFpsMax - 1 - (int)inv[tid + k * kFpsThreads]) << 16) | (uint32_t)(k * 64 + lane);
    const uint64_t gk = wave_max_u64(v ? ((uint64_t)(__float_as_uint(dist[k]) + 1u) << 32) | klo : 0ull);
    if (lane == k) {
      g_lo[0] = l0;
      g_lo[1] = l1;
      g_lo[2] = l2;
      g_hi[0] = h0;
      g_hi[1] = h1;
      g_hi[2] = h2;
      gkey = gk;
      gbest = gk ? 3.0e38f : -1.f;
    }
  }
  // the wave's best: (key, coordinates), recomputed when a group changes
  uint64_t wkey = 0ull;
  float wx = 0.f, wy = 0.f, wz = 0.f;

  int cur = 0;
  float qx = s_q[0], qy = s_q[1], qz = s_q[2];
  for (int s = 0; s < kP; ++s) {
    if (tid == 0) {
      s_sel[s][0] = qx - cx;
      s_sel[s][1] = qy - cy;
      s_sel[s][2] = qz - cz;
      s_cur[s] = cur;
    }
    if (s + 1 == kP) break;
    // groups that q can reach: rounded squared box distance below the group's
    // largest distance (lanes 0-15; a group of padding slots only, gbest -1,
    // never changes)
    const float gx = fmaxf(fmaxf(g_lo[0] - qx, qx - g_hi[0]), 0.f);
    const float gy = fmaxf(fmaxf(g_lo[1] - qy, qy - g_hi[1]), 0.f);
    const float gz = fmaxf(fmaxf(g_lo[2] - qz, qz - g_hi[2]), 0.f);
    float lb = gx * gx;
    lb = lb + gy * gy;
    lb = lb + gz * gz;
    const uint64_t need = __ballot(lane < kFpsPer && gbest >= 0.f && !(lb >= gbest && lb < INFINITY));
    if (need) {
      // uniform branches over constant slots: the distances stay in their
      // own registers
#pragma unroll
      for (int k = 0; k < kFpsPer; ++k) {
        if (need & (1ull << k)) {
          const uint64_t key = fps_group_update(px[k], py[k], pz[k], dist[k], inv, w * 64 + k * kFpsThreads, lane, k,
                                                qx, qy, qz, readlane_u64(gkey, k));
          if (lane == k) {
            gkey = key;
            gbest = key ? __uint_as_float((uint32_t)(key >> 32) - 1u) : -1.f;
          }
        }
      }
      wkey = readlane_u64(row_max_u64(lane < kFpsPer ? gkey : 0ull), 0);
      if (wkey) {
        const int slot = (int)(wkey & 0xffffull);
        const int kb = slot >> 6, lb = slot & 63;
        // one scalar branch taken (the asm keeps the compiler from turning
        // the branches into 48 vector selects)
#pragma unroll
        for (int k = 0; k < kFpsPer; ++k)
          if (kb == k) {
            asm volatile("" ::"v"(px[k]));
            wx = readlane_f32(px[k], lb);
            wy = readlane_f32(py[k], lb);
            wz = readlane_f32(pz[k], lb);
          }
      }
    }
    const int buf = s & 1;
    if (lane == 0) {
      red_k[buf][w] = wkey;
      red_p[buf][w][0] = wx;
      red_p[buf][w][1] = wy;
      red_p[buf][w][2] = wz;
    }
    __syncthreads();
    uint64_t gk = red_k[buf][0];
    int gw = 0;
#pragma unroll
    for (int v = 1; v < kFpsWaves; ++v) {
      const uint64_t o = red_k[buf][v];
      gw = o > gk ? v : gw;
      gk = o > gk ? o : gk;
    }
    if (gk) {
      cur = kFpsMax - 1 - (int)((uint32_t)gk >> 16);
      qx = red_p[buf][gw][0];
      qy = red_p[buf][gw][1];
      qz = red_p[buf][gw][2];
    } else {  // no subsample point: repeat the first one
      cur = 0;
      qx = s_q[0];
      qy = s_q[1];
      qz = s_q[2];
    }
  }
  __syncthreads();
  for (int s = tid; s < kP; s += kFpsThreads) {
    const size_t o = (size_t)c * kP + s;
    pts_out[3 * o] = s_sel[s][0];
    pts_out[3 * o + 1] = s_sel[s][1];
    pts_out[3 * o + 2] = s_sel[s][2];
    if (sel_out) sel_out[o] = s_cur[s];
  }
  for (int t = tid; t < kP * kC0; t += kFpsThreads) {
    const int s = t / kC0, ch = t % kC0;
    x0[(size_t)c * kP * kC0 + t] = f2bf(ch < 3 ? s_sel[s][ch] : 0.f);
  }
}

// ---------------------------------------------------------------------------
// pooling, head, losses, optimiser
// ---------------------------------------------------------------------------

// Max over the kP points of each cloud, per channel; argmax = lowest row.
__global__ void k_maxpool(const uint16_t* Y, int C, int n_clouds, float* G, int32_t* arg) {
  const int c = blockIdx.x, ch = threadIdx.x + blockIdx.y * blockDim.x;
  if (ch >= C) return;
  const uint16_t* y = Y + (size_t)c * kP * C;
  float best = -3.0e38f;
  int bi = 0;
  for (int r = 0; r < kP; ++r) {
    const float v = bf2f(y[(size_t)r * C + ch]);
    if (v > best) {
      best = v;
      bi = r;
    }
  }
  G[(size_t)c * C + ch] = best;
  arg[(size_t)c * C + ch] = bi;
}

// F[pair] = (G[src cloud], G[tgt cloud]) as bf16; clouds are ordered
// (src 0, tgt 0, src 1, tgt 1, ...).
__global__ void k_concat(const float* G, int n_pairs, uint16_t* F) {
  const int p = blockIdx.x, ch = threadIdx.x;  // 512 threads
  if (p >= n_pairs) return;
  const float v = ch < kC3 ? G[(size_t)(2 * p) * kC3 + ch] : G[(size_t)(2 * p + 1) * kC3 + (ch - kC3)];
  F[(size_t)p * (2 * kC3) + ch] = f2bf(v);
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// 6D -> rotation by Gram-Schmidt (columns b1, b2, b3 = b1 x b2), translation
// o[6..8].  Loss per pair: mean over the source's FPS points s of
// ||(R - R*) s + (t - t*)||^2 = tr(dR S dR^T) + 2 dt^T dR m + |dt|^2 with
// S = mean s s^T, m = mean s.  Writes R, t (poses) and, when labels are
// given, the loss and dL/dO.
__global__ void k_head(const float* O, int n_pairs, const float* pts /* cloud*kP*3 */, const double* lab /* n*12 */,
                       double* pose /* n*12 (R row-major, t) */, float* dO, float* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const float* o = O + (size_t)p * kOut;
  double a1[3] = {o[0], o[1], o[2]}, a2[3] = {o[3], o[4], o[5]};
  const double n1 = sqrt(dot3(a1, a1)) + 1e-12;
  double b1[3] = {a1[0] / n1, a1[1] / n1, a1[2] / n1};
  const double c = dot3(b1, a2);
  double u2[3] = {a2[0] - c * b1[0], a2[1] - c * b1[1], a2[2] - c * b1[2]};
  const double n2 = sqrt(dot3(u2, u2)) + 1e-12;
  double b2[3] = {u2[0] / n2, u2[1] / n2, u2[2] / n2}, b3[3];
  cross3(b1, b2, b3);
  double R[9];  // row-major, columns b1 b2 b3
  for (int r = 0; r < 3; ++r) {
    R[r * 3] = b1[r];
    R[r * 3 + 1] = b2[r];
    R[r * 3 + 2] = b3[r];
  }
  const double t[3] = {o[6], o[7], o[8]};
  for (int k = 0; k < 9; ++k) pose[(size_t)p * 12 + k] = R[k];
  for (int k = 0; k < 3; ++k) pose[(size_t)p * 12 + 9 + k] = t[k];
  if (!lab) return;
  // moments of the source FPS points (cloud 2p)
  double S[9] = {0}, mm[3] = {0};
  const float* sp = pts + (size_t)(2 * p) * kP * 3;
  for (int i = 0; i < kP; ++i) {
    const double s[3] = {sp[3 * i], sp[3 * i + 1], sp[3 * i + 2]};
    for (int a = 0; a < 3; ++a) {
      mm[a] += s[a];
      for (int b = 0; b < 3; ++b) S[a * 3 + b] += s[a] * s[b];
    }
  }
  for (int a = 0; a < 3; ++a) mm[a] /= kP;
  for (int k = 0; k < 9; ++k) S[k] /= kP;
  const double* L = lab + (size_t)p * 12;
  double dR[9], dt[3];
  for (int k = 0; k < 9; ++k) dR[k] = R[k] - L[k];
  for (int k = 0; k < 3; ++k) dt[k] = t[k] - L[9 + k];
  double l = 0;
  for (int r = 0; r < 3; ++r)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) l += dR[r * 3 + a] * S[a * 3 + b] * dR[r * 3 + b];
  double dRm[3];
  for (int r = 0; r < 3; ++r) dRm[r] = dR[r * 3] * mm[0] + dR[r * 3 + 1] * mm[1] + dR[r * 3 + 2] * mm[2];
  l += 2.0 * dot3(dt, dRm) + dot3(dt, dt);
  loss[p] = (float)l;
  // gradients (the batch loss is the mean over pairs)
  const double inv = 1.0 / n_pairs;
  double gR[9];
  for (int r = 0; r < 3; ++r)
    for (int a = 0; a < 3; ++a) {
      double v = 0;
      for (int b = 0; b < 3; ++b) v += dR[r * 3 + b] * S[b * 3 + a];
      gR[r * 3 + a] = (2.0 * v + 2.0 * dt[r] * mm[a]) * inv;
    }
  double gt[3];
  for (int r = 0; r < 3; ++r) gt[r] = (2.0 * dt[r] + 2.0 * dRm[r]) * inv;
  // back through Gram-Schmidt: columns of gR are dL/db1, dL/db2, dL/db3
  double g1[3] = {gR[0], gR[3], gR[6]}, g2[3] = {gR[1], gR[4], gR[7]}, g3[3] = {gR[2], gR[5], gR[8]}, tmp[3];
  cross3(b2, g3, tmp);  // b3 = b1 x b2
  for (int a = 0; a < 3; ++a) g1[a] += tmp[a];
  cross3(g3, b1, tmp);
  for (int a = 0; a < 3; ++a) g2[a] += tmp[a];
  const double gb2 = dot3(b2, g2);
  double dU[3];
  for (int a = 0; a < 3; ++a) dU[a] = (g2[a] - b2[a] * gb2) / n2;
  const double dUb1 = dot3(dU, b1);
  double da2[3];
  for (int a = 0; a < 3; ++a) {
    da2[a] = dU[a] - dUb1 * b1[a];
    g1[a] += -dUb1 * a2[a] - c * dU[a];
  }
  const double gb1 = dot3(b1, g1);
  float* d = dO + (size_t)p * kOut;
  for (int a = 0; a < 3; ++a) {
    d[a] = (float)((g1[a] - b1[a] * gb1) / n1);
    d[3 + a] = (float)da2[a];
    d[6 + a] = (float)gt[a];
  }
}

// dY(pre-activation) = dY(post) * (Y > 0), as bf16 for the next product.
__global__ void k_relu_back(const float* dY, const uint16_t* Y, size_t n, uint16_t* out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = bf2f(Y[i]) > 0.f ? f2bf(dY[i]) : (uint16_t)0;
}

// Max-pool + ReLU backward fused: the gradient of the pre-activation of
// layer 3 is dG at each channel's argmax row when the pooled value is
// positive (ReLU passes it), zero elsewhere; written as bf16.
__global__ void k_pool_back(const float* dG /* cloud x C */, const int32_t* arg, const float* G, int C,
                            size_t rows, uint16_t* out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows * C; i += (size_t)gridDim.x * blockDim.x) {
    const size_t row = i / C;
    const int ch = (int)(i - row * C);
    const size_t cl = row / kP;
    const size_t gi = cl * C + ch;
    const bool hit = arg[gi] == (int32_t)(row - cl * kP) && G[gi] > 0.f;
    out[i] = hit ? f2bf(dG[gi]) : (uint16_t)0;
  }
}

// dF (pairs x 512) -> dG (clouds x 256)
__global__ void k_split(const float* dF, int n_pairs, float* dG) {
  const int p = blockIdx.x, ch = threadIdx.x;
  if (p >= n_pairs) return;
  const float v = dF[(size_t)p * (2 * kC3) + ch];
  if (ch < kC3)
    dG[(size_t)(2 * p) * kC3 + ch] = v;
  else
    dG[(size_t)(2 * p + 1) * kC3 + (ch - kC3)] = v;
}

// column sums of a bf16 matrix (bias gradients), split over row chunks
__global__ void k_colsum(const uint16_t* Y, size_t rows, int C, float* out) {
  const int ch = threadIdx.x + blockIdx.y * blockDim.x;
  if (ch >= C) return;
  const size_t chunk = (rows + gridDim.x - 1) / gridDim.x;
  const size_t r0 = blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
  float s = 0.f;
  for (size_t r = r0; r < r1; ++r) s += bf2f(Y[r * C + ch]);
  atomicAdd(out + ch, s);
}
__global__ void k_colsum_f32(const float* Y, size_t rows, int C, float* out) {
  const int ch = threadIdx.x + blockIdx.y * blockDim.x;
  if (ch >= C) return;
  const size_t chunk = (rows + gridDim.x - 1) / gridDim.x;
  const size_t r0 = blockIdx.x * chunk, r1 = min(rows, r0 + chunk);
  float s = 0.f;
  for (size_t r = r0; r < r1; ++r) s += Y[r * C + ch];
  atomicAdd(out + ch, s);
}

__global__ void k_to_bf16(const float* in, size_t n, uint16_t* out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = f2bf(in[i]);
}

// Hypothesis inputs: the n source FPS clouds rotated by each of the k
// rotations, then the n target clouds once:
// clouds [Q_0 src_0 .. Q_0 src_{n-1}, Q_1 src_0, .., Q_{k-1} src_{n-1}, tgt_0 .. tgt_{n-1}]
// (pts holds the FPS clouds as src0, tgt0, src1, ...; rots k row-major 3x3).
__global__ void k_pack_multi(const float* pts, int n, int k, const float* rots, uint16_t* x0) {
  const size_t total = (size_t)(k + 1) * n * kP * kC0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / kC0;
    const int c = (int)(i - r * kC0);
    const size_t cl = r / kP, row = r - cl * kP;
    float v = 0.f;
    if (c < 3) {
      if (cl < (size_t)k * n) {
        const size_t j = cl / n, p = cl - j * n;
        const float* q = pts + ((2 * p) * kP + row) * 3;
        const float* Q = rots + 9 * j + 3 * c;
        v = Q[0] * q[0];
        v = v + Q[1] * q[1];
        v = v + Q[2] * q[2];
      } else {
        const size_t p = cl - (size_t)k * n;
        v = pts[((2 * p + 1) * kP + row) * 3 + c];
      }
    }
    x0[i] = f2bf(v);
  }
}

// F row r = j*n + p: (G of rotated source cloud r, G of target cloud k*n + p)
__global__ void k_concat_multi(const float* G, int n, int k, uint16_t* F) {
  const int r = blockIdx.x, ch = threadIdx.x;  // 512 threads
  if (r >= k * n) return;
  const int p = r % n;
  const float v = ch < kC3 ? G[(size_t)r * kC3 + ch] : G[((size_t)k * n + p) * kC3 + (ch - kC3)];
  F[(size_t)r * (2 * kC3) + ch] = f2bf(v);
}

// Candidate score: truncated chamfer between the pair's source FPS points
// moved by the candidate (R_r Q_j, t_r) and its target FPS points, both in
// their centred frames: mean over the 1024 source points of
// min(min_q |R s + t - q|^2, tau2).  One workgroup per candidate r = j*n + p;
// the target points are staged in LDS as float4, and each thread moves its
// four source points together through them (one LDS read per target point
// for four distances; the distances are computed as before).
__global__ __launch_bounds__(256) void k_score(const float* pts, const double* pose, const float* rots, int n,
                                               int k, float tau2, double* score) {
  constexpr int kPer = kP / 256;
  __shared__ float4 tq[kP];
  __shared__ double red[4];
  const int r = blockIdx.x;
  if (r >= k * n) return;
  const int j = r / n, p = r - j * n;
  const float* tp = pts + (size_t)(2 * p + 1) * kP * 3;
  for (int i = threadIdx.x; i < kP; i += 256) tq[i] = make_float4(tp[3 * i], tp[3 * i + 1], tp[3 * i + 2], 0.f);
  const double* P = pose + (size_t)r * 12;
  const float* Q = rots + 9 * j;
  float M[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      M[a * 3 + b] = (float)(P[a * 3] * Q[b] + P[a * 3 + 1] * Q[3 + b] + P[a * 3 + 2] * Q[6 + b]);
  const float t0 = (float)P[9], t1 = (float)P[10], t2 = (float)P[11];
  __syncthreads();
  const float* sp = pts + (size_t)(2 * p) * kP * 3;
  float qx[kPer], qy[kPer], qz[kPer], best[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = threadIdx.x + u * 256;
    const float sx = sp[3 * i], sy = sp[3 * i + 1], sz = sp[3 * i + 2];
    qx[u] = M[0] * sx + M[1] * sy + M[2] * sz + t0;
    qy[u] = M[3] * sx + M[4] * sy + M[5] * sz + t1;
    qz[u] = M[6] * sx + M[7] * sy + M[8] * sz + t2;
    best[u] = tau2;
  }
  for (int q = 0; q < kP; ++q) {
    const float4 c = tq[q];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const float dx = qx[u] - c.x, dy = qy[u] - c.y, dz = qz[u] - c.z;
      best[u] = fminf(best[u], dx * dx + dy * dy + dz * dz);
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < kPer; ++u) acc += best[u];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) score[r] = (red[0] + red[1] + red[2] + red[3]) / kP;
}

// xyz rows (FP32) -> network input rows of kC0 bf16 (channels 3.. zero)
__global__ void k_pack_x0(const float* xyz, size_t rows, uint16_t* x0) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows * kC0; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / kC0;
    const int c = (int)(i - r * kC0);
    x0[i] = c < 3 ? f2bf(xyz[3 * r + c]) : (uint16_t)0;
  }
}

__global__ void k_zero(float* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint16_t* wb) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------


struct Proposer {
  int device = 0, max_pairs = 0;
  int n_pairs = 0;     // pairs of the batch in the network buffers
  int n_resident = 0;  // pairs of the resident raw clouds (set_clouds)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ParamLayout L;
  int64_t step = 0;
  DevBuf w, wb, g, m, v;
  DevBuf raw, roff, rcnt;           // resident raw clouds for proposals
  DevBuf pts, x0, cen, sel;         // FPS outputs
  DevBuf y1, y2, y3, G, arg, F, h1, O, pose, lab, dO, loss;
  DevBuf dtmp, dbf, dG, dF, dH1, dH1b, dOb;
  DevBuf rots, score;               // hypothesis rotations, candidate scores
  bool train_ready = false;
  int fused_blocks = 256;           // workgroups of the fused point MLP (one per CU)
  bool fused_off = false;           // inference through the layer-by-layer products
  std::vector<uint32_t> h_off, h_cnt;
  float last_ms = 0.f;
  void* pin = nullptr;  // pinned staging of the proposals read back each call
  size_t pin_bytes = 0;

  int ensure_pinned(size_t bytes) {
    if (pin && bytes <= pin_bytes) return RTR_OK;
    if (pin) (void)hipHostFree(pin);
    pin = nullptr;
    pin_bytes = 0;
    RTR_HIP(hipHostMalloc(&pin, bytes, hipHostMallocDefault));
    pin_bytes = bytes;
    return RTR_OK;
  }

  int init(int dev, int maxp) {
    device = dev;
    max_pairs = maxp;
    int n = 0;
    RTR_HIP(hipGetDeviceCount(&n));
    if (dev < 0 || dev >= n) {
      set_error("device %d not available", dev);
      return RTR_ERR_DEVICE;
    }
    RTR_HIP(hipSetDevice(device));
    RTR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    RTR_HIP(hipEventCreate(&ev0));
    RTR_HIP(hipEventCreate(&ev1));
    RTR_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_pool),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(FusedShared)));
    int cus = 0;
    RTR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    fused_blocks = cus > 0 ? cus : 256;
    const size_t P = L.total;
    const size_t clouds = 2 * (size_t)maxp, rows = clouds * kP;
    int rc = w.ensure(P * 4);
    if (!rc) rc = wb.ensure(P * 2);
    if (!rc) rc = pts.ensure(rows * 12);
    if (!rc) rc = x0.ensure(rows * kC0 * 2);
    if (!rc) rc = cen.ensure(clouds * 12);
    if (!rc) rc = sel.ensure(rows * 4);
    if (!rc) rc = y1.ensure(rows * kC1 * 2);
    if (!rc) rc = y2.ensure(rows * kC2 * 2);
    if (!rc) rc = y3.ensure(rows * kC3 * 2);
    if (!rc) rc = G.ensure(clouds * kC3 * 4);
    if (!rc) rc = arg.ensure(clouds * kC3 * 4);
    if (!rc) rc = F.ensure((size_t)maxp * 2 * kC3 * 2);
    if (!rc) rc = h1.ensure((size_t)maxp * kH * 2);
    if (!rc) rc = O.ensure((size_t)maxp * kOut * 4);
    if (!rc) rc = pose.ensure((size_t)maxp * 12 * 8);
    if (!rc) rc = rots.ensure(kMaxHyp * 9 * 4);
    if (!rc) rc = score.ensure((size_t)maxp * 8);
    return rc;
  }

  // Gradient, optimiser and backward buffers: allocated on the first
  // training call only (an inference-only proposer never holds them).
  int ensure_train() {
    if (train_ready) return RTR_OK;
    const size_t P = L.total;
    const size_t clouds = 2 * (size_t)max_pairs, rows = clouds * kP;
    const size_t maxp = (size_t)max_pairs;
    int rc = g.ensure(P * 4);
    if (!rc) rc = m.ensure(P * 4);
    if (!rc) rc = v.ensure(P * 4);
    if (!rc) rc = lab.ensure(maxp * 12 * 8);
    if (!rc) rc = dO.ensure(maxp * kOut * 4);
    if (!rc) rc = loss.ensure(maxp * 4);
    if (!rc) rc = dtmp.ensure(rows * kC3 * 4);
    if (!rc) rc = dbf.ensure(rows * kC3 * 2);
    if (!rc) rc = dG.ensure(clouds * kC3 * 4);
    if (!rc) rc = dF.ensure(maxp * 2 * kC3 * 4);
    if (!rc) rc = dH1.ensure(maxp * kH * 4);
    if (!rc) rc = dH1b.ensure(maxp * kH * 2);
    if (!rc) rc = dOb.ensure(maxp * 16 * 2);
    if (rc) return rc;
    RTR_HIP(hipMemset(m.p, 0, P * 4));
    RTR_HIP(hipMemset(v.p, 0, P * 4));
    train_ready = true;
    return RTR_OK;
  }

  ~Proposer() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    DevBuf* all[] = {&w,  &wb,  &g,  &m,  &v,   &raw, &roff, &rcnt, &pts, &x0,  &cen,  &sel,  &y1,
                     &y2, &y3,  &G,  &arg, &F,  &h1,  &O,    &pose, &lab, &dO,  &loss, &dtmp, &dbf,
                     &dG, &dF,  &dH1, &dH1b, &dOb, &rots, &score};
    for (DevBuf* b : all) b->release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (pin) (void)hipHostFree(pin);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int set_weights(const float* host) {
    RTR_HIP(hipSetDevice(device));
    RTR_HIP(hipMemcpy(w.p, host, L.total * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, stream, w.as<float>(), L.total, wb.as<uint16_t>());
    RTR_HIP(hipGetLastError());
    RTR_HIP(hipStreamSynchronize(stream));
    return RTR_OK;
  }
  int get_weights(float* host) {
    RTR_HIP(hipSetDevice(device));
    RTR_HIP(hipMemcpy(host, w.p, L.total * 4, hipMemcpyDeviceToHost));
    return RTR_OK;
  }

  int set_clouds(int np, const float* const* src, const size_t* ns, const float* const* tgt, const size_t* nt) {
    RTR_HIP(hipSetDevice(device));
    if (np < 1 || np > max_pairs) {
      set_error("pairs %d outside [1, %d]", np, max_pairs);
      return RTR_ERR_INVALID_ARG;
    }
    n_pairs = np;
    n_resident = 0;
    h_off.assign(2 * np, 0);
    h_cnt.assign(2 * np, 0);
    size_t tot = 0;
    for (int p = 0; p < np; ++p) {
      h_off[2 * p] = (uint32_t)tot;
      h_cnt[2 * p] = (uint32_t)ns[p];
      tot += ns[p];
      h_off[2 * p + 1] = (uint32_t)tot;
      h_cnt[2 * p + 1] = (uint32_t)nt[p];
      tot += nt[p];
    }
    if (tot >= (1ull << 32)) {
      set_error("clouds too large (%zu points)", tot);
      return RTR_ERR_CAPACITY;
    }
    int rc = raw.ensure(tot * 12);
    if (!rc) rc = roff.ensure(2 * np * 4);
    if (!rc) rc = rcnt.ensure(2 * np * 4);
    if (rc) return rc;
    {
      // pairs interleaved: source p, then target p
      std::vector<const float*> cl(2 * np);
      std::vector<size_t> cn(2 * np), co(2 * np);
      for (int p = 0; p < np; ++p) {
        cl[2 * p] = src[p];
        cl[2 * p + 1] = tgt[p];
      }
      for (int c = 0; c < 2 * np; ++c) {
        cn[c] = h_cnt[c];
        co[c] = h_off[c];
      }
      RTR_TRY(upload_xyz_staged(stream, &pin, &pin_bytes, 2 * np, cl.data(), cn.data(), nullptr, co.data(), tot,
                                raw.p));
    }
    RTR_HIP(hipMemcpyAsync(roff.p, h_off.data(), 2 * np * 4, hipMemcpyHostToDevice, stream));
    RTR_HIP(hipMemcpyAsync(rcnt.p, h_cnt.data(), 2 * np * 4, hipMemcpyHostToDevice, stream));
    RTR_HIP(hipStreamSynchronize(stream));
    n_resident = np;
    return RTR_OK;
  }

  // The same resident clouds from packed xyz already on this device: every
  // source, then every target, each group in pair order (two device copies).
  int set_device_clouds(int np, const float* src_xyz, const size_t* ns, const float* tgt_xyz, const size_t* nt) {
    RTR_HIP(hipSetDevice(device));
    if (np < 1 || np > max_pairs) {
      set_error("pairs %d outside [1, %d]", np, max_pairs);
      return RTR_ERR_INVALID_ARG;
    }
    size_t ts = 0, tt = 0;
    for (int p = 0; p < np; ++p) {
      ts += ns[p];
      tt += nt[p];
    }
    if (ts + tt >= (1ull << 32)) {
      set_error("clouds too large (%zu points)", ts + tt);
      return RTR_ERR_CAPACITY;
    }
    n_pairs = np;
    n_resident = 0;
    h_off.assign(2 * np, 0);
    h_cnt.assign(2 * np, 0);
    size_t a = 0, b = ts;
    for (int p = 0; p < np; ++p) {
      h_off[2 * p] = (uint32_t)a;
      h_cnt[2 * p] = (uint32_t)ns[p];
      a += ns[p];
      h_off[2 * p + 1] = (uint32_t)b;
      h_cnt[2 * p + 1] = (uint32_t)nt[p];
      b += nt[p];
    }
    int rc = raw.ensure((ts + tt) * 12);
    if (!rc) rc = roff.ensure(2 * np * 4);
    if (!rc) rc = rcnt.ensure(2 * np * 4);
    if (rc) return rc;
    RTR_HIP(hipMemcpyAsync(raw.p, src_xyz, ts * 12, hipMemcpyDeviceToDevice, stream));
    RTR_HIP(hipMemcpyAsync(raw.as<float>() + 3 * ts, tgt_xyz, tt * 12, hipMemcpyDeviceToDevice, stream));
    RTR_HIP(hipMemcpyAsync(roff.p, h_off.data(), 2 * np * 4, hipMemcpyHostToDevice, stream));
    RTR_HIP(hipMemcpyAsync(rcnt.p, h_cnt.data(), 2 * np * 4, hipMemcpyHostToDevice, stream));
    RTR_HIP(hipStreamSynchronize(stream));
    n_resident = np;
    return RTR_OK;
  }

  // FPS of the resident clouds into the network buffers (n_pairs becomes the
  // resident count).
  void enqueue_fps() {
    n_pairs = n_resident;
    hipLaunchKernelGGL(k_fps, dim3(2 * n_pairs), dim3(kFpsThreads), 0, stream, raw.as<float>(), roff.as<uint32_t>(),
                       rcnt.as<uint32_t>(), pts.as<float>(), x0.as<uint16_t>(), cen.as<float>(),
                       sel.as<int32_t>());
  }

  // FPS points given directly (training): centred xyz, n_pairs * 2 * kP
  int set_points(int np, const float* xyz) {
    RTR_HIP(hipSetDevice(device));
    if (np < 1 || np > max_pairs) {
      set_error("pairs %d outside [1, %d]", np, max_pairs);
      return RTR_ERR_INVALID_ARG;
    }
    n_pairs = np;
    const size_t rows = (size_t)2 * np * kP;
    RTR_HIP(hipMemcpyAsync(pts.p, xyz, rows * 12, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_pack_x0, dim3(1024), dim3(256), 0, stream, pts.as<float>(), rows, x0.as<uint16_t>());
    RTR_HIP(hipGetLastError());
    return RTR_OK;
  }

  const uint16_t* wbp(size_t off) const { return wb.as<uint16_t>() + off; }
  const float* wp(size_t off) const { return w.as<float>() + off; }

  // pooled features G of C clouds from x0: fused for inference, layer by
  // layer (activations and argmax kept for the backward) for training
  void enqueue_point_mlp(int C, bool keep_activations) {
    if (!keep_activations) {
      const int grid = C < fused_blocks ? C : fused_blocks;
      hipLaunchKernelGGL(k_mlp_pool, dim3(grid), dim3(kFThreads), sizeof(FusedShared), stream, x0.as<uint16_t>(), C,
                         wb.as<uint16_t>(), w.as<float>(), L, G.as<float>());
      return;
    }
    const int rows = C * kP;
    gemm(stream, x0.as<uint16_t>(), kC0, 0, wbp(L.W1), kC1, 0, y1.p, kC1, rows, kC1, kC0, wp(L.b1), 1, 1);
    gemm(stream, y1.as<uint16_t>(), kC1, 0, wbp(L.W2), kC2, 0, y2.p, kC2, rows, kC2, kC1, wp(L.b2), 1, 1);
    gemm(stream, y2.as<uint16_t>(), kC2, 0, wbp(L.W3), kC3, 0, y3.p, kC3, rows, kC3, kC2, wp(L.b3), 1, 1);
    hipLaunchKernelGGL(k_maxpool, dim3(C, 1), dim3(kC3), 0, stream, y3.as<uint16_t>(), kC3, C, G.as<float>(),
                       arg.as<int32_t>());
  }

  void enqueue_forward(bool with_labels, bool keep_activations) {
    const int B = n_pairs, C = 2 * B;
    enqueue_point_mlp(C, keep_activations);
    hipLaunchKernelGGL(k_concat, dim3(B), dim3(2 * kC3), 0, stream, G.as<float>(), B, F.as<uint16_t>());
    gemm(stream, F.as<uint16_t>(), 2 * kC3, 0, wbp(L.W4), kH, 0, h1.p, kH, B, kH, 2 * kC3, wp(L.b4), 1, 1);
    gemm(stream, h1.as<uint16_t>(), kH, 0, wbp(L.W5), kOut, 0, O.p, kOut, B, kOut, kH, wp(L.b5), 0, 0);
    hipLaunchKernelGGL(k_head, dim3((B + 63) / 64), dim3(64), 0, stream, O.as<float>(), B, pts.as<float>(),
                       with_labels ? lab.as<double>() : nullptr, pose.as<double>(), dO.as<float>(),
                       loss.as<float>());
  }

  // Network over k hypotheses per resident pair (after enqueue_fps): point
  // MLP over (k+1)*n clouds, head over k*n rows, then candidate scores.
  void enqueue_forward_multi(int k, float tau2) {
    const int n = n_pairs, C = (k + 1) * n, B = k * n;
    hipLaunchKernelGGL(k_pack_multi, dim3(2048), dim3(256), 0, stream, pts.as<float>(), n, k, rots.as<float>(),
                       x0.as<uint16_t>());
    enqueue_point_mlp(C, fused_off);
    hipLaunchKernelGGL(k_concat_multi, dim3(B), dim3(2 * kC3), 0, stream, G.as<float>(), n, k, F.as<uint16_t>());
    gemm(stream, F.as<uint16_t>(), 2 * kC3, 0, wbp(L.W4), kH, 0, h1.p, kH, B, kH, 2 * kC3, wp(L.b4), 1, 1);
    gemm(stream, h1.as<uint16_t>(), kH, 0, wbp(L.W5), kOut, 0, O.p, kOut, B, kOut, kH, wp(L.b5), 0, 0);
    hipLaunchKernelGGL(k_head, dim3((B + 63) / 64), dim3(64), 0, stream, O.as<float>(), B, pts.as<float>(),
                       nullptr, pose.as<double>(), nullptr, nullptr);
    hipLaunchKernelGGL(k_score, dim3(B), dim3(256), 0, stream, pts.as<float>(), pose.as<double>(),
                       rots.as<float>(), n, k, tau2, score.as<double>());
  }

  // gradients of the mean loss into g (FP32)
  void enqueue_backward() {
    const int B = n_pairs, C = 2 * B;
    const size_t rows = (size_t)C * kP;
    float* gp = g.as<float>();
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, stream, gp, L.total);
    const unsigned cs_blocks = 64;
    // head layer 2: O = H1 W5 + b5
    hipLaunchKernelGGL(k_to_bf16, dim3(64), dim3(256), 0, stream, dO.as<float>(), (size_t)B * kOut,
                       dOb.as<uint16_t>());
    gemm(stream, h1.as<uint16_t>(), kH, 1, dOb.as<uint16_t>(), kOut, 0, gp + L.W5, kOut, kH, kOut, B, nullptr, 0, 0);
    hipLaunchKernelGGL(k_colsum_f32, dim3(1, 1), dim3(64), 0, stream, dO.as<float>(), (size_t)B, kOut, gp + L.b5);
    gemm(stream, dOb.as<uint16_t>(), kOut, 0, wbp(L.W5), kOut, 1, dH1.p, kH, B, kH, kOut, nullptr, 0, 0);
    hipLaunchKernelGGL(k_relu_back, dim3(64), dim3(256), 0, stream, dH1.as<float>(), h1.as<uint16_t>(),
                       (size_t)B * kH, dH1b.as<uint16_t>());
    // head layer 1: H1 = relu(F W4 + b4)
    gemm(stream, F.as<uint16_t>(), 2 * kC3, 1, dH1b.as<uint16_t>(), kH, 0, gp + L.W4, kH, 2 * kC3, kH, B, nullptr,
         0, 0);
    hipLaunchKernelGGL(k_colsum, dim3(1, 1), dim3(kH), 0, stream, dH1b.as<uint16_t>(), (size_t)B, kH, gp + L.b4);
    gemm(stream, dH1b.as<uint16_t>(), kH, 0, wbp(L.W4), kH, 1, dF.p, 2 * kC3, B, 2 * kC3, kH, nullptr, 0, 0);
    hipLaunchKernelGGL(k_split, dim3(B), dim3(2 * kC3), 0, stream, dF.as<float>(), B, dG.as<float>());
    // max-pool + ReLU: dY3 nonzero at the argmax rows only
    hipLaunchKernelGGL(k_pool_back, dim3(2048), dim3(256), 0, stream, dG.as<float>(), arg.as<int32_t>(),
                       G.as<float>(), kC3, rows, dbf.as<uint16_t>());
    const int ks = (int)std::min<size_t>(256, rows / 1024 + 1);
    // layer 3: Y3 = relu(Y2 W3 + b3)
    gemm(stream, y2.as<uint16_t>(), kC2, 1, dbf.as<uint16_t>(), kC3, 0, gp + L.W3, kC3, kC2, kC3, (int)rows, nullptr,
         0, 0, ks);
    hipLaunchKernelGGL(k_colsum, dim3(cs_blocks, 1), dim3(kC3), 0, stream, dbf.as<uint16_t>(), rows, kC3, gp + L.b3);
    gemm(stream, dbf.as<uint16_t>(), kC3, 0, wbp(L.W3), kC3, 1, dtmp.p, kC2, (int)rows, kC2, kC3, nullptr, 0, 0);
    hipLaunchKernelGGL(k_relu_back, dim3(1024), dim3(256), 0, stream, dtmp.as<float>(), y2.as<uint16_t>(),
                       rows * kC2, dbf.as<uint16_t>());
    // layer 2
    gemm(stream, y1.as<uint16_t>(), kC1, 1, dbf.as<uint16_t>(), kC2, 0, gp + L.W2, kC2, kC1, kC2, (int)rows, nullptr,
         0, 0, ks);
    hipLaunchKernelGGL(k_colsum, dim3(cs_blocks, 1), dim3(kC2), 0, stream, dbf.as<uint16_t>(), rows, kC2, gp + L.b2);
    gemm(stream, dbf.as<uint16_t>(), kC2, 0, wbp(L.W2), kC2, 1, dtmp.p, kC1, (int)rows, kC1, kC2, nullptr, 0, 0);
    hipLaunchKernelGGL(k_relu_back, dim3(1024), dim3(256), 0, stream, dtmp.as<float>(), y1.as<uint16_t>(),
                       rows * kC1, dbf.as<uint16_t>());
    // layer 1
    gemm(stream, x0.as<uint16_t>(), kC0, 1, dbf.as<uint16_t>(), kC1, 0, gp + L.W1, kC1, kC0, kC1, (int)rows, nullptr,
         0, 0, ks);
    hipLaunchKernelGGL(k_colsum, dim3(cs_blocks, 1), dim3(kC1), 0, stream, dbf.as<uint16_t>(), rows, kC1, gp + L.b1);
  }

  int upload_labels(const double* labels) {
    RTR_HIP(hipMemcpyAsync(lab.p, labels, (size_t)n_pairs * 12 * 8, hipMemcpyHostToDevice, stream));
    return RTR_OK;
  }
};

}  // namespace rtr

using rtr::Proposer;
using rtr::set_error;

struct rtr_prop {
  Proposer p;
};

#define PROP_GUARD(x)                     \
  try {                                   \
    x                                     \
  } catch (...) {                         \
    set_error("host allocation failed");  \
    return RTR_ERR_CAPACITY;              \
  }

extern "C" {

int rtr_prop_create(rtr_prop** out, int device, int max_pairs) {
  PROP_GUARD(if (!out || max_pairs < 1) return RTR_ERR_INVALID_ARG; *out = nullptr; rtr_prop* h = new rtr_prop();
             int rc = h->p.init(device, max_pairs); if (rc) {
               delete h;
               return rc;
             } *out = h;
             return RTR_OK;)
}

int rtr_prop_destroy(rtr_prop* h) {
  delete h;
  return RTR_OK;
}

int rtr_prop_num_params(size_t* n) {
  if (!n) return RTR_ERR_INVALID_ARG;
  *n = rtr::ParamLayout().total;
  return RTR_OK;
}

int rtr_prop_set_weights(rtr_prop* h, const float* w) {
  if (!h || !w) return RTR_ERR_INVALID_ARG;
  return h->p.set_weights(w);
}

int rtr_prop_get_weights(rtr_prop* h, float* w) {
  if (!h || !w) return RTR_ERR_INVALID_ARG;
  return h->p.get_weights(w);
}

int rtr_prop_set_clouds(rtr_prop* h, int n_pairs, const float* const* src, const size_t* ns,
                        const float* const* tgt, const size_t* nt) {
  PROP_GUARD(if (!h || !src || !ns || !tgt || !nt) return RTR_ERR_INVALID_ARG;
             return h->p.set_clouds(n_pairs, src, ns, tgt, nt);)
}

int rtr_prop_set_device_clouds(rtr_prop* h, int n_pairs, const float* src_xyz, const size_t* ns,
                               const float* tgt_xyz, const size_t* nt) {
  PROP_GUARD(if (!h || !src_xyz || !ns || !tgt_xyz || !nt) return RTR_ERR_INVALID_ARG;
             return h->p.set_device_clouds(n_pairs, src_xyz, ns, tgt_xyz, nt);)
}

// FPS + forward on the resident clouds; poses (n*16 column-major) map each
// source onto its target.
int rtr_prop_propose(rtr_prop* h, double* poses) {
  using namespace rtr;
  if (!h || !poses) return RTR_ERR_INVALID_ARG;
  Proposer& P = h->p;
  RTR_HIP(hipSetDevice(P.device));
  if (P.n_resident < 1) {
    set_error("no clouds set");
    return RTR_ERR_NOT_READY;
  }
  RTR_HIP(hipEventRecord(P.ev0, P.stream));
  P.enqueue_fps();
  P.enqueue_forward(false, P.fused_off);
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipEventRecord(P.ev1, P.stream));
  std::vector<double> pose((size_t)P.n_pairs * 12);
  std::vector<float> cen((size_t)2 * P.n_pairs * 3);
  RTR_HIP(hipMemcpyAsync(pose.data(), P.pose.p, pose.size() * 8, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipMemcpyAsync(cen.data(), P.cen.p, cen.size() * 4, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipStreamSynchronize(P.stream));
  RTR_HIP(hipEventElapsedTime(&P.last_ms, P.ev0, P.ev1));
  for (int p = 0; p < P.n_pairs; ++p) {
    const double* R = &pose[(size_t)p * 12];
    const double* t = R + 9;
    const float* cs = &cen[(size_t)(2 * p) * 3];
    const float* ct = &cen[(size_t)(2 * p + 1) * 3];
    double* T = poses + (size_t)p * 16;
    // T = Tr(ct) [R, t] Tr(-cs)
    for (int i = 0; i < 16; ++i) T[i] = 0.0;
    T[15] = 1.0;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) T[c * 4 + r] = R[r * 3 + c];
      T[12 + r] = t[r] + ct[r] - (R[r * 3] * cs[0] + R[r * 3 + 1] * cs[1] + R[r * 3 + 2] * cs[2]);
    }
  }
  return RTR_OK;
}

int rtr_prop_set_fused(rtr_prop* h, int on) {
  if (!h) return RTR_ERR_INVALID_ARG;
  h->p.fused_off = (on == 0);
  return RTR_OK;
}

int rtr_prop_last_ms(rtr_prop* h, float* ms) {
  if (!h || !ms) return RTR_ERR_INVALID_ARG;
  *ms = h->p.last_ms;
  return RTR_OK;
}

// k hypotheses per resident pair: the network on the source FPS points
// rotated by each rots[j] (k row-major 3x3), candidate j's pose composed with
// rots[j], every candidate scored by the truncated chamfer (tau) of the FPS
// points; outputs pair-major (p*k + j): poses (n*k*16 column-major, nullable),
// scores (n*k, nullable), and the best-scoring pose per pair (n*16; lowest j
// on ties).
int rtr_prop_propose_multi(rtr_prop* h, int k, const double* rots_in, float tau, double* poses, double* scores,
                           double* best) {
  using namespace rtr;
  if (!h || !rots_in || !best || k < 1 || k > kMaxHyp || !(tau > 0.f)) return RTR_ERR_INVALID_ARG;
  Proposer& P = h->p;
  RTR_HIP(hipSetDevice(P.device));
  if (P.n_resident < 1) {
    set_error("no clouds set");
    return RTR_ERR_NOT_READY;
  }
  const int n = P.n_resident;
  if ((size_t)k * n > (size_t)P.max_pairs || (size_t)(k + 1) * n > 2 * (size_t)P.max_pairs) {
    set_error("%d hypotheses x %d pairs exceed the proposer's capacity (%d pairs)", k, n, P.max_pairs);
    return RTR_ERR_CAPACITY;
  }
  float rf[kMaxHyp * 9];
  for (int i = 0; i < 9 * k; ++i) rf[i] = (float)rots_in[i];
  RTR_HIP(hipMemcpyAsync(P.rots.p, rf, (size_t)k * 9 * 4, hipMemcpyHostToDevice, P.stream));
  RTR_HIP(hipEventRecord(P.ev0, P.stream));
  P.enqueue_fps();
  P.enqueue_forward_multi(k, tau * tau);
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipEventRecord(P.ev1, P.stream));
  // read back through pinned staging (poses, scores, centroids)
  const size_t npose = (size_t)k * n * 12, nsc = (size_t)k * n, ncen = (size_t)2 * n * 3;
  RTR_TRY(P.ensure_pinned((npose + nsc) * 8 + ncen * 4));
  double* const pose = static_cast<double*>(P.pin);
  double* const sc = pose + npose;
  float* const cen = reinterpret_cast<float*>(sc + nsc);
  RTR_HIP(hipMemcpyAsync(pose, P.pose.p, npose * 8, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipMemcpyAsync(sc, P.score.p, nsc * 8, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipMemcpyAsync(cen, P.cen.p, ncen * 4, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipStreamSynchronize(P.stream));
  RTR_HIP(hipEventElapsedTime(&P.last_ms, P.ev0, P.ev1));
  for (int p = 0; p < n; ++p) {
    const float* cs = &cen[(size_t)(2 * p) * 3];
    const float* ct = &cen[(size_t)(2 * p + 1) * 3];
    int jb = 0;
    for (int j = 0; j < k; ++j) {
      const size_t r = (size_t)j * n + p;
      if (sc[r] < sc[(size_t)jb * n + p]) jb = j;
      if (scores) scores[(size_t)p * k + j] = sc[r];
    }
    for (int j = 0; j < k; ++j) {
      if (!poses && j != jb) continue;
      const size_t r = (size_t)j * n + p;
      const double* Rn = &pose[r * 12];
      const double* t = Rn + 9;
      const double* Q = rots_in + 9 * j;
      double Rm[9];  // R = Rn Q
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Rm[a * 3 + b] = Rn[a * 3] * Q[b] + Rn[a * 3 + 1] * Q[3 + b] + Rn[a * 3 + 2] * Q[6 + b];
      double T[16];
      for (int i = 0; i < 16; ++i) T[i] = 0.0;
      T[15] = 1.0;
      for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) T[b * 4 + a] = Rm[a * 3 + b];
        T[12 + a] = t[a] + ct[a] - (Rm[a * 3] * cs[0] + Rm[a * 3 + 1] * cs[1] + Rm[a * 3 + 2] * cs[2]);
      }
      if (poses) memcpy(poses + ((size_t)p * k + j) * 16, T, sizeof(T));
      if (j == jb) memcpy(best + (size_t)p * 16, T, sizeof(T));
    }
  }
  return RTR_OK;
}

// FPS of the resident clouds: centred points (2*n_pairs*1024*3), centroids
// (2*n_pairs*3) and selected subsample indices (nullable).
int rtr_prop_fps(rtr_prop* h, float* pts, float* centroids, int32_t* sel) {
  using namespace rtr;
  if (!h || !pts) return RTR_ERR_INVALID_ARG;
  Proposer& P = h->p;
  RTR_HIP(hipSetDevice(P.device));
  if (P.n_resident < 1) {
    set_error("no clouds set");
    return RTR_ERR_NOT_READY;
  }
  P.enqueue_fps();
  RTR_HIP(hipGetLastError());
  const size_t rows = (size_t)2 * P.n_pairs * kP;
  RTR_HIP(hipMemcpyAsync(pts, P.pts.p, rows * 12, hipMemcpyDeviceToHost, P.stream));
  if (centroids)
    RTR_HIP(hipMemcpyAsync(centroids, P.cen.p, (size_t)2 * P.n_pairs * 12, hipMemcpyDeviceToHost, P.stream));
  if (sel) RTR_HIP(hipMemcpyAsync(sel, P.sel.p, rows * 4, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipStreamSynchronize(P.stream));
  return RTR_OK;
}

// Forward on given centred FPS points (n_pairs*2*1024*3, clouds src0, tgt0,
// src1, ...); outputs: raw head outputs O (n*9, nullable), R|t (n*12, nullable),
// and with labels (n*12: R row-major, t, in the centred frames) the mean loss
// and the gradient of the mean loss (nullable).
int rtr_prop_forward_backward(rtr_prop* h, int n_pairs, const float* xyz, const double* labels, float* O_out,
                              double* rt_out, float* loss_out, float* grad_out) {
  using namespace rtr;
  if (!h || !xyz) return RTR_ERR_INVALID_ARG;
  Proposer& P = h->p;
  int rc = labels ? P.ensure_train() : RTR_OK;
  if (!rc) rc = P.set_points(n_pairs, xyz);
  if (rc) return rc;
  if (labels) RTR_HIP(hipMemcpyAsync(P.lab.p, labels, (size_t)n_pairs * 12 * 8, hipMemcpyHostToDevice, P.stream));
  P.enqueue_forward(labels != nullptr, labels != nullptr || P.fused_off);
  if (labels && grad_out) P.enqueue_backward();
  RTR_HIP(hipGetLastError());
  if (O_out) RTR_HIP(hipMemcpyAsync(O_out, P.O.p, (size_t)n_pairs * kOut * 4, hipMemcpyDeviceToHost, P.stream));
  if (rt_out) RTR_HIP(hipMemcpyAsync(rt_out, P.pose.p, (size_t)n_pairs * 12 * 8, hipMemcpyDeviceToHost, P.stream));
  std::vector<float> l(labels ? n_pairs : 0);
  if (labels) RTR_HIP(hipMemcpyAsync(l.data(), P.loss.p, (size_t)n_pairs * 4, hipMemcpyDeviceToHost, P.stream));
  if (labels && grad_out) RTR_HIP(hipMemcpyAsync(grad_out, P.g.p, P.L.total * 4, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipStreamSynchronize(P.stream));
  if (loss_out && labels) {
    double s = 0;
    for (float v : l) s += v;
    *loss_out = (float)(s / n_pairs);
  }
  return RTR_OK;
}

// One Adam step on a batch (same inputs as rtr_prop_forward_backward).
int rtr_prop_train_step(rtr_prop* h, int n_pairs, const float* xyz, const double* labels, float lr, float* loss_out) {
  using namespace rtr;
  if (!h || !xyz || !labels) return RTR_ERR_INVALID_ARG;
  Proposer& P = h->p;
  int rc = P.ensure_train();
  if (!rc) rc = P.set_points(n_pairs, xyz);
  if (rc) return rc;
  RTR_HIP(hipMemcpyAsync(P.lab.p, labels, (size_t)n_pairs * 12 * 8, hipMemcpyHostToDevice, P.stream));
  P.enqueue_forward(true, true);
  P.enqueue_backward();
  P.step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)P.step), c2 = 1.f - powf(b2, (float)P.step);
  hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, P.stream, P.w.as<float>(), P.g.as<float>(), P.m.as<float>(),
                     P.v.as<float>(), P.L.total, lr, b1, b2, 1e-8f, c1, c2, P.wb.as<uint16_t>());
  RTR_HIP(hipGetLastError());
  std::vector<float> l(n_pairs);
  RTR_HIP(hipMemcpyAsync(l.data(), P.loss.p, (size_t)n_pairs * 4, hipMemcpyDeviceToHost, P.stream));
  RTR_HIP(hipStreamSynchronize(P.stream));
  if (loss_out) {
    double s = 0;
    for (float v : l) s += v;
    *loss_out = (float)(s / n_pairs);
  }
  return RTR_OK;
}

int rtr_bf16_conversion_check(int device, unsigned long long* mismatches) {
  using namespace rtr;
  if (!mismatches) return RTR_ERR_INVALID_ARG;
  RTR_HIP(hipSetDevice(device));
  DevBuf d;
  RTR_TRY(d.ensure(sizeof(unsigned long long)));
  RTR_HIP(hipMemset(d.p, 0, sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_bf16_check, dim3(1u << 20), dim3(256), 0, nullptr, d.as<unsigned long long>());
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipMemcpy(mismatches, d.p, sizeof(unsigned long long), hipMemcpyDeviceToHost));
  d.release();
  return RTR_OK;
}

// Plain BF16 MFMA product for tests: C (M x N, FP32) = op(A) op(B), A/B given
// as FP32 and rounded to BF16 on the host.
int rtr_gemm_bf16(int device, const float* A, int ta, const float* B, int tb, float* C, int M, int N, int K,
                  int k_split) {
  using namespace rtr;
  if (!A || !B || !C || M < 1 || N < 1 || K < 1 || k_split < 1) return RTR_ERR_INVALID_ARG;
  RTR_HIP(hipSetDevice(device));
  std::vector<uint16_t> a((size_t)M * K), b((size_t)K * N);
  for (size_t i = 0; i < a.size(); ++i) a[i] = f2bf(A[i]);
  for (size_t i = 0; i < b.size(); ++i) b[i] = f2bf(B[i]);
  DevBuf da, db, dc;
  int rc = da.ensure(a.size() * 2);
  if (!rc) rc = db.ensure(b.size() * 2);
  if (!rc) rc = dc.ensure((size_t)M * N * 4);
  if (rc) return rc;
  RTR_HIP(hipMemcpy(da.p, a.data(), a.size() * 2, hipMemcpyHostToDevice));
  RTR_HIP(hipMemcpy(db.p, b.data(), b.size() * 2, hipMemcpyHostToDevice));
  RTR_HIP(hipMemset(dc.p, 0, (size_t)M * N * 4));
  gemm(nullptr, da.as<uint16_t>(), ta ? M : K, ta, db.as<uint16_t>(), tb ? K : N, tb, dc.p, N, M, N, K, nullptr, 0, 0,
       k_split);
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipMemcpy(C, dc.p, (size_t)M * N * 4, hipMemcpyDeviceToHost));
  da.release();
  db.release();
  dc.release();
  return RTR_OK;
}

}  // extern "C"
