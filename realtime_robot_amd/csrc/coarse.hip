// Keypoint + TDF + yaw-search coarse proposer: the reference's own initial
// alignment, restated for the GPU as the baseline the learned proposer is
// measured against.  Reference: RealTimeRobot.cpp:45-105 (pipeline),
// key_point.h:251-334 (model keypoint TDF: a 30^3 grid of 1 cm voxels about
// the keypoint, squared voxel distance capped at 900), key_point.h:112-161
// (scan keypoint occupancy: 1 cm voxels within +-10 cm), matching.h:122-222
// (get_Distance: translate one keypoint onto the other, 36 yaw steps of 10
// degrees about z, score = mean TDF over the rotated occupied voxels, keep
// the best yaw), function.h:35-109 (Ransac: keep the pair whose transform the
// most other pairs agree with) and function.h:171-178 (occupancy-count filter).
//
// Here, per pair (source = the moved cloud, target = the scene):
//   * K keypoints per cloud by farthest-point sampling of a strided subsample
//     (the reference uses Harris 3D keypoints from PCL, not available here);
//   * source keypoint j: the list of occupied 1 cm voxels within +-10 cm;
//   * target keypoint i: the 30^3 TDF about it, built in LDS (exact separable
//     squared distance transform, as in tdf.hip), capped at 900;
//   * score(i, j, yaw) = mean over j's occupied voxel centres v of
//     TDF_i(Rz(yaw) v), voxels that land outside the grid counting 900;
//     pairs whose +-10 cm occupied-voxel counts differ by more than 2x are
//     rejected (the intent of function.h:171-178; the reference's integer
//     division makes its version nearly always pass);
//   * the best yaw per (i, j), the M best (i, j) by score become candidate
//     poses T = Trans(k_i) Rz(yaw) Trans(-k_j), and the candidate with the
//     lowest truncated chamfer between 1024-point subsamples wins (the
//     consensus step; the reference counts agreeing keypoint pairs instead).
// The reference squares its TDF values again when scoring (distance^4) and
// scores the model keypoint's own occupancy against its own TDF
// (matching.h:149-189); the intended keypoint-to-keypoint comparison is used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rtr.h"
#include "engine.hpp"

namespace rtr {

constexpr int kCkKeys = 32;        // keypoints per cloud
constexpr int kCkSub = 8192;       // strided subsample used for keypoints and occupancy
constexpr int kCkOcc = 20;         // source occupancy grid: +-10 cm at 1 cm
constexpr int kCkTdf = 30;         // target TDF grid: +-15 cm at 1 cm
constexpr int kCkYaw = 36;         // yaw steps of 10 degrees
constexpr int kCkMaxVox = 2048;    // occupied voxels kept per source keypoint
constexpr int kCkCons = 1024;      // consensus subsample per cloud
constexpr float kCkCap = 900.f;    // TDF cap (squared voxel units)

// Strided subsample index i of m from a cloud of n points.
__device__ __forceinline__ uint32_t ck_sub(int i, int m, uint32_t n) {
  return (uint32_t)((uint64_t)i * n / (uint64_t)m);
}

// K keypoints per cloud: farthest-point sampling over a strided subsample of
// at most kCkSub points, starting at subsample point 0, ties to the lowest
// index.  One workgroup per cloud, 256 threads, 32 points per thread.
__global__ __launch_bounds__(256) void k_ck_keypoints(const float* raw, const uint32_t* off, const uint32_t* cnt,
                                                      float* keys /* cloud * kCkKeys * 3 */) {
  constexpr int kPer = kCkSub / 256;
  const int c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t n = cnt[c];
  const float* p = raw + 3 * (size_t)off[c];
  const int m = (int)min((uint32_t)kCkSub, n);
  __shared__ float s_q[3];
  __shared__ float red_v[4];
  __shared__ int red_i[4];
  float px[kPer], py[kPer], pz[kPer], d[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * 256;
    px[k] = py[k] = pz[k] = 0.f;
    d[k] = -1.f;
    if (i < m) {
      const uint32_t s = ck_sub(i, m, n);
      px[k] = p[3 * s];
      py[k] = p[3 * s + 1];
      pz[k] = p[3 * s + 2];
      d[k] = 3.0e38f;
    }
  }
  if (tid == 0) {
    s_q[0] = px[0];
    s_q[1] = py[0];
    s_q[2] = pz[0];
  }
  __syncthreads();
  for (int s = 0; s < kCkKeys; ++s) {
    const float qx = s_q[0], qy = s_q[1], qz = s_q[2];
    if (tid == 0) {
      float* o = keys + ((size_t)c * kCkKeys + s) * 3;
      o[0] = qx;
      o[1] = qy;
      o[2] = qz;
    }
    float bv = -1.f;
    int bi = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      if (d[k] >= 0.f) {
        const float dx = px[k] - qx, dy = py[k] - qy, dz = pz[k] - qz;
        float e = dx * dx;
        e = e + dy * dy;
        e = e + dz * dz;
        d[k] = fminf(d[k], e);
        if (d[k] > bv) {
          bv = d[k];
          bi = tid + k * 256;
        }
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(bv, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[w] = bv;
      red_i[w] = bi;
    }
    __syncthreads();
    float gv = red_v[0];
    int gi = red_i[0];
    for (int k = 1; k < 4; ++k)
      if (red_v[k] > gv || (red_v[k] == gv && red_i[k] < gi)) {
        gv = red_v[k];
        gi = red_i[k];
      }
    const int cur = (gi >= 0 && gi < m) ? gi : 0;
    if (tid == (cur % 256)) {
      const int kk = cur / 256;
#pragma unroll
      for (int k = 0; k < kPer; ++k)
        if (k == kk) {
          s_q[0] = px[k];
          s_q[1] = py[k];
          s_q[2] = pz[k];
        }
    }
    __syncthreads();
  }
}

// Source keypoint occupancy: the occupied 1 cm voxels within +-10 cm of the
// keypoint (voxel x + 20 y + 400 z of floor((p - k + 0.1) / res)), in
// ascending voxel order, at most kCkMaxVox.  One workgroup per (pair, j).
__global__ __launch_bounds__(256) void k_ck_occupancy(const float* raw, const uint32_t* off, const uint32_t* cnt,
                                                      const float* keys, float res, uint16_t* vox, int* nvox) {
  constexpr int kBits = kCkOcc * kCkOcc * kCkOcc, kWords = (kBits + 31) / 32;
  const int b = blockIdx.x, pair = b / kCkKeys, j = b - pair * kCkKeys, tid = threadIdx.x;
  const int c = 2 * pair;  // the source cloud of the pair
  __shared__ uint32_t bits[kWords];
  __shared__ uint32_t base[kWords];
  for (int t = tid; t < kWords; t += 256) bits[t] = 0u;
  __syncthreads();
  const uint32_t n = cnt[c];
  const float* p = raw + 3 * (size_t)off[c];
  const int m = (int)min((uint32_t)kCkSub, n);
  const float* kp = keys + ((size_t)c * kCkKeys + j) * 3;
  const float h = 0.5f * kCkOcc * res, inv = 1.f / res;
  for (int i = tid; i < m; i += 256) {
    const uint32_t s = ck_sub(i, m, n);
    const float fx = (p[3 * s] - kp[0] + h) * inv, fy = (p[3 * s + 1] - kp[1] + h) * inv,
                fz = (p[3 * s + 2] - kp[2] + h) * inv;
    if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f && fx < (float)kCkOcc && fy < (float)kCkOcc && fz < (float)kCkOcc))
      continue;
    const int v = (int)fx + kCkOcc * ((int)fy + kCkOcc * (int)fz);
    atomicOr(&bits[v >> 5], 1u << (v & 31));
  }
  __syncthreads();
  if (tid == 0) {  // exclusive prefix of the word popcounts (250 words)
    uint32_t s = 0;
    for (int t = 0; t < kWords; ++t) {
      base[t] = s;
      s += (uint32_t)__popc(bits[t]);
    }
    nvox[b] = (int)min(s, (uint32_t)kCkMaxVox);
  }
  __syncthreads();
  uint16_t* o = vox + (size_t)b * kCkMaxVox;
  for (int t = tid; t < kWords; t += 256) {
    uint32_t w = bits[t], k = base[t];
    while (w) {
      const int bit = __ffs(w) - 1;
      w &= w - 1u;
      if (k < (uint32_t)kCkMaxVox) o[k] = (uint16_t)(t * 32 + bit);
      ++k;
    }
  }
}

// cos/sin of the yaw steps, rounded to float on the host from double
struct YawTable {
  float c[kCkYaw], s[kCkYaw];
};

// One 1-D pass of the squared distance transform over the 30^3 LDS grid along
// axis A; one line per thread, held in registers.
template <int A>
__device__ __forceinline__ void ck_edt_pass(float* g, float cap) {
  constexpr int N = kCkTdf;
  for (int line = threadIdx.x; line < N * N; line += blockDim.x) {
    const int u = line % N, v = line / N;
    int base, stride;
    if (A == 0) {
      base = N * (u + N * v);
      stride = 1;
    } else if (A == 1) {
      base = u + N * N * v;
      stride = N;
    } else {
      base = u + N * v;
      stride = N * N;
    }
    float in[N];
#pragma unroll
    for (int t = 0; t < N; ++t) in[t] = g[base + t * stride];
#pragma unroll 2
    for (int a = 0; a < N; ++a) {
      float best = 1.0e30f;
#pragma unroll
      for (int t = 0; t < N; ++t) {
        const float d = (float)(a - t);
        best = fminf(best, d * d + in[t]);
      }
      g[base + a * stride] = cap > 0.f ? fminf(best, cap) : best;
    }
  }
}

// Target keypoint i of every pair: its TDF in LDS, then score(i, j, yaw) for
// every source keypoint j and yaw step; writes the best yaw and its score per
// (pair, i, j).  One workgroup per (pair, i).
__global__ __launch_bounds__(256) void k_ck_score(const float* raw, const uint32_t* off, const uint32_t* cnt,
                                                  const float* keys, float res, const uint16_t* vox,
                                                  const int* nvox, YawTable yt, float* best_score, int* best_yaw) {
  constexpr int N = kCkTdf, kCells = N * N * N;
  extern __shared__ float dyn[];
  float* g = dyn;                          // kCells
  float* sc = dyn + kCells;                // kCkKeys * kCkYaw
  __shared__ int s_inner;
  const int b = blockIdx.x, pair = b / kCkKeys, i = b - pair * kCkKeys, tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int c = 2 * pair + 1;  // the target cloud of the pair
  for (int t = tid; t < kCells; t += 256) g[t] = 1.0e30f;
  if (tid == 0) s_inner = 0;
  __syncthreads();
  const uint32_t n = cnt[c];
  const float* p = raw + 3 * (size_t)off[c];
  const int m = (int)min((uint32_t)kCkSub, n);
  const float* kp = keys + ((size_t)c * kCkKeys + i) * 3;
  const float h = 0.5f * N * res, inv = 1.f / res;
  for (int q = tid; q < m; q += 256) {
    const uint32_t s = ck_sub(q, m, n);
    const float fx = (p[3 * s] - kp[0] + h) * inv, fy = (p[3 * s + 1] - kp[1] + h) * inv,
                fz = (p[3 * s + 2] - kp[2] + h) * inv;
    if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f && fx < (float)N && fy < (float)N && fz < (float)N)) continue;
    g[(int)fx + N * ((int)fy + N * (int)fz)] = 0.f;
  }
  __syncthreads();
  // occupied voxels of the inner +-10 cm box (the source side's extent), for
  // the occupancy-count filter
  {
    int cntv = 0;
    const int lo = (N - kCkOcc) / 2, hi = lo + kCkOcc;
    for (int t = tid; t < kCells; t += 256) {
      const int x = t % N, y = (t / N) % N, z = t / (N * N);
      if (g[t] == 0.f && x >= lo && x < hi && y >= lo && y < hi && z >= lo && z < hi) ++cntv;
    }
    for (int o = 32; o > 0; o >>= 1) cntv += __shfl_xor(cntv, o, 64);
    if (lane == 0) atomicAdd(&s_inner, cntv);
  }
  __syncthreads();
  ck_edt_pass<0>(g, 0.f);
  __syncthreads();
  ck_edt_pass<1>(g, 0.f);
  __syncthreads();
  ck_edt_pass<2>(g, kCkCap);
  __syncthreads();
  const int n_i = s_inner;
  // one wave per (j, yaw) combination at a time, lanes over j's voxels
  for (int combo = w; combo < kCkKeys * kCkYaw; combo += 4) {
    const int j = combo / kCkYaw, yaw = combo - j * kCkYaw;
    const int sb = pair * kCkKeys + j;
    const int nv = nvox[sb];
    float sum = 0.f;
    const bool ok = nv > 0 && n_i > 0 && !(n_i > 2 * nv || nv > 2 * n_i);
    if (ok) {
      const float ca = yt.c[yaw], sa = yt.s[yaw];
      const uint16_t* lv = vox + (size_t)sb * kCkMaxVox;
      for (int t = lane; t < nv; t += 64) {
        const int v = lv[t];
        const int vx = v % kCkOcc, vy = (v / kCkOcc) % kCkOcc, vz = v / (kCkOcc * kCkOcc);
        // voxel centre relative to the source keypoint
        const float cx = ((float)vx + 0.5f) * res - 0.5f * kCkOcc * res;
        const float cy = ((float)vy + 0.5f) * res - 0.5f * kCkOcc * res;
        const float cz = ((float)vz + 0.5f) * res - 0.5f * kCkOcc * res;
        const float rx = ca * cx - sa * cy, ry = sa * cx + ca * cy;
        const float fx = (rx + h) * inv, fy = (ry + h) * inv, fz = (cz + h) * inv;
        float val = kCkCap;
        if (fx >= 0.f && fy >= 0.f && fz >= 0.f && fx < (float)N && fy < (float)N && fz < (float)N)
          val = g[(int)fx + N * ((int)fy + N * (int)fz)];
        sum += val;
      }
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (lane == 0) sc[combo] = ok ? sum / (float)nv : 3.0e38f;
  }
  __syncthreads();
  if (tid < kCkKeys) {
    const int j = tid;
    float bs = sc[j * kCkYaw];
    int by = 0;
    for (int y = 1; y < kCkYaw; ++y)
      if (sc[j * kCkYaw + y] < bs) {
        bs = sc[j * kCkYaw + y];
        by = y;
      }
    const size_t o = ((size_t)pair * kCkKeys + i) * kCkKeys + j;
    best_score[o] = bs;
    best_yaw[o] = by;
  }
}

// Consensus: truncated chamfer of candidate r = pair * M + k, mean over a
// 1024-point strided subsample of the source moved by the candidate of
// min(min over the target subsample |q - t|^2, tau2).
__global__ __launch_bounds__(256) void k_ck_consensus(const float* raw, const uint32_t* off, const uint32_t* cnt,
                                                      const double* cand, int M, float tau2, double* score) {
  __shared__ float tx[kCkCons], ty[kCkCons], tz[kCkCons];
  __shared__ double red[4];
  const int r = blockIdx.x, pair = r / M;
  const int ct = 2 * pair + 1, cs = 2 * pair;
  const uint32_t nt = cnt[ct], ns = cnt[cs];
  const int mt = (int)min((uint32_t)kCkCons, nt), ms = (int)min((uint32_t)kCkCons, ns);
  const float* tp = raw + 3 * (size_t)off[ct];
  for (int i = threadIdx.x; i < mt; i += 256) {
    const uint32_t s = ck_sub(i, mt, nt);
    tx[i] = tp[3 * s];
    ty[i] = tp[3 * s + 1];
    tz[i] = tp[3 * s + 2];
  }
  const double* T = cand + (size_t)r * 16;  // column-major
  const float m00 = (float)T[0], m01 = (float)T[4], m02 = (float)T[8], m03 = (float)T[12];
  const float m10 = (float)T[1], m11 = (float)T[5], m12 = (float)T[9], m13 = (float)T[13];
  const float m20 = (float)T[2], m21 = (float)T[6], m22 = (float)T[10], m23 = (float)T[14];
  __syncthreads();
  const float* sp = raw + 3 * (size_t)off[cs];
  double acc = 0.0;
  for (int i = threadIdx.x; i < ms; i += 256) {
    const uint32_t s = ck_sub(i, ms, ns);
    const float x = sp[3 * s], y = sp[3 * s + 1], z = sp[3 * s + 2];
    const float qx = m00 * x + m01 * y + m02 * z + m03;
    const float qy = m10 * x + m11 * y + m12 * z + m13;
    const float qz = m20 * x + m21 * y + m22 * z + m23;
    float best = tau2;
    for (int q = 0; q < mt; ++q) {
      const float dx = qx - tx[q], dy = qy - ty[q], dz = qz - tz[q];
      best = fminf(best, dx * dx + dy * dy + dz * dz);
    }
    acc += best;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) score[r] = ms > 0 ? (red[0] + red[1] + red[2] + red[3]) / ms : 3.0e38;
}

struct Coarse {
  int device = 0;
  int n_pairs = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  DevBuf raw, off, cnt, keys, vox, nvox, bscore, byaw, cand, cscore;
  std::vector<float> hkeys;

  int init(int dev) {
    device = dev;
    RTR_HIP(hipSetDevice(dev));
    RTR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    RTR_HIP(hipEventCreate(&ev0));
    RTR_HIP(hipEventCreate(&ev1));
    RTR_HIP(hipFuncSetAttribute((const void*)k_ck_score, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (kCkTdf * kCkTdf * kCkTdf + kCkKeys * kCkYaw) * (int)sizeof(float)));
    return RTR_OK;
  }
  ~Coarse() {
    if (stream) (void)hipStreamSynchronize(stream);
    DevBuf* b[] = {&raw, &off, &cnt, &keys, &vox, &nvox, &bscore, &byaw, &cand, &cscore};
    for (DevBuf* x : b) x->release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int set_clouds(int n, const float* const* src, const size_t* ns, const float* const* tgt, const size_t* nt) {
    RTR_HIP(hipSetDevice(device));
    std::vector<uint32_t> ho(2 * n), hc(2 * n);
    size_t total = 0;
    for (int p = 0; p < n; ++p)
      for (int s = 0; s < 2; ++s) {
        const size_t k = s == 0 ? ns[p] : nt[p];
        if (k > 0xffffffffu || total + k > 0xffffffffu) {
          set_error("coarse proposer: clouds too large");
          return RTR_ERR_INVALID_ARG;
        }
        ho[2 * p + s] = (uint32_t)total;
        hc[2 * p + s] = (uint32_t)k;
        total += k;
      }
    RTR_TRY(raw.ensure(std::max<size_t>(total, 1) * 12));
    RTR_TRY(off.ensure((size_t)2 * n * 4));
    RTR_TRY(cnt.ensure((size_t)2 * n * 4));
    for (int p = 0; p < n; ++p) {
      if (ns[p]) RTR_HIP(hipMemcpy(raw.as<float>() + 3 * (size_t)ho[2 * p], src[p], ns[p] * 12, hipMemcpyHostToDevice));
      if (nt[p])
        RTR_HIP(hipMemcpy(raw.as<float>() + 3 * (size_t)ho[2 * p + 1], tgt[p], nt[p] * 12, hipMemcpyHostToDevice));
    }
    RTR_HIP(hipMemcpy(off.p, ho.data(), ho.size() * 4, hipMemcpyHostToDevice));
    RTR_HIP(hipMemcpy(cnt.p, hc.data(), hc.size() * 4, hipMemcpyHostToDevice));
    const size_t kk = (size_t)n * kCkKeys;
    RTR_TRY(keys.ensure((size_t)2 * kk * 3 * 4));
    RTR_TRY(vox.ensure(kk * kCkMaxVox * 2));
    RTR_TRY(nvox.ensure(kk * 4));
    RTR_TRY(bscore.ensure(kk * kCkKeys * 4));
    RTR_TRY(byaw.ensure(kk * kCkKeys * 4));
    n_pairs = n;
    return RTR_OK;
  }

  // keypoints, occupancy and the (i, j, yaw) scores; copies the per-pair
  // best scores and yaws to the host
  int score_pairs(float res, std::vector<float>& hs, std::vector<int>& hy) {
    const int n = n_pairs;
    hipLaunchKernelGGL(k_ck_keypoints, dim3(2 * n), dim3(256), 0, stream, raw.as<float>(), off.as<uint32_t>(),
                       cnt.as<uint32_t>(), keys.as<float>());
    hipLaunchKernelGGL(k_ck_occupancy, dim3(n * kCkKeys), dim3(256), 0, stream, raw.as<float>(), off.as<uint32_t>(),
                       cnt.as<uint32_t>(), keys.as<float>(), res, vox.as<uint16_t>(), nvox.as<int>());
    const size_t lds = (kCkTdf * kCkTdf * kCkTdf + kCkKeys * kCkYaw) * sizeof(float);
    YawTable yt;
    for (int k = 0; k < kCkYaw; ++k) {
      yt.c[k] = (float)std::cos(2.0 * M_PI * k / kCkYaw);
      yt.s[k] = (float)std::sin(2.0 * M_PI * k / kCkYaw);
    }
    hipLaunchKernelGGL(k_ck_score, dim3(n * kCkKeys), dim3(256), lds, stream, raw.as<float>(), off.as<uint32_t>(),
                       cnt.as<uint32_t>(), keys.as<float>(), res, vox.as<uint16_t>(), nvox.as<int>(), yt,
                       bscore.as<float>(), byaw.as<int>());
    RTR_HIP(hipGetLastError());
    const size_t nn = (size_t)n * kCkKeys * kCkKeys;
    hs.resize(nn);
    hy.resize(nn);
    hkeys.resize((size_t)2 * n * kCkKeys * 3);
    RTR_HIP(hipMemcpyAsync(hs.data(), bscore.p, nn * 4, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipMemcpyAsync(hy.data(), byaw.p, nn * 4, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipMemcpyAsync(hkeys.data(), keys.p, hkeys.size() * 4, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipStreamSynchronize(stream));
    return RTR_OK;
  }

  int propose(int M, float res, float tau, double* poses, double* out_score) {
    RTR_HIP(hipSetDevice(device));
    const int n = n_pairs;
    RTR_HIP(hipEventRecord(ev0, stream));
    std::vector<float> hs;
    std::vector<int> hy;
    RTR_TRY(score_pairs(res, hs, hy));
    // the M best (i, j) of every pair become candidate poses
    std::vector<double> hc((size_t)n * M * 16);
    std::vector<int> order(kCkKeys * kCkKeys);
    for (int p = 0; p < n; ++p) {
      const float* s = &hs[(size_t)p * kCkKeys * kCkKeys];
      for (int k = 0; k < kCkKeys * kCkKeys; ++k) order[k] = k;
      std::partial_sort(order.begin(), order.begin() + M, order.end(),
                        [&](int a, int b) { return s[a] < s[b] || (s[a] == s[b] && a < b); });
      for (int c = 0; c < M; ++c) {
        const int ij = order[c], i = ij / kCkKeys, j = ij - i * kCkKeys;
        const double ang = (double)hy[(size_t)p * kCkKeys * kCkKeys + ij] * (2.0 * M_PI / kCkYaw);
        const float* ki = &hkeys[((size_t)(2 * p + 1) * kCkKeys + i) * 3];
        const float* kj = &hkeys[((size_t)(2 * p) * kCkKeys + j) * 3];
        const double ca = std::cos(ang), sa = std::sin(ang);
        double* T = &hc[((size_t)p * M + c) * 16];
        for (int q = 0; q < 16; ++q) T[q] = 0.0;
        T[0] = ca;
        T[1] = sa;
        T[4] = -sa;
        T[5] = ca;
        T[10] = 1.0;
        T[15] = 1.0;
        // t = k_i - Rz k_j
        T[12] = ki[0] - (ca * kj[0] - sa * kj[1]);
        T[13] = ki[1] - (sa * kj[0] + ca * kj[1]);
        T[14] = ki[2] - kj[2];
      }
    }
    RTR_TRY(cand.ensure(hc.size() * 8));
    RTR_TRY(cscore.ensure((size_t)n * M * 8));
    RTR_HIP(hipMemcpyAsync(cand.p, hc.data(), hc.size() * 8, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_ck_consensus, dim3(n * M), dim3(256), 0, stream, raw.as<float>(), off.as<uint32_t>(),
                       cnt.as<uint32_t>(), cand.as<double>(), M, tau * tau, cscore.as<double>());
    RTR_HIP(hipGetLastError());
    std::vector<double> cs((size_t)n * M);
    RTR_HIP(hipMemcpyAsync(cs.data(), cscore.p, cs.size() * 8, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipEventRecord(ev1, stream));
    RTR_HIP(hipStreamSynchronize(stream));
    RTR_HIP(hipEventElapsedTime(&last_ms, ev0, ev1));
    for (int p = 0; p < n; ++p) {
      int b = 0;
      for (int c = 1; c < M; ++c)
        if (cs[(size_t)p * M + c] < cs[(size_t)p * M + b]) b = c;
      std::memcpy(poses + (size_t)p * 16, &hc[((size_t)p * M + b) * 16], 16 * 8);
      if (out_score) out_score[p] = cs[(size_t)p * M + b];
    }
    return RTR_OK;
  }
};

}  // namespace rtr

struct rtr_coarse {
  rtr::Coarse c;
};

extern "C" {

int rtr_coarse_create(rtr_coarse** out, int device) {
  using namespace rtr;
  if (!out) return RTR_ERR_INVALID_ARG;
  *out = nullptr;
  try {
    rtr_coarse* h = new rtr_coarse();
    const int rc = h->c.init(device);
    if (rc) {
      delete h;
      return rc;
    }
    *out = h;
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_coarse_destroy(rtr_coarse* h) {
  delete h;
  return RTR_OK;
}

int rtr_coarse_set_clouds(rtr_coarse* h, int n, const float* const* src, const size_t* ns, const float* const* tgt,
                          const size_t* nt) {
  using namespace rtr;
  if (!h || n < 1 || !src || !ns || !tgt || !nt) return RTR_ERR_INVALID_ARG;
  for (int p = 0; p < n; ++p)
    if ((ns[p] && !src[p]) || (nt[p] && !tgt[p])) return RTR_ERR_INVALID_ARG;
  try {
    return h->c.set_clouds(n, src, ns, tgt, nt);
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_coarse_propose(rtr_coarse* h, int candidates, float resolution, float tau, double* poses, double* scores) {
  using namespace rtr;
  if (!h || !poses || candidates < 1 || candidates > kCkKeys * kCkKeys || !(resolution > 0.f) || !(tau > 0.f))
    return RTR_ERR_INVALID_ARG;
  if (h->c.n_pairs < 1) {
    set_error("coarse proposer: no clouds set");
    return RTR_ERR_NOT_READY;
  }
  try {
    return h->c.propose(candidates, resolution, tau, poses, scores);
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_coarse_pair_scores(rtr_coarse* h, float resolution, float* scores, int* yaw, float* keypoints) {
  using namespace rtr;
  if (!h || !scores || !yaw || !(resolution > 0.f)) return RTR_ERR_INVALID_ARG;
  if (h->c.n_pairs < 1) {
    set_error("coarse proposer: no clouds set");
    return RTR_ERR_NOT_READY;
  }
  try {
    RTR_HIP(hipSetDevice(h->c.device));
    std::vector<float> hs;
    std::vector<int> hy;
    RTR_TRY(h->c.score_pairs(resolution, hs, hy));
    std::memcpy(scores, hs.data(), hs.size() * 4);
    std::memcpy(yaw, hy.data(), hy.size() * 4);
    if (keypoints) std::memcpy(keypoints, h->c.hkeys.data(), h->c.hkeys.size() * 4);
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_coarse_occupancy(rtr_coarse* h, uint16_t* voxels, int* counts) {
  using namespace rtr;
  if (!h || !voxels || !counts) return RTR_ERR_INVALID_ARG;
  if (h->c.n_pairs < 1) return RTR_ERR_NOT_READY;
  RTR_HIP(hipSetDevice(h->c.device));
  const size_t kk = (size_t)h->c.n_pairs * kCkKeys;
  RTR_HIP(hipMemcpy(voxels, h->c.vox.p, kk * kCkMaxVox * 2, hipMemcpyDeviceToHost));
  RTR_HIP(hipMemcpy(counts, h->c.nvox.p, kk * 4, hipMemcpyDeviceToHost));
  return RTR_OK;
}

float rtr_coarse_last_ms(rtr_coarse* h) { return h ? h->c.last_ms : 0.f; }

}  // extern "C"
