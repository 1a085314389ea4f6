// Voxel-hash map: merges transformed clouds into a scene map.
//
// The reference "merges" by transforming a cloud, saving it as a PCD and
// overlaying files in a viewer (RealTimeRobot.cpp:104-114).  Here every
// point is transformed on the device (PCL transformPointCloud order), keyed
// by floor(p * inv_voxel) per axis (the truncation convention of
// key_point.h:143-160 / matching.h:181-184, floored for negative
// coordinates), and accumulated into an open-addressing hash table in HBM:
// per voxel a 64-bit key, the FP32 sum of (p - voxel corner) and a count.
// Offsets from the corner keep the FP32 sums small, so the centroid keeps
// full float precision however many frames land in the voxel.
//
// Integration kernel: each workgroup takes a contiguous run of tiles of 512
// consecutive points (a few image rows of a depth frame), stages each tile
// in LDS with coalesced loads, and merges the points voxel by voxel in an LDS
// hash table that lives across the run's tiles (neighbouring pixels, and
// neighbouring rows, share voxels); the table is flushed into the global one
// when it holds more than a quarter of its slots ahead of the next tile (so
// LDS probes stay short), and at the end of the run.  The global table then
// sees one insert and one set of atomics per distinct voxel of each flush,
// not per point.  Global probes are bounded (the table holds at most half its
// slots).  Many frames (each with its own transform) integrate in one
// launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rtr.h"
#include "engine.hpp"
#include "geom.hpp"

namespace rtr {

constexpr uint64_t kEmpty = ~0ull;
constexpr int kBias = 1 << 20;

struct MapDev {
  unsigned long long* keys;  // capacity slots, kEmpty when free
  float4* acc;               // (sum dx, sum dy, sum dz, count bits)
  uint32_t* counters;        // [0] occupied slots, [1] dropped points, [2] overflow
  uint64_t mask;             // capacity - 1 (power of two)
  uint32_t max_voxels;
  float h, inv_h;
};

__device__ __forceinline__ uint64_t mix64(uint64_t k) {
  // SplitMix64 finaliser (Steele, Lea, Flood 2014)
  k ^= k >> 30;
  k *= 0xbf58476d1ce4e5b9ull;
  k ^= k >> 27;
  k *= 0x94d049bb133111ebull;
  k ^= k >> 31;
  return k;
}

RTR_HD bool voxel_key(float x, float y, float z, float inv_h, uint64_t* key, int* k3) {
  const float v[3] = {x * inv_h, y * inv_h, z * inv_h};
  uint64_t kk = 0;
  for (int a = 0; a < 3; ++a) {
    if (!(v[a] - v[a] == 0.f)) return false;  // inf or NaN
    const float f = floorf(v[a]);
    if (f < -(float)kBias || f >= (float)kBias) return false;
    const int k = (int)f;
    k3[a] = k;
    kk |= (uint64_t)(uint32_t)(k + kBias) << (21 * a);
  }
  *key = kk;
  return true;
}

constexpr int kMapBlock = 256;
constexpr int kMapTile = 512;    // points per workgroup tile
constexpr int kMapSlots = 2048;  // the run's LDS hash table
constexpr int kMapFlushAt = kMapSlots / 2 - kMapTile;  // flush before a tile when more keys are held
constexpr int kMapMinRun = 4;    // tiles per workgroup at least (the merging spans the run)
constexpr int kMaxProbe = 128;   // global probe bound (the table is at most half full)
static_assert(kMapFlushAt + kMapTile < kMapSlots, "every tile key must find a free LDS slot");

// One frame of an integration launch: packed xyz on the device, its point
// count, its first tile in the launch and the float rows of its transform.
struct MapFrame {
  const float* xyz;
  uint32_t n;
  uint32_t tile0;
  float m[12];
  int has_T;
  int pad_;
};

// One merged voxel of a tile into the global table: (sum of corner offsets,
// count).  False when the probe bound is reached (table too full).
__device__ __forceinline__ bool global_add(const MapDev& m, uint64_t key, float sx, float sy, float sz, uint32_t cnt) {
  uint64_t slot = mix64(key) & m.mask;
  for (int probe = 0; probe < kMaxProbe; ++probe) {
    // a stale empty read only costs a CAS; a key once written never changes
    unsigned long long cur = m.keys[slot];
    if (cur == kEmpty) {
      cur = atomicCAS(&m.keys[slot], (unsigned long long)kEmpty, (unsigned long long)key);
      if (cur == kEmpty) {
        const uint32_t occ = atomicAdd(&m.counters[0], 1u);
        if (occ >= m.max_voxels) atomicOr(&m.counters[2], 1u);
        cur = key;
      }
    }
    if (cur == key) {
      float* a = reinterpret_cast<float*>(&m.acc[slot]);
      atomicAdd(a + 0, sx);
      atomicAdd(a + 1, sy);
      atomicAdd(a + 2, sz);
      atomicAdd(reinterpret_cast<uint32_t*>(a + 3), cnt);
      return true;
    }
    slot = (slot + 1) & m.mask;
  }
  return false;
}

// The run's merged voxels into the global table; the LDS table is left
// empty.  Every thread of the workgroup calls it.
__device__ __forceinline__ void map_flush(const MapDev& m, unsigned long long* s_key, float4* s_acc,
                                          uint32_t& dropped, uint32_t& overflow) {
  for (int i = threadIdx.x; i < kMapSlots; i += kMapBlock) {
    const unsigned long long k = s_key[i];
    if (k == kEmpty) continue;
    const float4 a = s_acc[i];
    const uint32_t c = __float_as_uint(a.w);
    if (!global_add(m, k, a.x, a.y, a.z, c)) {
      dropped += c;
      overflow = 1;
    }
    s_key[i] = kEmpty;
    s_acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

__global__ __launch_bounds__(kMapBlock) void k_map_integrate(MapDev m, const MapFrame* frames, int n_frames,
                                                             uint32_t n_tiles, uint32_t run) {
  __shared__ float s_xyz[3 * kMapTile];
  __shared__ unsigned long long s_key[kMapSlots];
  __shared__ float4 s_acc[kMapSlots];
  __shared__ int s_frame;
  __shared__ uint32_t s_used;
  const int tid = threadIdx.x;
  uint32_t dropped = 0, overflow = 0;
  for (int i = tid; i < kMapSlots; i += kMapBlock) {
    s_key[i] = kEmpty;
    s_acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (tid == 0) s_used = 0u;
  const uint32_t t0 = blockIdx.x * run;
  const uint32_t t1 = min(t0 + run, n_tiles);
  for (uint32_t tile = t0; tile < t1; ++tile) {
    __syncthreads();  // the previous tile's inserts (and s_used) are complete
    if (s_used > (uint32_t)kMapFlushAt) {
      map_flush(m, s_key, s_acc, dropped, overflow);
      __syncthreads();
      if (tid == 0) s_used = 0u;
    }
    if (tid == 0) {
      // the frame of this tile: the last one whose first tile is <= tile
      int lo = 0, hi = n_frames - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (frames[mid].tile0 <= tile)
          lo = mid;
        else
          hi = mid - 1;
      }
      s_frame = lo;
    }
    __syncthreads();
    const MapFrame* f = frames + s_frame;
    const uint32_t base = (tile - f->tile0) * (uint32_t)kMapTile;
    const uint32_t cnt = min((uint32_t)kMapTile, f->n - base);
    // coalesced staging: consecutive lanes load consecutive floats
    const float* src = f->xyz + 3 * (size_t)base;
    for (uint32_t j = tid; j < 3 * cnt; j += kMapBlock) s_xyz[j] = src[j];
    const int has_T = f->has_T;
    float mt[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) mt[k] = f->m[k];
    __syncthreads();
    for (uint32_t p = tid; p < cnt; p += kMapBlock) {
      float px = s_xyz[3 * p], py = s_xyz[3 * p + 1], pz = s_xyz[3 * p + 2];
      if (has_T) apply_f32(mt, px, py, pz, px, py, pz);
      uint64_t key;
      int k3[3];
      if (!voxel_key(px, py, pz, m.inv_h, &key, k3)) {
        ++dropped;
        continue;
      }
      const float dx = px - (float)k3[0] * m.h;
      const float dy = py - (float)k3[1] * m.h;
      const float dz = pz - (float)k3[2] * m.h;
      // the table holds at most kMapFlushAt + kMapTile < kMapSlots keys: a
      // free or matching slot always exists within kMapSlots probes
      uint32_t sl = (uint32_t)(mix64(key) & (uint64_t)(kMapSlots - 1));
      for (int probe = 0; probe < kMapSlots; ++probe) {
        unsigned long long cur = s_key[sl];
        if (cur == kEmpty) {
          cur = atomicCAS(&s_key[sl], (unsigned long long)kEmpty, (unsigned long long)key);
          if (cur == kEmpty) {
            atomicAdd(&s_used, 1u);
            cur = key;
          }
        }
        if (cur == key) {
          atomicAdd(&s_acc[sl].x, dx);
          atomicAdd(&s_acc[sl].y, dy);
          atomicAdd(&s_acc[sl].z, dz);
          atomicAdd(reinterpret_cast<uint32_t*>(&s_acc[sl].w), 1u);
          break;
        }
        sl = (sl + 1) & (uint32_t)(kMapSlots - 1);
      }
    }
  }
  __syncthreads();
  map_flush(m, s_key, s_acc, dropped, overflow);
  if (dropped) atomicAdd(&m.counters[1], dropped);
  if (overflow) atomicOr(&m.counters[2], 1u);
}

__global__ void k_map_compact(MapDev m, unsigned long long* okeys, float4* oacc, uint32_t* ocount,
                              uint32_t cap) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= m.mask;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = m.keys[s];
    if (k == kEmpty) continue;
    const uint32_t o = atomicAdd(ocount, 1u);
    if (o < cap) {
      okeys[o] = k;
      oacc[o] = m.acc[s];
    }
  }
}

__global__ void k_map_clear(MapDev m) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= m.mask;
       s += (uint64_t)gridDim.x * blockDim.x) {
    m.keys[s] = kEmpty;
    m.acc[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) m.counters[threadIdx.x] = 0u;
}

struct VoxelMap {
  int device;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool profiling = false;
  float last_ms = 0.f;
  int cu_count = 0;
  DevBuf keys, acc, counters, in, T, okeys, oacc, ocount;
  MapDev dev;
  size_t capacity = 0;
  uint64_t dropped_total = 0;

  int init(int dev_id, float voxel, size_t cap) {
    device = dev_id;
    int n = 0;
    RTR_HIP(hipGetDeviceCount(&n));
    if (dev_id < 0 || dev_id >= n) {
      set_error("device %d not available", dev_id);
      return RTR_ERR_DEVICE;
    }
    RTR_HIP(hipSetDevice(device));
    RTR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    RTR_HIP(hipEventCreate(&ev[0]));
    RTR_HIP(hipEventCreate(&ev[1]));
    RTR_HIP(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, device));
    uint64_t slots = 1024;
    while (slots < 2 * (uint64_t)cap) slots <<= 1;
    capacity = cap;
    int rc = keys.ensure(slots * 8);
    if (!rc) rc = acc.ensure(slots * 16);
    if (!rc) rc = counters.ensure(16);
    if (!rc) rc = T.ensure(64);
    if (rc) return rc;
    dev.keys = keys.as<unsigned long long>();
    dev.acc = acc.as<float4>();
    dev.counters = counters.as<uint32_t>();
    dev.mask = slots - 1;
    dev.max_voxels = (uint32_t)cap;
    dev.h = voxel;
    dev.inv_h = (float)(1.0 / (double)voxel);
    return clear();
  }
  int clear() {
    RTR_HIP(hipSetDevice(device));
    hipLaunchKernelGGL(k_map_clear, dim3(1024), dim3(256), 0, stream, dev);
    RTR_HIP(hipGetLastError());
    RTR_HIP(hipStreamSynchronize(stream));
    dropped_total = 0;
    return RTR_OK;
  }
  ~VoxelMap() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    DevBuf* b[] = {&keys, &acc, &counters, &in, &T, &okeys, &oacc, &ocount};
    for (DevBuf* x : b) x->release();
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int read_counters(uint32_t c[4]) {
    RTR_HIP(hipMemcpyAsync(c, counters.p, 16, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipStreamSynchronize(stream));
    return RTR_OK;
  }
  // nf frames in one launch: frame f's points (packed, or with a byte
  // stride) moved by T[16 f ..] (column-major; T or its entries nullable
  // means identity)
  int integrate_frames(int nf, const float* const* xyz, const size_t* n, size_t stride, const double* Tcm) {
    RTR_HIP(hipSetDevice(device));
    if (stride == 0) stride = 12;
    size_t total = 0, tiles = 0;
    for (int f = 0; f < nf; ++f) {
      total += n[f];
      tiles += (n[f] + kMapTile - 1) / kMapTile;
    }
    if (total == 0) return RTR_OK;
    if (total >= (1ull << 32) || tiles >= (1ull << 32)) {
      set_error("too many points in one call");
      return RTR_ERR_CAPACITY;
    }
    RTR_TRY(in.ensure(total * 12));
    RTR_TRY(T.ensure(sizeof(MapFrame) * (size_t)nf));
    std::vector<MapFrame> fr(nf);
    std::vector<float> packed;
    size_t off = 0;
    uint32_t tile0 = 0;
    for (int f = 0; f < nf; ++f) {
      MapFrame& d = fr[f];
      memset(&d, 0, sizeof(d));
      d.xyz = in.as<float>() + 3 * off;
      d.n = (uint32_t)n[f];
      d.tile0 = tile0;
      d.has_T = Tcm ? 1 : 0;
      if (Tcm) tf_rows_f32(Tcm + 16 * (size_t)f, d.m);
      if (n[f]) {
        const float* src = xyz[f];
        if (stride != 12) {
          packed.resize(n[f] * 3);
          const char* b = reinterpret_cast<const char*>(xyz[f]);
          for (size_t i = 0; i < n[f]; ++i) memcpy(packed.data() + 3 * i, b + i * stride, 12);
          src = packed.data();
        }
        // synchronous for pageable memory: the staging buffer is reused
        RTR_HIP(hipMemcpyAsync(in.as<float>() + 3 * off, src, n[f] * 12, hipMemcpyHostToDevice, stream));
        if (stride != 12) RTR_HIP(hipStreamSynchronize(stream));
      }
      off += n[f];
      tile0 += (uint32_t)((n[f] + kMapTile - 1) / kMapTile);
    }
    RTR_HIP(hipMemcpyAsync(T.p, fr.data(), sizeof(MapFrame) * (size_t)nf, hipMemcpyHostToDevice, stream));
    // contiguous runs of at least kMapMinRun tiles, at most 8 workgroups per CU
    size_t blocks = std::min<size_t>((tiles + kMapMinRun - 1) / kMapMinRun, (size_t)std::max(cu_count, 1) * 8);
    const uint32_t run = (uint32_t)((tiles + blocks - 1) / blocks);
    blocks = (tiles + run - 1) / run;
    if (profiling) RTR_HIP(hipEventRecord(ev[0], stream));
    hipLaunchKernelGGL(k_map_integrate, dim3((unsigned)blocks), dim3(kMapBlock), 0, stream, dev, T.as<MapFrame>(), nf,
                       (uint32_t)tiles, run);
    RTR_HIP(hipGetLastError());
    if (profiling) RTR_HIP(hipEventRecord(ev[1], stream));
    uint32_t after[4];
    RTR_TRY(read_counters(after));
    if (profiling) RTR_HIP(hipEventElapsedTime(&last_ms, ev[0], ev[1]));
    dropped_total = after[1];
    if (after[2]) {
      set_error("voxel map capacity exceeded (%zu voxels)", capacity);
      return RTR_ERR_CAPACITY;
    }
    return RTR_OK;
  }
  int integrate(const float* xyz, size_t n, size_t stride, const double* Tcm) {
    return integrate_frames(1, &xyz, &n, stride, Tcm);
  }
  int size(size_t* n) {
    uint32_t c[4];
    int rc = read_counters(c);
    if (rc) return rc;
    *n = c[0];
    return RTR_OK;
  }
  // host copy of occupied voxels sorted by key: keys, absolute sums (double), counts
  int gather(std::vector<uint64_t>& k, std::vector<double>& sums, std::vector<uint32_t>& cnt) {
    RTR_HIP(hipSetDevice(device));
    size_t nv;
    int rc = size(&nv);
    if (rc) return rc;
    if (nv == 0) {
      k.clear();
      sums.clear();
      cnt.clear();
      return RTR_OK;
    }
    rc = okeys.ensure(nv * 8);
    if (!rc) rc = oacc.ensure(nv * 16);
    if (!rc) rc = ocount.ensure(4);
    if (rc) return rc;
    launch_fill_u32(ocount.as<uint32_t>(), 0u, 1, stream);
    hipLaunchKernelGGL(k_map_compact, dim3(1024), dim3(256), 0, stream, dev,
                       okeys.as<unsigned long long>(), oacc.as<float4>(), ocount.as<uint32_t>(),
                       (uint32_t)nv);
    RTR_HIP(hipGetLastError());
    std::vector<uint64_t> hk(nv);
    std::vector<float4> ha(nv);
    RTR_HIP(hipMemcpyAsync(hk.data(), okeys.p, nv * 8, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipMemcpyAsync(ha.data(), oacc.p, nv * 16, hipMemcpyDeviceToHost, stream));
    RTR_HIP(hipStreamSynchronize(stream));
    std::vector<uint32_t> order(nv);
    for (size_t i = 0; i < nv; ++i) order[i] = (uint32_t)i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return hk[a] < hk[b]; });
    k.resize(nv);
    sums.resize(3 * nv);
    cnt.resize(nv);
    for (size_t i = 0; i < nv; ++i) {
      const uint32_t o = order[i];
      k[i] = hk[o];
      uint32_t c;
      memcpy(&c, &ha[o].w, 4);
      cnt[i] = c;
      const float s[3] = {ha[o].x, ha[o].y, ha[o].z};
      for (int a = 0; a < 3; ++a) {
        const int kk = (int)((hk[o] >> (21 * a)) & 0x1fffffull) - kBias;
        const float corner = (float)kk * dev.h;
        sums[3 * i + a] = (double)corner * c + (double)s[a];
      }
    }
    return RTR_OK;
  }
};

}  // namespace rtr

using rtr::set_error;
using rtr::VoxelMap;

struct rtr_map {
  VoxelMap m;
};

extern "C" {

int rtr_map_create(rtr_map** out, int device, float voxel_size, size_t capacity) {
  try {
    if (!out || !(voxel_size > 0) || capacity == 0 || capacity > (1ull << 31)) {
      set_error("invalid voxel map arguments (voxel %g, capacity %zu)", voxel_size, capacity);
      return RTR_ERR_INVALID_ARG;
    }
    *out = nullptr;
    rtr_map* m = new rtr_map();
    int rc = m->m.init(device, voxel_size, capacity);
    if (rc) {
      delete m;
      return rc;
    }
    *out = m;
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_map_destroy(rtr_map* m) {
  delete m;
  return RTR_OK;
}

int rtr_map_clear(rtr_map* m) {
  if (!m) return RTR_ERR_INVALID_ARG;
  return m->m.clear();
}

int rtr_map_integrate(rtr_map* m, const float* xyz, size_t n, size_t stride, const double T[16]) {
  if (!m || (n && !xyz) || (stride != 0 && stride < 12)) return RTR_ERR_INVALID_ARG;
  try {
    return m->m.integrate(xyz, n, stride, T);
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_map_integrate_frames(rtr_map* m, int n_frames, const float* const* xyz, const size_t* n, size_t stride,
                             const double* T) {
  if (!m || n_frames < 1 || !xyz || !n || (stride != 0 && stride < 12)) return RTR_ERR_INVALID_ARG;
  for (int f = 0; f < n_frames; ++f)
    if (n[f] && !xyz[f]) return RTR_ERR_INVALID_ARG;
  try {
    return m->m.integrate_frames(n_frames, xyz, n, stride, T);
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_map_set_profiling(rtr_map* m, int on) {
  if (!m) return RTR_ERR_INVALID_ARG;
  m->m.profiling = on != 0;
  return RTR_OK;
}

int rtr_map_last_ms(rtr_map* m, float* ms) {
  if (!m || !ms) return RTR_ERR_INVALID_ARG;
  *ms = m->m.last_ms;
  return RTR_OK;
}

int rtr_map_size(rtr_map* m, size_t* n) {
  if (!m || !n) return RTR_ERR_INVALID_ARG;
  return m->m.size(n);
}

int rtr_map_dropped(rtr_map* m, uint64_t* n) {
  if (!m || !n) return RTR_ERR_INVALID_ARG;
  *n = m->m.dropped_total;
  return RTR_OK;
}

int rtr_map_extract(rtr_map* m, int64_t* keys, float* xyz, uint32_t* counts, size_t cap,
                    size_t* n_out, int as_centroids) {
  if (!m || !n_out) return RTR_ERR_INVALID_ARG;
  try {
    std::vector<uint64_t> k;
    std::vector<double> s;
    std::vector<uint32_t> c;
    int rc = m->m.gather(k, s, c);
    if (rc) return rc;
    *n_out = k.size();
    if (k.size() > cap) {
      set_error("extract capacity %zu < %zu voxels", cap, k.size());
      return RTR_ERR_CAPACITY;
    }
    for (size_t i = 0; i < k.size(); ++i) {
      if (keys) keys[i] = (int64_t)k[i];
      if (counts) counts[i] = c[i];
      if (xyz)
        for (int a = 0; a < 3; ++a)
          xyz[3 * i + a] = as_centroids ? (float)(s[3 * i + a] / (double)c[i]) : (float)s[3 * i + a];
    }
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_voxel_downsample(int device, const float* xyz, size_t n, size_t stride, float leaf,
                         float* out, size_t cap, size_t* n_out) {
  if ((n && !xyz) || !(leaf > 0) || !n_out) return RTR_ERR_INVALID_ARG;
  try {
    VoxelMap vm;
    int rc = vm.init(device, leaf, n > 0 ? n : 1);
    if (rc) return rc;
    rc = vm.integrate(xyz, n, stride, nullptr);
    if (rc) return rc;
    std::vector<uint64_t> k;
    std::vector<double> s;
    std::vector<uint32_t> c;
    rc = vm.gather(k, s, c);
    if (rc) return rc;
    const size_t nv = k.size();
    *n_out = nv;
    if (nv > cap) {
      set_error("downsample capacity %zu < %zu voxels", cap, nv);
      return RTR_ERR_CAPACITY;
    }
    // PCL VoxelGrid order: linear index over the ijk bounding box, x fastest
    int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
    std::vector<int64_t> ijk(3 * nv);
    for (size_t i = 0; i < nv; ++i)
      for (int a = 0; a < 3; ++a) {
        const int64_t v = (int64_t)((k[i] >> (21 * a)) & 0x1fffffull) - (1 << 20);
        ijk[3 * i + a] = v;
        lo[a] = std::min(lo[a], v);
        hi[a] = std::max(hi[a], v);
      }
    const int64_t dx = hi[0] - lo[0] + 1, dy = hi[1] - lo[1] + 1;
    std::vector<std::pair<int64_t, uint32_t>> lin(nv);
    for (size_t i = 0; i < nv; ++i)
      lin[i] = {(ijk[3 * i] - lo[0]) + (ijk[3 * i + 1] - lo[1]) * dx + (ijk[3 * i + 2] - lo[2]) * dx * dy,
                (uint32_t)i};
    std::sort(lin.begin(), lin.end());
    for (size_t o = 0; o < nv; ++o) {
      const uint32_t i = lin[o].second;
      for (int a = 0; a < 3; ++a) out[3 * o + a] = (float)(s[3 * i + a] / (double)c[i]);
    }
    return RTR_OK;
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

}  // extern "C"
