// Truncated distance field of an occupancy grid: for every voxel of a
// dx x dy x dz grid (x fastest: index x + y*dx + z*dx*dy) the squared
// Euclidean distance, in voxel units, to the nearest occupied voxel, capped
// at trunc2.  The reference computes the same quantity by brute force over
// every occupied voxel (kernel.cu:9-31: a 30^3 grid, cap 900, called from
// key_point.h:287-308 on a key point's occupied octree voxels and read at
// matching.h:169-188).  Here the squared distance transform is separable:
// three 1-D passes, each d'(a) = min_t (a - t)^2 + d(t) along one axis.  Every
// value below 2^24 is an integer held exactly in FP32 (larger ones only lose
// low bits and can never undercut a smaller exact value), so capped results
// equal the brute force bit for bit.  Grid dimensions are arbitrary; occupied
// coordinates outside the grid are ignored.  The points variant maps a point
// to floor((p - origin) * (1 / resolution)) in FP32.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/rtr.h"
#include "engine.hpp"

namespace rtr {

constexpr float kTdfInf = 1.0e30f;  // far above any squared grid distance

__global__ void k_tdf_fill(float* g, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    g[i] = kTdfInf;
}

__global__ void k_tdf_mark(const int32_t* occ, size_t n_occ, int dx, int dy, int dz, float* g) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_occ; i += (size_t)gridDim.x * blockDim.x) {
    const int x = occ[3 * i], y = occ[3 * i + 1], z = occ[3 * i + 2];
    if (x < 0 || y < 0 || z < 0 || x >= dx || y >= dy || z >= dz) continue;
    g[x + (size_t)dx * ((size_t)y + (size_t)dy * z)] = 0.f;
  }
}

// Points -> occupied voxels: floor((p - origin) / res) inside the grid.
__global__ void k_tdf_mark_points(const float* xyz, size_t n, float ox, float oy, float oz, float inv_res, int dx,
                                  int dy, int dz, float* g) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float fx = (xyz[3 * i] - ox) * inv_res, fy = (xyz[3 * i + 1] - oy) * inv_res,
                fz = (xyz[3 * i + 2] - oz) * inv_res;
    if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f && fx < (float)dx && fy < (float)dy && fz < (float)dz)) continue;
    const int x = min((int)fx, dx - 1), y = min((int)fy, dy - 1), z = min((int)fz, dz - 1);
    g[x + (size_t)dx * ((size_t)y + (size_t)dy * z)] = 0.f;
  }
}

// One pass along axis A (0 = x, 1 = y, 2 = z): out(a) = min_t (a - t)^2 + in(t)
// over the line through the voxel; the line is staged in LDS one 1024-value
// tile at a time.  The last pass applies the cap.
template <int A>
__global__ __launch_bounds__(256) void k_tdf_pass(const float* in, float* out, int dx, int dy, int dz, float cap) {
  __shared__ float line[1024];
  const int len = A == 0 ? dx : (A == 1 ? dy : dz);
  // lines are enumerated by the two other coordinates
  const size_t n_lines = (size_t)dx * dy * dz / (size_t)len;
  for (size_t li = blockIdx.x; li < n_lines; li += gridDim.x) {
    size_t base, stride;
    if (A == 0) {
      base = li * dx;
      stride = 1;
    } else if (A == 1) {
      const size_t x = li % dx, z = li / dx;
      base = x + (size_t)dx * dy * z;
      stride = dx;
    } else {
      base = li;
      stride = (size_t)dx * dy;
    }
    for (int a0 = 0; a0 < len; a0 += 256) {
      const int a = a0 + threadIdx.x;
      float best = kTdfInf;
      for (int t0 = 0; t0 < len; t0 += 1024) {
        const int tn = min(1024, len - t0);
        __syncthreads();
        for (int t = threadIdx.x; t < tn; t += 256) line[t] = in[base + (size_t)(t0 + t) * stride];
        __syncthreads();
        if (a < len) {
          for (int t = 0; t < tn; ++t) {
            const float d = (float)(a - (t0 + t));
            best = fminf(best, d * d + line[t]);
          }
        }
      }
      if (a < len) out[base + (size_t)a * stride] = cap > 0.f ? fminf(best, cap) : best;
    }
  }
}

static int tdf_run(int device, const int dims[3], float trunc2, float* out, const int32_t* occ, size_t n_occ,
                   const float* xyz, size_t n_pts, size_t stride, const float origin[3], float res) {
  RTR_HIP(hipSetDevice(device));
  const int dx = dims[0], dy = dims[1], dz = dims[2];
  const size_t n = (size_t)dx * dy * dz;
  DevBuf a, b, src;
  int rc = a.ensure(n * 4);
  if (!rc) rc = b.ensure(n * 4);
  const size_t src_bytes = occ ? n_occ * 12 : n_pts * 12;
  if (!rc && src_bytes) rc = src.ensure(src_bytes);
  if (rc) return rc;
  hipStream_t s = nullptr;
  hipLaunchKernelGGL(k_tdf_fill, dim3(1024), dim3(256), 0, s, a.as<float>(), n);
  if (occ && n_occ) {
    RTR_HIP(hipMemcpy(src.p, occ, n_occ * 12, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tdf_mark, dim3(256), dim3(256), 0, s, src.as<int32_t>(), n_occ, dx, dy, dz, a.as<float>());
  } else if (xyz && n_pts) {
    std::vector<float> packed(3 * n_pts);
    const char* base = reinterpret_cast<const char*>(xyz);
    for (size_t i = 0; i < n_pts; ++i) memcpy(&packed[3 * i], base + i * stride, 12);
    RTR_HIP(hipMemcpy(src.p, packed.data(), n_pts * 12, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tdf_mark_points, dim3(256), dim3(256), 0, s, src.as<float>(), n_pts, origin[0], origin[1],
                       origin[2], 1.0f / res, dx, dy, dz, a.as<float>());
  }
  const unsigned blocks = 2048;
  hipLaunchKernelGGL(k_tdf_pass<0>, dim3(blocks), dim3(256), 0, s, a.as<float>(), b.as<float>(), dx, dy, dz, 0.f);
  hipLaunchKernelGGL(k_tdf_pass<1>, dim3(blocks), dim3(256), 0, s, b.as<float>(), a.as<float>(), dx, dy, dz, 0.f);
  hipLaunchKernelGGL(k_tdf_pass<2>, dim3(blocks), dim3(256), 0, s, a.as<float>(), b.as<float>(), dx, dy, dz,
                     trunc2);
  RTR_HIP(hipGetLastError());
  RTR_HIP(hipMemcpy(out, b.p, n * 4, hipMemcpyDeviceToHost));
  a.release();
  b.release();
  src.release();
  return RTR_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant (rtr_tdf_noalloc): the caller's device buffers hold
// the occupied voxels and the field; the intermediate grid of the separable
// passes is this file-scope device array, one per GPU, held until the process
// ends.  Host code never passes `scratch` itself to a kernel: its device
// address, per device, comes from hipGetSymbolAddress on the first call on
// that device.
// ---------------------------------------------------------------------------
constexpr size_t kTdfScratchMax = (size_t)1 << 21;  // floats: 8 MiB, grids of up to 2^21 voxels (128^3)
__device__ __align__(256) float scratch[kTdfScratchMax];

constexpr int kMaxDevices = 64;
static float* g_scratch_dev[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mu;

static int scratch_of(int device, float** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    RTR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("hipGetSymbolAddress returned no address for the TDF scratch array");
      return RTR_ERR_DEVICE;
    }
    g_scratch_dev[device] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[device];
  return RTR_OK;
}

static bool tdf_dims_ok(const int dims[3]) {
  if (!dims) return false;
  size_t n = 1;
  for (int k = 0; k < 3; ++k) {
    if (dims[k] < 1 || dims[k] > 4096) return false;
    n *= (size_t)dims[k];
  }
  return n <= ((size_t)1 << 31);
}

}  // namespace rtr

extern "C" {

int rtr_tdf(int device, const int32_t* occ, size_t n_occ, const int dims[3], float trunc2, float* out) {
  using namespace rtr;
  if (!out || (n_occ && !occ) || !tdf_dims_ok(dims) || !(trunc2 > 0.f)) {
    set_error("rtr_tdf: invalid arguments");
    return RTR_ERR_INVALID_ARG;
  }
  try {
    return tdf_run(device, dims, trunc2, out, occ, n_occ, nullptr, 0, 12, nullptr, 1.f);
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

int rtr_tdf_noalloc(int device, const int32_t* occ_dev, size_t n_occ, const int dims[3], float trunc2, float* out_dev,
                    void* stream) {
  using namespace rtr;
  if (!out_dev || (n_occ && !occ_dev) || !tdf_dims_ok(dims) || !(trunc2 > 0.f)) {
    set_error("rtr_tdf_noalloc: invalid arguments");
    return RTR_ERR_INVALID_ARG;
  }
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  if (n > kTdfScratchMax) {
    set_error("rtr_tdf_noalloc: a %d x %d x %d grid has %zu voxels, above the allocation-free limit of %zu",
              dims[0], dims[1], dims[2], n, kTdfScratchMax);
    return RTR_ERR_CAPACITY;
  }
  int n_dev = 0;
  RTR_HIP(hipGetDeviceCount(&n_dev));
  if (device < 0 || device >= n_dev || device >= kMaxDevices) {
    set_error("rtr_tdf_noalloc: device %d not available", device);
    return RTR_ERR_DEVICE;
  }
  RTR_HIP(hipSetDevice(device));
  float* g = nullptr;
  RTR_TRY(scratch_of(device, &g));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int dx = dims[0], dy = dims[1], dz = dims[2];
  // every scratch voxel is written by the fill before it is read: nothing
  // left by an earlier call on this device is used
  hipLaunchKernelGGL(k_tdf_fill, dim3(1024), dim3(256), 0, s, g, n);
  if (n_occ) hipLaunchKernelGGL(k_tdf_mark, dim3(256), dim3(256), 0, s, occ_dev, n_occ, dx, dy, dz, g);
  const unsigned blocks = 2048;
  hipLaunchKernelGGL(k_tdf_pass<0>, dim3(blocks), dim3(256), 0, s, g, out_dev, dx, dy, dz, 0.f);
  hipLaunchKernelGGL(k_tdf_pass<1>, dim3(blocks), dim3(256), 0, s, out_dev, g, dx, dy, dz, 0.f);
  hipLaunchKernelGGL(k_tdf_pass<2>, dim3(blocks), dim3(256), 0, s, g, out_dev, dx, dy, dz, trunc2);
  RTR_HIP(hipGetLastError());
  return RTR_OK;
}

int rtr_tdf_points(int device, const float* xyz, size_t n, size_t stride_bytes, const float origin[3],
                   float resolution, const int dims[3], float trunc2, float* out) {
  using namespace rtr;
  if (!out || (n && !xyz) || !origin || !(resolution > 0.f) || !tdf_dims_ok(dims) || !(trunc2 > 0.f)) {
    set_error("rtr_tdf_points: invalid arguments");
    return RTR_ERR_INVALID_ARG;
  }
  if (stride_bytes == 0) stride_bytes = 12;
  if (stride_bytes < 12) {
    set_error("rtr_tdf_points: stride below 12 bytes");
    return RTR_ERR_INVALID_ARG;
  }
  try {
    return tdf_run(device, dims, trunc2, out, nullptr, 0, xyz, n, stride_bytes, origin, resolution);
  } catch (...) {
    set_error("host allocation failed");
    return RTR_ERR_CAPACITY;
  }
}

}  // extern "C"
