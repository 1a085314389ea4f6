// HBM copy probe: a plain streaming copy of a known number of bytes.  It
// calibrates the HBM byte counters read by the profiler (FETCH_SIZE and
// WRITE_SIZE of this kernel against its known traffic) and measures the
// bandwidth the device actually sustains.
#include <hip/hip_runtime.h>

#include "../../include/rtr.h"
#include "engine.hpp"

namespace rtr {

__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace rtr

extern "C" int rtr_copy_probe(int device, size_t bytes, int reps, float* ms_avg) {
  using namespace rtr;
  if (bytes < 4096 || reps < 1 || !ms_avg) return RTR_ERR_INVALID_ARG;
  RTR_HIP(hipSetDevice(device));
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s = nullptr;
  int rc = RTR_OK;
  float ms = 0.f;
  hipError_t err = hipMalloc(&a, n * 16);
  if (err == hipSuccess) err = hipMalloc(&b, n * 16);
  if (err == hipSuccess) err = hipMemset(a, 0, n * 16);
  if (err == hipSuccess) err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  int cus = 256;
  if (err == hipSuccess) err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  const unsigned blocks = (unsigned)(cus * 8);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, (const float4*)a, (float4*)b, n);
    err = hipEventRecord(e0, s);
  }
  for (int r = 0; err == hipSuccess && r < reps; ++r) {
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, (const float4*)a, (float4*)b, n);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e1, s);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  if (err != hipSuccess) {
    set_error("copy probe: %s", hipGetErrorString(err));
    rc = RTR_ERR_DEVICE;
  } else {
    *ms_avg = ms / (float)reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}
