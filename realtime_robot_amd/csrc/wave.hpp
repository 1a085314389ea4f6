// Wave64 reductions built from DPP row operations (gfx9 encodings).
//
// Within each 16-lane row, four DPP steps give every lane the row's total:
// quad_perm [1,0,3,2] and [2,3,0,1] (partners at lane distance 1 and 2), then
// row_half_mirror (lane i of an 8-lane half row reads lane 7 - i, the other
// quad) and row_mirror (lane i reads 15 - i, the other half row).  Each step
// combines a lane with a partner whose value is already identical across its
// group, and the operations are commutative, so after the four steps every
// lane of a row holds the same value.  The four rows are then combined from
// v_readlane in a fixed order ((r0 op r1) op (r2 op r3)), so the result is
// wave-uniform and does not depend on scheduling.  Every lane of the wave
// must be active (call these from converged control flow).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtr {

constexpr int kDppQuadXor1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = dpp_i32<CTRL>(__double2loint(v));
  const int hi = dpp_i32<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Sum over the wave in FP64, the same value in every lane.
__device__ __forceinline__ double wave_sum_f64(double v) {
  v = v + dpp_f64<kDppQuadXor1>(v);
  v = v + dpp_f64<kDppQuadXor2>(v);
  v = v + dpp_f64<kDppRowHalfMirror>(v);
  v = v + dpp_f64<kDppRowMirror>(v);
  const double r0 = readlane_f64(v, 0), r1 = readlane_f64(v, 16);
  const double r2 = readlane_f64(v, 32), r3 = readlane_f64(v, 48);
  return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
  int x = (int)v;
  x = x + dpp_i32<kDppQuadXor1>(x);
  x = x + dpp_i32<kDppQuadXor2>(x);
  x = x + dpp_i32<kDppRowHalfMirror>(x);
  x = x + dpp_i32<kDppRowMirror>(x);
  return (uint32_t)(__builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16)) +
         (uint32_t)(__builtin_amdgcn_readlane(x, 32) + __builtin_amdgcn_readlane(x, 48));
}

__device__ __forceinline__ int wave_min_i32(int v) {
  v = min(v, dpp_i32<kDppQuadXor1>(v));
  v = min(v, dpp_i32<kDppQuadXor2>(v));
  v = min(v, dpp_i32<kDppRowHalfMirror>(v));
  v = min(v, dpp_i32<kDppRowMirror>(v));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ __forceinline__ int wave_max_i32(int v) {
  v = max(v, dpp_i32<kDppQuadXor1>(v));
  v = max(v, dpp_i32<kDppQuadXor2>(v));
  v = max(v, dpp_i32<kDppRowHalfMirror>(v));
  v = max(v, dpp_i32<kDppRowMirror>(v));
  return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(dpp_i32<CTRL>(__float_as_int(v)));
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ float wave_min_f32(float v) {
  v = fminf(v, dpp_f32<kDppQuadXor1>(v));
  v = fminf(v, dpp_f32<kDppQuadXor2>(v));
  v = fminf(v, dpp_f32<kDppRowHalfMirror>(v));
  v = fminf(v, dpp_f32<kDppRowMirror>(v));
  return fminf(fminf(readlane_f32(v, 0), readlane_f32(v, 16)), fminf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

__device__ __forceinline__ float wave_max_f32(float v) {
  v = fmaxf(v, dpp_f32<kDppQuadXor1>(v));
  v = fmaxf(v, dpp_f32<kDppQuadXor2>(v));
  v = fmaxf(v, dpp_f32<kDppRowHalfMirror>(v));
  v = fmaxf(v, dpp_f32<kDppRowMirror>(v));
  return fmaxf(fmaxf(readlane_f32(v, 0), readlane_f32(v, 16)), fmaxf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)dpp_i32<CTRL>((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)dpp_i32<CTRL>((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// Every lane of each 16-lane row gets the row's largest value.
__device__ __forceinline__ uint64_t row_max_u64(uint64_t v) {
  uint64_t o = dpp_u64<kDppQuadXor1>(v);
  v = o > v ? o : v;
  o = dpp_u64<kDppQuadXor2>(v);
  v = o > v ? o : v;
  o = dpp_u64<kDppRowHalfMirror>(v);
  v = o > v ? o : v;
  o = dpp_u64<kDppRowMirror>(v);
  return o > v ? o : v;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  v = row_max_u64(v);
  const uint64_t a = readlane_u64(v, 0), b = readlane_u64(v, 16);
  const uint64_t c = readlane_u64(v, 32), d = readlane_u64(v, 48);
  const uint64_t ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}

}  // namespace rtr
