// Wave-wide (64-lane) reductions on DPP row permutations for gfx950 kernels: xor 1, xor 2,
// half-row and row mirrors reduce each row of 16 lanes, and four readlanes combine the
// rows, so every lane gets the result.  Every lane of the wave must be active.
#pragma once

#include <hip/hip_runtime.h>

namespace sslab {
namespace wave {

template <int kCtrl>
__device__ __forceinline__ int dpp_i32(int x) {
  // every lane has a valid source lane for the permutations used here
  return __builtin_amdgcn_mov_dpp(x, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ double dpp_e(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_i32<kCtrl>((int)b);
  const int hi = dpp_i32<kCtrl>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int kCtrl>
__device__ __forceinline__ float dpp_e(float x) {
  return __int_as_float(dpp_i32<kCtrl>(__float_as_int(x)));
}
__device__ __forceinline__ double readlane_e(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane_e(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
constexpr int kDppXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;

template <typename E>
__device__ __forceinline__ E max_e(E a, E b) { return a > b ? a : b; }

// Wave maximum of a value that is never NaN.
template <typename E>
__device__ __forceinline__ E wave_max(E v) {
  v = max_e(v, dpp_e<kDppXor1>(v));
  v = max_e(v, dpp_e<kDppXor2>(v));
  v = max_e(v, dpp_e<kDppHalfMirror>(v));
  v = max_e(v, dpp_e<kDppMirror>(v));
  const E a = max_e(readlane_e(v, 0), readlane_e(v, 16));
  const E b = max_e(readlane_e(v, 32), readlane_e(v, 48));
  return max_e(a, b);
}
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, dpp_i32<kDppXor1>(v));
  v = min(v, dpp_i32<kDppXor2>(v));
  v = min(v, dpp_i32<kDppHalfMirror>(v));
  v = min(v, dpp_i32<kDppMirror>(v));
  const int a = min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16));
  const int b = min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48));
  return min(a, b);
}
// Wave maximum of unsigned 32-bit keys.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, (unsigned)dpp_i32<kDppXor1>((int)v));
  v = max(v, (unsigned)dpp_i32<kDppXor2>((int)v));
  v = max(v, (unsigned)dpp_i32<kDppHalfMirror>((int)v));
  v = max(v, (unsigned)dpp_i32<kDppMirror>((int)v));
  const unsigned a = max((unsigned)__builtin_amdgcn_readlane((int)v, 0),
                         (unsigned)__builtin_amdgcn_readlane((int)v, 16));
  const unsigned b = max((unsigned)__builtin_amdgcn_readlane((int)v, 32),
                         (unsigned)__builtin_amdgcn_readlane((int)v, 48));
  return max(a, b);
}
// Sum over each row of 16 lanes (lanes 16 q .. 16 q + 15), the same in the row's lanes.
template <typename E>
__device__ __forceinline__ E row16_sum(E v) {
  v += dpp_e<kDppXor1>(v);
  v += dpp_e<kDppXor2>(v);
  v += dpp_e<kDppHalfMirror>(v);
  v += dpp_e<kDppMirror>(v);
  return v;
}
// Wave sum (the same in every lane).
template <typename E>
__device__ __forceinline__ E wave_sum(E v) {
  v += dpp_e<kDppXor1>(v);
  v += dpp_e<kDppXor2>(v);
  v += dpp_e<kDppHalfMirror>(v);
  v += dpp_e<kDppMirror>(v);
  return (readlane_e(v, 0) + readlane_e(v, 16)) + (readlane_e(v, 32) + readlane_e(v, 48));
}

}  // namespace wave
}  // namespace sslab
