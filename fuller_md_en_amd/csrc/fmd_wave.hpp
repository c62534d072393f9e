// fmd_wave.hpp -- wavefront-level helpers shared by the gfx950 kernels:
// DPP lane permutations and readlanes of doubles, and a fixed-order
// wavefront sum that needs no LDS traffic.
#pragma once

#include <hip/hip_runtime.h>

namespace fmd {

// One DPP lane permutation of a double (both halves), all rows and banks.
template <int CTRL>
__device__ __forceinline__ double dpp64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum of a double over the 64 lanes of a wavefront, the same value in every
// lane and a fixed order: DPP pairings inside each row of 16 (lane ^ 1,
// lane ^ 2, 8-lane and 16-lane mirrors; each pairing adds two equal-size
// partial sums, so both partners get bitwise the same value), then the four
// row sums.  Must be called with the whole wavefront active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp64<0x141>(v);  // row_half_mirror
  v += dpp64<0x140>(v);  // row_mirror
  return (readlane64(v, 0) + readlane64(v, 16)) + (readlane64(v, 32) + readlane64(v, 48));
}

}  // namespace fmd
