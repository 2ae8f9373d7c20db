// Cross-lane sums that stay in the vector registers: DPP row operations
// inside a row of 16 lanes and lane reads across the wave's four rows, in
// place of shuffles through the LDS crossbar (ds_bpermute / ds_swizzle).
#pragma once

#include <hip/hip_runtime.h>

namespace ttf {

// DPP controls of the CDNA row operations
constexpr int kDppXor1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;         // quad_perm [2, 3, 0, 1]
constexpr int kDppRowMirror = 0x140;   // lane i <- lane 15 - i within the row
constexpr int kDppRowHalfMirror = 0x141;   // lane i <- lane 7 - i within each half row

// x of the lane that CTRL selects within the caller's row.  Every control
// above reads a lane of the same row and all rows and banks are enabled, so
// no lane falls back to the bound-control zero; the move carries no old
// value, which spares a register copy per 32-bit half before each move.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 16 lanes of the caller's DPP row (lanes 16k .. 16k + 15, all
// active).  Every step adds two partial sums that mirror each other, so all
// 16 lanes end with bit-identical sums.
__device__ __forceinline__ double row_sum16(double x) {
  x += dpp_f64<kDppXor1>(x);
  x += dpp_f64<kDppXor2>(x);
  x += dpp_f64<kDppRowHalfMirror>(x);
  x += dpp_f64<kDppRowMirror>(x);
  return x;
}

// Sum over the 4 lanes of the caller's quad (lanes 4k .. 4k + 3, all
// active); all four end with the same value.
__device__ __forceinline__ double quad_sum(double x) {
  x += dpp_f64<kDppXor1>(x);
  x += dpp_f64<kDppXor2>(x);
  return x;
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, lane);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the whole wave (all 64 lanes active); the result is uniform.
__device__ __forceinline__ double wave_sum64(double x) {
  x = row_sum16(x);
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}

}  // namespace ttf
