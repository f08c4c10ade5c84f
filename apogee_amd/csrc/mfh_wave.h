// Wave-wide reductions of 32-bit unsigned values through DPP (gfx950,
// wave64), every lane receiving the result: four DPP steps leave each row of
// 16 lanes holding its row's extremum (xor 1 and xor 2 by quad permutes, then
// the mirrored half row and the mirrored row), and the four row results are
// read into scalars with v_readlane.  No LDS permutes.  Nonnegative floats
// order like their bit patterns, so these serve float maxima as well.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfh {

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_umax(uint32_t v) {
  return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false));
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_umin(uint32_t v) {
  return min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ uint32_t row_readlanes_max(uint32_t v) {
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
  const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
  const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
  const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
  return max(max(a, b), max(c, d));
}

__device__ __forceinline__ uint32_t row_readlanes_min(uint32_t v) {
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
  const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
  const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
  const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
  return min(min(a, b), min(c, d));
}

__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
  v = dpp_umax<0xB1>(v);   // quad_perm(1, 0, 3, 2)
  v = dpp_umax<0x4E>(v);   // quad_perm(2, 3, 0, 1)
  v = dpp_umax<0x141>(v);  // row_half_mirror
  v = dpp_umax<0x140>(v);  // row_mirror
  return row_readlanes_max(v);
}

__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
  v = dpp_umin<0xB1>(v);
  v = dpp_umin<0x4E>(v);
  v = dpp_umin<0x141>(v);
  v = dpp_umin<0x140>(v);
  return row_readlanes_min(v);
}

}  // namespace mfh
