// MFMA fragment conventions used by the coarse screen (gfx950, wave64).
//
// For v_mfma_f32_16x16x32_{f16,bf16} (D = A(16x32) * B(32x16) + C):
//   A: lane l holds row (l % 16), k = 8*(l / 16) + i, i = 0..7
//   B: lane l holds col (l % 16), k = 8*(l / 16) + i, i = 0..7
//   C/D: lane l holds col (l % 16), rows 4*(l / 16) + i, i = 0..3
// For v_mfma_f32_16x16x16{f16,bf16_1k}: the same with k = 4*(l / 16) + i,
// i = 0..3.  mfh_selftest_mfma() runs each instruction on given matrices so
// the GPU tests pin these conventions against numpy.
#pragma once
#include <hip/hip_runtime.h>

namespace mfh {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4v mfma16x16x32(half8 a, half8 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4v mfma16x16x32(bf16x8 a, bf16x8 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4v mfma16x16x16(half4 a, half4 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4v mfma16x16x16(short4v a, short4v b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

}  // namespace mfh
