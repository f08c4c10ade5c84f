// Register-resident inverse DFT building blocks for gfx950 (wave64).
//
// Everything here is the inverse sign, w_R = exp(+2 pi i / R), because the
// matched filter evaluates z[k] = sum_f P[f] exp(+2 pi i f k / n).  Each
// routine maps its arguments in natural order to natural order; the 4x4 and
// 4x2 decompositions' output transposes are compile-time register renames.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfh {

struct cf {
  float x, y;
};

__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
// a * (+i)
__device__ __forceinline__ cf cmul_i(cf a) { return {-a.y, a.x}; }
// a * w
__device__ __forceinline__ cf cmul(cf a, cf w) {
  return {__builtin_fmaf(a.x, w.x, -(a.y * w.y)), __builtin_fmaf(a.x, w.y, a.y * w.x)};
}
// a * conj(b)
__device__ __forceinline__ cf cmul_conj(cf a, cf b) {
  return {__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -(a.x * b.y))};
}

constexpr float kSqrtHalf = 0.70710678118654752440f;
constexpr float kCos8 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float kSin8 = 0.38268343236508977173f;  // sin(pi/8)

// a * exp(+i pi/4)
__device__ __forceinline__ cf rot_w8_1(cf a) {
  return {(a.x - a.y) * kSqrtHalf, (a.x + a.y) * kSqrtHalf};
}
// a * exp(+3 i pi/4)
__device__ __forceinline__ cf rot_w8_3(cf a) {
  return {-(a.x + a.y) * kSqrtHalf, (a.x - a.y) * kSqrtHalf};
}
// a * exp(+i pi/8)
__device__ __forceinline__ cf rot_w16_1(cf a) { return cmul(a, cf{kCos8, kSin8}); }
// a * exp(+3 i pi/8)
__device__ __forceinline__ cf rot_w16_3(cf a) { return cmul(a, cf{kSin8, kCos8}); }
// a * exp(+9 i pi/8) = -(a * exp(+i pi/8))
__device__ __forceinline__ cf rot_w16_9(cf a) { return cmul(a, cf{-kCos8, -kSin8}); }

__device__ __forceinline__ void dft2(cf &a0, cf &a1) {
  const cf s = cadd(a0, a1), d = csub(a0, a1);
  a0 = s;
  a1 = d;
}

__device__ __forceinline__ void dft4(cf &a0, cf &a1, cf &a2, cf &a3) {
  const cf s0 = cadd(a0, a2), d0 = csub(a0, a2);
  const cf s1 = cadd(a1, a3), d1 = cmul_i(csub(a1, a3));
  a0 = cadd(s0, s1);
  a2 = csub(s0, s1);
  a1 = cadd(d0, d1);
  a3 = csub(d0, d1);
}

// 8-point inverse DFT as 2 (inner) x 4 (outer): x[4 n1 + n2].
__device__ __forceinline__ void dft8(cf *v) {
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) dft2(v[n2], v[4 + n2]);
  // v[4 k1 + n2] *= w8^(n2 k1); only k1 = 1 is non-trivial
  v[5] = rot_w8_1(v[5]);
  v[6] = cmul_i(v[6]);
  v[7] = rot_w8_3(v[7]);
  dft4(v[0], v[1], v[2], v[3]);
  dft4(v[4], v[5], v[6], v[7]);
  // X[k1 + 2 k2] sits at v[4 k1 + k2]
  cf o[8];
#pragma unroll
  for (int k1 = 0; k1 < 2; k1++)
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) o[k1 + 2 * k2] = v[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = o[i];
}

// 16-point inverse DFT as 4 x 4: x[4 n1 + n2].
__device__ __forceinline__ void dft16(cf *v) {
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
  // Y[n2][k1] at v[4 k1 + n2] *= w16^(n2 k1)
  v[5] = rot_w16_1(v[5]);
  v[6] = rot_w8_1(v[6]);
  v[7] = rot_w16_3(v[7]);
  v[9] = rot_w8_1(v[9]);
  v[10] = cmul_i(v[10]);
  v[11] = rot_w8_3(v[11]);
  v[13] = rot_w16_3(v[13]);
  v[14] = rot_w8_3(v[14]);
  v[15] = rot_w16_9(v[15]);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
  // X[k1 + 4 k2] sits at v[4 k1 + k2]
  cf o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; i++) v[i] = o[i];
}

// R-point inverse DFT on v[0..R) in natural order.
template <int R>
__device__ __forceinline__ void dftR(cf *v) {
  if constexpr (R == 2) {
    dft2(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    dft8(v);
  } else {
    static_assert(R == 16, "radix");
    dft16(v);
  }
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

}  // namespace mfh
