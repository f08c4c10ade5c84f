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
constexpr float kTan8 = 0.41421356237309504880f;  // tan(pi/8) = sqrt(2) - 1

// a * exp(+i pi/4)
__device__ __forceinline__ cf rot_w8_1(cf a) {
  return {(a.x - a.y) * kSqrtHalf, (a.x + a.y) * kSqrtHalf};
}
// a * exp(+3 i pi/4)
__device__ __forceinline__ cf rot_w8_3(cf a) {
  return {-(a.x + a.y) * kSqrtHalf, (a.x - a.y) * kSqrtHalf};
}

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

// The second-layer 4-point DFT for k1 = 2, inputs owing 1, w8, i, w8^3:
// with u = b1 + i b3 and v = b1 - i b3, a1 + a3 = w8 u and
// i (a1 - a3) = w8^3 v, and both rotations fold into the output FMAs.
__device__ __forceinline__ void dft4_k2(cf &a0, cf &b1, cf &b2, cf &b3) {
  const cf s0 = {a0.x - b2.y, a0.y + b2.x};  // a0 + i b2
  const cf d0 = {a0.x + b2.y, a0.y - b2.x};  // a0 - i b2
  const cf u = {b1.x - b3.y, b1.y + b3.x};
  const cf v = {b1.x + b3.y, b1.y - b3.x};
  const float pu = u.x - u.y, qu = u.x + u.y;  // w8 u = (pu k, qu k)
  const float pv = v.x - v.y, qv = v.x + v.y;  // w8^3 v = (-qv k, pv k)
  a0 = cf{__builtin_fmaf(pu, kSqrtHalf, s0.x), __builtin_fmaf(qu, kSqrtHalf, s0.y)};
  b2 = cf{__builtin_fmaf(-pu, kSqrtHalf, s0.x), __builtin_fmaf(-qu, kSqrtHalf, s0.y)};
  b1 = cf{__builtin_fmaf(-qv, kSqrtHalf, d0.x), __builtin_fmaf(pv, kSqrtHalf, d0.y)};
  b3 = cf{__builtin_fmaf(qv, kSqrtHalf, d0.x), __builtin_fmaf(-pv, kSqrtHalf, d0.y)};
}

// The second-layer 4-point DFT for k1 = 1 (SIGN = +1) or k1 = 3 (SIGN =
// -1), whose inputs owe 1, w16^k1, w8^(2 - SIGN) and w16^(3 k1).  Every odd
// power of w16 is cos(pi/8) times (1 + i tan(pi/8)) or (tan(pi/8) + i), up to
// sign: w16 = c (1 + i t), w16^3 = c (t + i), w16^9 = -c (1 + i t) with
// c = cos(pi/8), t = tan(pi/8).  So b1 and b3 are rotated by the unscaled
// factors (two FMAs each, against four instructions for a full complex
// multiply) and the common c folds into the output FMAs, as the w8 factor
// of b2 does.
template <int SIGN>
__device__ __forceinline__ void dft4_w16(cf &a0, cf &b1, cf &b2, cf &b3) {
  const float p = SIGN > 0 ? b2.x - b2.y : -(b2.x + b2.y);
  const float q = SIGN > 0 ? b2.x + b2.y : b2.x - b2.y;
  const cf s0 = {__builtin_fmaf(p, kSqrtHalf, a0.x), __builtin_fmaf(q, kSqrtHalf, a0.y)};
  const cf d0 = {__builtin_fmaf(-p, kSqrtHalf, a0.x), __builtin_fmaf(-q, kSqrtHalf, a0.y)};
  // (1 + i t) x = (x.x - t x.y, x.y + t x.x); (t + i) x = (t x.x - x.y, t x.y + x.x)
  const cf e = SIGN > 0 ? b1 : b3;  // the input rotated by (1 + i t)
  const cf f = SIGN > 0 ? b3 : b1;  // the input rotated by (t + i)
  const cf pe = {__builtin_fmaf(-kTan8, e.y, e.x), __builtin_fmaf(kTan8, e.x, e.y)};
  const cf pf = {__builtin_fmaf(kTan8, f.x, -f.y), __builtin_fmaf(kTan8, f.y, f.x)};
  // a1 + a3 = c su and i (a1 - a3) = c i dv, where for k1 = 1 a1 = c pe and
  // a3 = c pf, and for k1 = 3 a1 = c pf and a3 = -c pe
  const cf su = SIGN > 0 ? cadd(pe, pf) : csub(pf, pe);
  const cf dv = SIGN > 0 ? csub(pe, pf) : cadd(pf, pe);
  a0 = cf{__builtin_fmaf(su.x, kCos8, s0.x), __builtin_fmaf(su.y, kCos8, s0.y)};
  b2 = cf{__builtin_fmaf(-su.x, kCos8, s0.x), __builtin_fmaf(-su.y, kCos8, s0.y)};
  b1 = cf{__builtin_fmaf(-dv.y, kCos8, d0.x), __builtin_fmaf(dv.x, kCos8, d0.y)};
  b3 = cf{__builtin_fmaf(dv.y, kCos8, d0.x), __builtin_fmaf(-dv.x, kCos8, d0.y)};
}

// 16-point inverse DFT as 4 x 4: x[4 n1 + n2].
__device__ __forceinline__ void dft16(cf *v) {
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
  // Y[n2][k1] at v[4 k1 + n2] owes w16^(n2 k1); every such factor is folded
  // into the second layer's arithmetic
  dft4(v[0], v[1], v[2], v[3]);
  dft4_w16<+1>(v[4], v[5], v[6], v[7]);
  dft4_k2(v[8], v[9], v[10], v[11]);
  dft4_w16<-1>(v[12], v[13], v[14], v[15]);
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
