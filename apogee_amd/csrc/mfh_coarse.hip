// Coarse band screen on the matrix cores (gfx950 MFMA), band m = 512.
//
// The coarse statistic is max_j |z_c[j]| over the coarse window, with
//   z_c[j] = sum_{f<m} D[f] conj(ct[f]) exp(+2 pi i f j / m)
// (the band-point flat filter of src/hmf.c:139-143 and :246-250 in the
// reference checkout).  Here the m-point DFT is factored as 32 x 16
// (f = 16 a + b, j = c + 32 e):
//   Y[b, c]  = sum_a P[16 a + b] w32^(a c)            step 1: 16x16x32 MFMA
//   Y'[b, c] = Y[b, c] * w512^(b c)                   twiddle, FP32
//   z[c+32e] = sum_b w16^(e b) Y'[b, c]               step 3: 16x16x16 MFMA
// so a pair costs 0.2 MFLOP on MFMA instead of 2.1 for the dense DFT.  The
// step-1 accumulator layout (lane l: column c = l % 16, rows b = 4 (l/16) + i)
// IS the B-operand layout of the 16x16x16 instruction, so step 3 takes its
// input straight from registers.
//
// Inputs are scaled to unit maximum per template and per data row before the
// conversion to fp16/bf16 (FP32 accumulation throughout); the result is
// rescaled.  The statistic's relative error is bounded by a measured factor
// (mfh_hier_set_coarse_precision) by which the gate is widened, so the
// screen keeps a superset of the FP32 screen's pairs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <vector>

#include "mfh_internal.h"
#include "mfh_mfma.h"
#include "mfhip.h"

namespace mfh {

namespace {

// ---------------------------------------------------------------- self-test

__global__ void mfma_selftest_kernel(int variant, const float *A, const float *B, float *C) {
  const int l = threadIdx.x;
  float4v acc = {0.f, 0.f, 0.f, 0.f};
  if (variant == 0 || variant == 1) {  // 16x16x32: A 16x32, B 32x16
    half8 a, b;
    bf16x8 ab, bb;
    for (int i = 0; i < 8; i++) {
      const float av = A[(l % 16) * 32 + 8 * (l / 16) + i];
      const float bv = B[(8 * (l / 16) + i) * 16 + (l % 16)];
      a[i] = (_Float16)av;
      b[i] = (_Float16)bv;
      ab[i] = (__bf16)av;
      bb[i] = (__bf16)bv;
    }
    acc = variant == 0 ? mfma16x16x32(a, b, acc) : mfma16x16x32(ab, bb, acc);
  } else {  // 16x16x16: A 16x16, B 16x16
    half4 a, b;
    short4v as, bs;
    for (int i = 0; i < 4; i++) {
      const float av = A[(l % 16) * 16 + 4 * (l / 16) + i];
      const float bv = B[(4 * (l / 16) + i) * 16 + (l % 16)];
      a[i] = (_Float16)av;
      b[i] = (_Float16)bv;
      __bf16 x = (__bf16)av, y = (__bf16)bv;
      as[i] = __builtin_bit_cast(short, x);
      bs[i] = __builtin_bit_cast(short, y);
    }
    acc = variant == 2 ? mfma16x16x16(a, b, acc) : mfma16x16x16(as, bs, acc);
  }
  for (int i = 0; i < 4; i++) C[(4 * (l / 16) + i) * 16 + (l % 16)] = acc[i];
}

// ------------------------------------------------------------ coarse kernel

template <bool BF16>
struct Frag;
template <>
struct Frag<false> {
  using v8 = half8;
  using v4 = half4;
  static __device__ __forceinline__ _Float16 cvt8(float x) { return (_Float16)x; }
  static __device__ __forceinline__ _Float16 cvt4(float x) { return (_Float16)x; }
};
template <>
struct Frag<true> {
  using v8 = bf16x8;
  using v4 = short4v;
  static __device__ __forceinline__ __bf16 cvt8(float x) { return (__bf16)x; }
  static __device__ __forceinline__ short cvt4(float x) {
    __bf16 b = (__bf16)x;
    return __builtin_bit_cast(short, b);
  }
};

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// One wave per template, walking `rows_per_wave` data rows; four waves per
// workgroup.  Writes the coarse maximum |z_c|^2 over lags [cs, ce) per pair,
// in units of the unscaled inputs.
template <bool BF16>
__global__ void __launch_bounds__(256) coarse_mfma_kernel(const cf *data, long long data_stride, const cf *ctmpl,
                                                          int nd, int nt, int rows_per_wave, int cs, int ce,
                                                          const cf *w32, const cf *w512, const cf *w16,
                                                          float *cmax2, int *zero) {
  using F = Frag<BF16>;
  using v8 = typename F::v8;
  using v4 = typename F::v4;
  constexpr int M = 512;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  // the refine-list counter of the gate that follows starts at zero; no
  // kernel appends to it before this one has finished
  if (zero && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *zero = 0;
  if (t >= nt) return;  // whole waves exit; no barriers in this kernel
  const int q = lane >> 4, r = lane & 15;

  // A-fragment positions of step 1: row b = r, k = a = 8q + i -> f = 16 a + b
  cf ct[8];
  float smax = 0.f;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    ct[i] = ctmpl[(long long)t * M + 16 * (8 * q + i) + r];
    smax = fmaxf(smax, fmaxf(fabsf(ct[i].x), fabsf(ct[i].y)));
  }
  smax = wave_max_f32(smax);
  const float st = smax > 0.f ? 1.0f / smax : 0.f;
#pragma unroll
  for (int i = 0; i < 8; i++) ct[i] = cf{ct[i].x * st, ct[i].y * st};

  // step-1 B operand: F32[a][c], a = 8q + i, c = r + 16 h (two column tiles h)
  v8 fr[2], fi[2], fn[2];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const cf w = w32[((8 * q + i) * (r + 16 * h)) & 31];
      fr[h][i] = F::cvt8(w.x);
      fi[h][i] = F::cvt8(w.y);
      fn[h][i] = F::cvt8(-w.y);
    }
  // twiddles w512^(b c) at the accumulator positions b = 4q + i, c = r + 16 h
  cf tw[2][4];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int i = 0; i < 4; i++) tw[h][i] = w512[((4 * q + i) * (r + 16 * h)) & 511];
  // step-3 A operand: F16[e][b], e = r, b = 4q + i
  v4 gr, gi, gn;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const cf w = w16[(r * (4 * q + i)) & 15];
    gr[i] = F::cvt4(w.x);
    gi[i] = F::cvt4(w.y);
    gn[i] = F::cvt4(-w.y);
  }

  const bool full = cs == 0 && ce == M;  // whole coarse window: no per-lag test
  const int d0 = blockIdx.y * rows_per_wave;
  const int d1 = min(nd, d0 + rows_per_wave);
  // the next row's samples are loaded while this row is transformed
  cf xn[8];
#pragma unroll
  for (int i = 0; i < 8; i++) xn[i] = data[(long long)d0 * data_stride + 16 * (8 * q + i) + r];
  for (int d = d0; d < d1; d++) {
    cf x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = xn[i];
    if (d + 1 < d1) {
      const cf *Dn = data + (long long)(d + 1) * data_stride;
#pragma unroll
      for (int i = 0; i < 8; i++) xn[i] = Dn[16 * (8 * q + i) + r];
    }
    cf p[8];
    float dmax = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      dmax = fmaxf(dmax, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
      p[i] = cmul_conj(x[i], ct[i]);
    }
    dmax = wave_max_f32(dmax);
    const float sd = dmax > 0.f ? 0.5f / dmax : 0.f;  // |P| <= 1 after scaling
    v8 xr, xi;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      xr[i] = F::cvt8(p[i].x * sd);
      xi[i] = F::cvt8(p[i].y * sd);
    }
    float best = 0.f;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float4v yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      yr = mfma16x16x32(xr, fr[h], yr);
      yr = mfma16x16x32(xi, fn[h], yr);
      yi = mfma16x16x32(xr, fi[h], yi);
      yi = mfma16x16x32(xi, fr[h], yi);
      v4 br, bi;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const cf y = cmul(cf{yr[i], yi[i]}, tw[h][i]);
        br[i] = F::cvt4(y.x);
        bi[i] = F::cvt4(y.y);
      }
      float4v zr = {0.f, 0.f, 0.f, 0.f}, zi = {0.f, 0.f, 0.f, 0.f};
      zr = mfma16x16x16(gr, br, zr);
      zr = mfma16x16x16(gn, bi, zr);
      zi = mfma16x16x16(gr, bi, zi);
      zi = mfma16x16x16(gi, br, zi);
      // z at j = c + 32 e with c = r + 16 h, e = 4q + i
      if (full) {
#pragma unroll
        for (int i = 0; i < 4; i++) best = fmaxf(best, __builtin_fmaf(zr[i], zr[i], zi[i] * zi[i]));
      } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int j = r + 16 * h + 32 * (4 * q + i);
          const float m2 = __builtin_fmaf(zr[i], zr[i], zi[i] * zi[i]);
          if (j >= cs && j < ce) best = fmaxf(best, m2);
        }
      }
    }
    best = wave_max_f32(best);
    if (lane == 0) {
      const float inv = (st > 0.f && sd > 0.f) ? 1.0f / (st * sd) : 0.f;
      cmax2[(long long)d * nt + t] = best * inv * inv;
    }
  }
}

// exp(+2 pi i k / L), k < L, FP64-generated
__global__ void unit_roots_kernel(cf *out, int L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < L) {
    double s, c;
    sincospi(2.0 * (double)k / (double)L, &s, &c);
    out[k] = cf{(float)c, (float)s};
  }
}

}  // namespace

// Roots of unity of orders 32, 512, 16 for the factored DFT, built once.
static const cf *coarse_roots(Ctx *c) {
  cf *r = (cf *)c->scratch("coarse_roots", (32 + 512 + 16) * sizeof(cf));
  if (!r) return nullptr;
  if (!c->coarse_roots_ready) {
    hipLaunchKernelGGL(unit_roots_kernel, dim3(1), dim3(64), 0, c->stream, r, 32);
    hipLaunchKernelGGL(unit_roots_kernel, dim3(2), dim3(256), 0, c->stream, r + 32, 512);
    hipLaunchKernelGGL(unit_roots_kernel, dim3(1), dim3(64), 0, c->stream, r + 544, 16);
    if (!hip_ok(hipGetLastError(), "unit_roots_kernel")) return nullptr;
    c->coarse_roots_ready = true;
  }
  return r;
}

int launch_coarse_mfma(Ctx *c, const cf *data, long long data_stride, const cf *ctmpl, int nd, int nt, int cs, int ce,
                       bool bf16, float *cmax2, int *zero) {
  if (c->coarse_band != 512) return fail("the MFMA coarse screen is built for band 512");
  const cf *roots = coarse_roots(c);
  if (!roots) return -1;
  const int rows = std::max(1, c->coarse_rows);
  dim3 grid((unsigned)((nt + 3) / 4), (unsigned)((nd + rows - 1) / rows));
  if (grid.y > 65535) return fail("too many data rows per launch (%d)", nd);
  Ctx::ProfScope ps(c, "coarse");
  if (bf16)
    hipLaunchKernelGGL(coarse_mfma_kernel<true>, grid, dim3(256), 0, c->stream, data, data_stride, ctmpl, nd, nt, rows,
                       cs, ce, roots, roots + 32, roots + 544, cmax2, zero);
  else
    hipLaunchKernelGGL(coarse_mfma_kernel<false>, grid, dim3(256), 0, c->stream, data, data_stride, ctmpl, nd, nt,
                       rows, cs, ce, roots, roots + 32, roots + 544, cmax2, zero);
  return hip_ok(hipGetLastError(), "coarse_mfma_kernel") ? 0 : -1;
}

}  // namespace mfh

using mfh::Ctx;
using mfh::fail;
using mfh::hip_ok;

extern "C" int mfh_selftest_mfma(int device, int variant, const float *A, const float *B, float *C) {
  if (variant < 0 || variant > 3) return fail("variant must be 0..3");
  if (!A || !B || !C) return fail("null argument");
  if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return -1;
  const size_t ka = variant < 2 ? 32 : 16;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  int rc = -1;
  if (hip_ok(hipMalloc(&dA, 16 * ka * sizeof(float)), "hipMalloc") &&
      hip_ok(hipMalloc(&dB, ka * 16 * sizeof(float)), "hipMalloc") &&
      hip_ok(hipMalloc(&dC, 16 * 16 * sizeof(float)), "hipMalloc") &&
      hip_ok(hipMemcpy(dA, A, 16 * ka * sizeof(float), hipMemcpyHostToDevice), "copy A") &&
      hip_ok(hipMemcpy(dB, B, ka * 16 * sizeof(float), hipMemcpyHostToDevice), "copy B")) {
    hipLaunchKernelGGL(mfh::mfma_selftest_kernel, dim3(1), dim3(64), 0, 0, variant, dA, dB, dC);
    if (hip_ok(hipGetLastError(), "selftest launch") &&
        hip_ok(hipMemcpy(C, dC, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost), "copy C"))
      rc = 0;
  }
  hipFree(dA);
  hipFree(dB);
  hipFree(dC);
  return rc;
}
