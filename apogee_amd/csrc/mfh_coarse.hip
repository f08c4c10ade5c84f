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
// Inputs are scaled by exact powers of two to below unit maximum per
// template and per data row before the conversion to fp16/bf16 (FP32
// accumulation throughout); the result is rescaled exactly.  The statistic's relative error is bounded by a measured factor
// (mfh_hier_set_coarse_precision) by which the gate is widened, so the
// screen keeps a superset of the FP32 screen's pairs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <vector>

#include "mfh_flat.h"
#include "mfh_internal.h"
#include "mfh_mfma.h"
#include "mfh_wave.h"
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

// The wave maximum of a nonnegative float, every lane receiving it.
__device__ __forceinline__ float wave_max_nonneg(float v) { return __uint_as_float(wave_umax(__float_as_uint(v))); }

// Exponent s with m * 2^-s < 1 and m * 2^-s >= 1/2 for a finite m > 0 (the
// biased exponent of m minus 126); m = 0 gives -126.
__device__ __forceinline__ int unit_exponent(float m) { return (int)(__float_as_uint(m) >> 23) - 126; }

constexpr int kBand = 512;

// Coarse templates in the step-1 A-fragment order, scaled by an exact power
// of two below unit maximum: lane l = 16 q + r holds the eight samples
// f = 128 q + 16 i + r, i < 8, contiguously at ctm[(t * 64 + l) * 8 + i],
// equal to ct[t][f] * 2^-s_t with s_t = sexp[t] = unit_exponent(max over the
// template of |Re|, |Im|).  One wave per template; built when the coarse
// templates are.
__global__ void __launch_bounds__(64) coarse_mfma_prep_kernel(const cf *ct, cf *ctm, int *sexp) {
  const int t = blockIdx.x, lane = threadIdx.x, q = lane >> 4, r = lane & 15;
  cf v[8];
  float m = 0.f;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    v[i] = ct[(long long)t * kBand + 128 * q + 16 * i + r];
    m = fmaxf(m, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
  }
  const int s = unit_exponent(wave_max_nonneg(m));
  cf *out = ctm + ((long long)t * 64 + lane) * 8;
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = cf{ldexpf(v[i].x, -s), ldexpf(v[i].y, -s)};
  if (lane == 0) sexp[t] = s;
}

// The per-lane constants of the factored 512-point DFT.
template <bool BF16>
struct CoarseDft {
  using F = Frag<BF16>;
  typename F::v8 fr[2], fi[2], fn[2];  // step-1 B operand F32[a][c], a = 8q + i, c = r + 16 h
  cf tw[2][4];                          // w512^(b c) at the accumulator positions b = 4q + i
  typename F::v4 gr, gi, gn;            // step-3 A operand F16[e][b], e = r, b = 4q + i

  // Per-lane table of all the constants above (kLaneWords 32-bit words per
  // lane, lane-major), built once per context by coarse_consts_kernel.
  static constexpr int kLaneWords = 48;
  __device__ __forceinline__ void load_table(const uint32_t *tab, int lane) {
    const uint4 *p = reinterpret_cast<const uint4 *>(tab + lane * kLaneWords);
    uint4 w[12];
#pragma unroll
    for (int i = 0; i < 12; i++) w[i] = p[i];
    __builtin_memcpy(&fr[0], &w[0], 16);
    __builtin_memcpy(&fr[1], &w[1], 16);
    __builtin_memcpy(&fi[0], &w[2], 16);
    __builtin_memcpy(&fi[1], &w[3], 16);
    __builtin_memcpy(&fn[0], &w[4], 16);
    __builtin_memcpy(&fn[1], &w[5], 16);
    __builtin_memcpy(&tw[0][0], &w[6], 64);
    __builtin_memcpy(&gr, &w[10], 8);
    __builtin_memcpy(&gi, reinterpret_cast<const char *>(&w[10]) + 8, 8);
    __builtin_memcpy(&gn, &w[11], 8);
  }
  __device__ __forceinline__ void store_table(uint32_t *tab, int lane) const {
    uint4 w[12] = {};
    __builtin_memcpy(&w[0], &fr[0], 16);
    __builtin_memcpy(&w[1], &fr[1], 16);
    __builtin_memcpy(&w[2], &fi[0], 16);
    __builtin_memcpy(&w[3], &fi[1], 16);
    __builtin_memcpy(&w[4], &fn[0], 16);
    __builtin_memcpy(&w[5], &fn[1], 16);
    __builtin_memcpy(&w[6], &tw[0][0], 64);
    __builtin_memcpy(&w[10], &gr, 8);
    __builtin_memcpy(reinterpret_cast<char *>(&w[10]) + 8, &gi, 8);
    __builtin_memcpy(&w[11], &gn, 8);
    uint4 *p = reinterpret_cast<uint4 *>(tab + lane * kLaneWords);
#pragma unroll
    for (int i = 0; i < 12; i++) p[i] = w[i];
  }

  __device__ __forceinline__ void load(const cf *w32, const cf *w512, const cf *w16, int q, int r) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const cf w = w32[((8 * q + i) * (r + 16 * h)) & 31];
        fr[h][i] = F::cvt8(w.x);
        fi[h][i] = F::cvt8(w.y);
        fn[h][i] = F::cvt8(-w.y);
      }
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int i = 0; i < 4; i++) tw[h][i] = w512[((4 * q + i) * (r + 16 * h)) & 511];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const cf w = w16[(r * (4 * q + i)) & 15];
      gr[i] = F::cvt4(w.x);
      gi[i] = F::cvt4(w.y);
      gn[i] = F::cvt4(-w.y);
    }
  }

  // The constants straight from a workgroup's LDS copy of the table, stored
  // word4-major ([w][lane], so a wave's 16-byte reads are conflict-free),
  // instead of held in registers: steps' constants are read where used.
  __device__ __forceinline__ static void lds_words(uint4 *w, const uint4 *tabT, int lane, int first, int count) {
#pragma unroll
    for (int i = 0; i < 12; i++)
      if (i >= first && i < first + count) w[i] = tabT[i * 64 + lane];
  }
  template <bool FULL>
  __device__ __forceinline__ uint32_t lane_max_lds(const uint4 *tabT, int lane, const cf *x, const cf *ct, int cs,
                                                   int ce, int q, int r) {
    // an opaque lane index per call keeps the reads inside the template loop
    int lo = lane;
    asm volatile("" : "+v"(lo));
    uint4 w[12];
    lds_words(w, tabT, lo, 10, 2);
    __builtin_memcpy(&gr, &w[10], 8);
    __builtin_memcpy(&gi, reinterpret_cast<const char *>(&w[10]) + 8, 8);
    __builtin_memcpy(&gn, &w[11], 8);
    typename F::v8 xr, xi;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const cf p = cmul_conj(x[i], ct[i]);
      xr[i] = F::cvt8(p.x);
      xi[i] = F::cvt8(p.y);
    }
    uint32_t best = 0u;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      w[h] = tabT[h * 64 + lo];
      w[2 + h] = tabT[(2 + h) * 64 + lo];
      w[4 + h] = tabT[(4 + h) * 64 + lo];
      w[6 + 2 * h] = tabT[(6 + 2 * h) * 64 + lo];
      w[7 + 2 * h] = tabT[(7 + 2 * h) * 64 + lo];
      __builtin_memcpy(&fr[h], &w[h], 16);
      __builtin_memcpy(&fi[h], &w[2 + h], 16);
      __builtin_memcpy(&fn[h], &w[4 + h], 16);
      __builtin_memcpy(&tw[h][0], &w[6 + 2 * h], 32);
      float4v yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      yr = mfma16x16x32(xr, fr[h], yr);
      yr = mfma16x16x32(xi, fn[h], yr);
      yi = mfma16x16x32(xr, fi[h], yi);
      yi = mfma16x16x32(xi, fr[h], yi);
      typename F::v4 br, bi;
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
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t m2 = __float_as_uint(__builtin_fmaf(zr[i], zr[i], zi[i] * zi[i]));
        if (FULL) {
          best = max(best, m2);
        } else {
          const int j = r + 16 * h + 32 * (4 * q + i);
          if (j >= cs && j < ce) best = max(best, m2);
        }
      }
    }
    return best;
  }

  // This lane's maximum of |z[j]|^2 over its 8 lags j = r + 16 h + 32 (4q + i)
  // that lie in [cs, ce) (all of them when FULL), for the scaled spectra x
  // (data) and ct (template) at the A-fragment positions.  |z|^2 >= 0, so
  // the maximum is taken on the bit patterns (no NaN canonicalisation).
  template <bool FULL>
  __device__ __forceinline__ uint32_t lane_max(const cf *x, const cf *ct, int cs, int ce, int q, int r) const {
    typename F::v8 xr, xi;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const cf p = cmul_conj(x[i], ct[i]);
      xr[i] = F::cvt8(p.x);
      xi[i] = F::cvt8(p.y);
    }
    uint32_t best = 0u;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float4v yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
      yr = mfma16x16x32(xr, fr[h], yr);
      yr = mfma16x16x32(xi, fn[h], yr);
      yi = mfma16x16x32(xr, fi[h], yi);
      yi = mfma16x16x32(xi, fr[h], yi);
      typename F::v4 br, bi;
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
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t m2 = __float_as_uint(__builtin_fmaf(zr[i], zr[i], zi[i] * zi[i]));
        if (FULL) {
          best = max(best, m2);
        } else {
          const int j = r + 16 * h + 32 * (4 * q + i);
          if (j >= cs && j < ce) best = max(best, m2);
        }
      }
    }
    return best;
  }
};

__device__ __forceinline__ void load_frag8(cf *ct, const cf *ctm, int t, int lane) {
  const float4 *src = (const float4 *)(ctm + ((long long)t * 64 + lane) * 8);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float4 v = src[i];
    ct[2 * i] = cf{v.x, v.y};
    ct[2 * i + 1] = cf{v.z, v.w};
  }
}

// Each wave holds R data rows, each scaled by an exact power of two below
// 1/2 in every component, and walks `tpw` templates, fetching the next
// template's fragments while it transforms the current one; the four waves
// of a workgroup take 4 R consecutive data rows and the same templates.
// With |Re|, |Im| < 1/2 for the data and < 1 for the template, |P| < 1,
// |Y| < 32 and |z| < 512: no fp16 overflow.  Writes the coarse maximum
// |z_c|^2 over lags [cs, ce) per pair, in units of the unscaled inputs (an
// exact power-of-two rescale).
//
// LDSC: the DFT constants are copied once per workgroup into LDS and read
// where used instead of held in registers (46 fewer VGPRs per lane, so more
// waves fit per SIMD).
//
// This is the screen of one workgroup, shared by the two kernels below: on
// return fired[k] holds, for data row d0 + k, bit j set iff template t0 + j
// passed the gate.  LDSC's LDS: ctab (12 x 64 words4, the DFT constants),
// tsexp (the workgroup's templates' scale exponents, tpw <= 64) and tfrag
// (the workgroup's templates' fragments, [template][quad][lane] in 16-byte
// units, conflict-free reads), staged once for the four waves.
template <bool BF16, int R, bool FULL, bool LDSC>
__device__ __forceinline__ void coarse_screen(const cf *data, long long data_stride, const cf *ctm, const int *sexp,
                                              int nd, int nt, int tpw, int cs, int ce, const uint32_t *consts,
                                              float *cmax2, const CoarseGate g, uint4 *ctab, int *tsexp,
                                              uint4 *tfrag, int d0, int t0, int t1, uint64_t *fired) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, r = lane & 15;
  // Every global load of the prologue is issued before anything waits on
  // one: the data rows, then (LDSC) the constants and the template
  // fragments, so the workgroup pays one memory latency, not one per loop
  // trip.  The data rows at the A-fragment positions f = 128 q + 16 i + r; a
  // row past nd (the wave's last, when nd is not a multiple of R) is read
  // from row nd - 1: it is transformed like any row, and its results are
  // never gated, counted or written.
  cf x[R][8];
#pragma unroll
  for (int k = 0; k < R; k++) {
    const cf *D = data + (long long)min(d0 + k, nd - 1) * data_stride;
#pragma unroll
    for (int i = 0; i < 8; i++) x[k][i] = D[128 * q + 16 * i + r];
  }
  if constexpr (LDSC) {
    // global table: lane-major, 12 words4 per lane (768 = 3 per thread);
    // LDS: word4-major
    const uint4 *src = reinterpret_cast<const uint4 *>(consts);
    const uint4 cw0 = src[threadIdx.x], cw1 = src[threadIdx.x + 256], cw2 = src[threadIdx.x + 512];
    const int tw0 = blockIdx.x * tpw, tcnt = min(nt, tw0 + tpw) - tw0;  // >= 1
    const int se = (int)threadIdx.x < tcnt ? sexp[tw0 + threadIdx.x] : 0;
    const uint4 *tsrc = reinterpret_cast<const uint4 *>(ctm) + (long long)tw0 * 256 + threadIdx.x;
    uint4 *tdst = tfrag + ((threadIdx.x & 3) * 64 + ((threadIdx.x >> 2) & 63));  // word4 256 j + threadIdx.x
    // four templates' words per trip, read from a clamped (valid) template
    // and stored only where the template exists
    for (int j0 = 0; j0 < tcnt; j0 += 4) {
      const int jl = tcnt - 1;
      const uint4 v0 = tsrc[min(j0, jl) * 256], v1 = tsrc[min(j0 + 1, jl) * 256];
      const uint4 v2 = tsrc[min(j0 + 2, jl) * 256], v3 = tsrc[min(j0 + 3, jl) * 256];
      if (j0 == 0) {
        if ((int)threadIdx.x < tcnt) tsexp[threadIdx.x] = se;
        ctab[(threadIdx.x % 12) * 64 + threadIdx.x / 12] = cw0;
        ctab[((threadIdx.x + 256) % 12) * 64 + (threadIdx.x + 256) / 12] = cw1;
        ctab[((threadIdx.x + 512) % 12) * 64 + (threadIdx.x + 512) / 12] = cw2;
      }
      tdst[j0 * 256] = v0;
      if (j0 + 1 < tcnt) tdst[(j0 + 1) * 256] = v1;
      if (j0 + 2 < tcnt) tdst[(j0 + 2) * 256] = v2;
      if (j0 + 3 < tcnt) tdst[(j0 + 3) * 256] = v3;
    }
    __syncthreads();
  }
  if (g.zero && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *g.zero = 0;

  // each row scaled by an exact power of two to components below 1/2.  The
  // exponent comes from the largest |component|, taken on the bit patterns
  // with the sign cleared (nonnegative floats order like their bits): no
  // NaN canonicalisation per component.  A NaN component gives the row the
  // largest exponent; its pairs' statistics are NaN either way.
  int sd[R];
#pragma unroll
  for (int k = 0; k < R; k++) {
    uint32_t mb = 0u;
#pragma unroll
    for (int i = 0; i < 8; i++)
      mb = max(mb, max(__float_as_uint(x[k][i].x) & 0x7fffffffu, __float_as_uint(x[k][i].y) & 0x7fffffffu));
    sd[k] = unit_exponent(__uint_as_float(wave_umax(mb))) + 1;  // components below 1/2
#pragma unroll
    for (int i = 0; i < 8; i++) x[k][i] = cf{ldexpf(x[k][i].x, -sd[k]), ldexpf(x[k][i].y, -sd[k])};
  }
  CoarseDft<BF16> dft;
  if constexpr (!LDSC) dft.load_table(consts, lane);

  // template t's pairs: the transforms of all R rows first (independent
  // chains the scheduler interleaves), then the reductions; gate decisions
  // are kept as wave-uniform bits per row and template (tpw <= 64) and
  // written out once, after the last template
#pragma unroll
  for (int k = 0; k < R; k++) fired[k] = 0;
  auto pairs = [&](const cf *ct, int t) {
    uint32_t lm[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
      if constexpr (LDSC)
        lm[k] = dft.template lane_max_lds<FULL>(ctab, lane, x[k], ct, cs, ce, q, r);
      else
        lm[k] = dft.template lane_max<FULL>(x[k], ct, cs, ce, q, r);
    }
    // the template's scale exponent: from LDS (staged with the fragments)
    // rather than a global load whose latency the pair chain would wait on
    const int st = LDSC ? tsexp[t - t0] : sexp[t];
#pragma unroll
    for (int k = 0; k < R; k++) {
      if (d0 + k >= nd) continue;
      bool fire;
      if (cmax2) {
        // the statistic itself is wanted: the wave maximum, rescaled
        const float m2 = ldexpf(__uint_as_float(wave_umax(lm[k])), 2 * (st + sd[k]));
        if (lane == 0) cmax2[(long long)(d0 + k) * nt + t] = m2;
        fire = g.thr2 <= 0.f || m2 >= g.thr2;
      } else {
        // only the decision: ldexp is monotonic, so "the rescaled wave
        // maximum reaches thr2" is "some lane's rescaled maximum does" --
        // unless a lane holds a NaN (bits above +inf), which the maximum
        // would return and no comparison passes
        const float m2l = ldexpf(__uint_as_float(lm[k]), 2 * (st + sd[k]));
        const bool any_nan = __ballot(lm[k] > 0x7f800000u) != 0ull;
        fire = !any_nan && (g.thr2 <= 0.f || __ballot(m2l >= g.thr2) != 0ull);
      }
      if (fire) fired[k] |= 1ull << (t - t0);
    }
  };

  if constexpr (LDSC) {
    for (int t = t0; t < t1; t++) {
      cf ct[8];
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) {
        const uint4 v = tfrag[((t - t0) * 4 + q4) * 64 + lane];
        ct[2 * q4] = cf{__uint_as_float(v.x), __uint_as_float(v.y)};
        ct[2 * q4 + 1] = cf{__uint_as_float(v.z), __uint_as_float(v.w)};
      }
      pairs(ct, t);
    }
  } else {
    // two fragment buffers in turn: the next template's loads are in flight
    // while the current one is transformed
    cf ca[8], cb[8];
    if (t0 < t1) load_frag8(ca, ctm, t0, lane);
    for (int t = t0; t < t1; t += 2) {
      if (t + 1 < t1) load_frag8(cb, ctm, t + 1, lane);
      pairs(ca, t);
      if (t + 1 >= t1) break;
      if (t + 2 < t1) load_frag8(ca, ctm, t + 2, lane);
      pairs(cb, t + 1);
    }
  }
}

// A wave's first data row d0 and its templates [t0, t1): the workgroup's
// four waves take 4 R consecutive rows and the same tpw templates.  A wave
// past the last data row walks no templates but stays for the workgroup's
// barriers.
template <int R>
__device__ __forceinline__ void coarse_tile(int nd, int nt, int tpw, int &d0, int &t0, int &t1) {
  d0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * R;
  t0 = blockIdx.x * tpw;
  t1 = d0 < nd ? min(nt, t0 + tpw) : t0;
}

template <bool BF16, int R, bool FULL, bool LDSC>
__global__ void __launch_bounds__(256) coarse_mfma_kernel(const cf *data, long long data_stride, const cf *ctm,
                                                          const int *sexp, int nd, int nt, int tpw, int cs, int ce,
                                                          const uint32_t *consts, float *cmax2, CoarseGate g) {
  __shared__ uint4 ctab[LDSC ? 12 * 64 : 1];
  __shared__ int tsexp[LDSC ? 64 : 1];
  extern __shared__ uint4 tfrag[];
  const int lane = threadIdx.x & 63;
  int d0, t0, t1;
  coarse_tile<R>(nd, nt, tpw, d0, t0, t1);
  uint64_t fired[R];
  coarse_screen<BF16, R, FULL, LDSC>(data, data_stride, ctm, sexp, nd, nt, tpw, cs, ce, consts, cmax2, g, ctab,
                                     tsexp, tfrag, d0, t0, t1, fired);
  if (!g.list) return;  // uniform over the grid

  // survivors: one atomic per workgroup reserves their list slots (a single
  // counter takes every workgroup's atomic, so fewer is faster); the other
  // pairs get -1 / 0 in all their bins, lanes across (template, bin)
  __shared__ int wave_cnt[4], wg_base;
  const int w = threadIdx.x >> 6;
  int total = 0;
#pragma unroll
  for (int k = 0; k < R; k++) total += __popcll(fired[k]);
  if (lane == 0) wave_cnt[w] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int sum = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    wg_base = sum ? atomicAdd(g.count, sum) : 0;
  }
  __syncthreads();
  int base = wg_base;
  for (int i = 0; i < w; i++) base += wave_cnt[i];
  const int cnt = t1 - t0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    if (d0 + k >= nd) break;
    const uint64_t f = fired[k];
    if (lane < cnt && ((f >> lane) & 1ull))
      g.list[base + __popcll(f & ((1ull << lane) - 1ull))] = int2{d0 + k, t0 + lane};
    base += __popcll(f);
    if (__popcll(f) == cnt) continue;
    const long long o = ((long long)(d0 + k) * nt + t0) * g.nb;
    if (g.nb == 1) {
      // one bin per pair (tpw <= 64): lane j writes template t0 + j
      if (lane < cnt && !((f >> lane) & 1ull)) {
        g.idx[o + lane] = -1;
        g.val[o + lane] = cf{0.f, 0.f};
      }
      continue;
    }
    for (int e = lane; e < cnt * g.nb; e += 64) {
      const int j = e / g.nb;
      if (!((f >> j) & 1ull)) {
        g.idx[o + e] = -1;
        g.val[o + e] = cf{0.f, 0.f};
      }
    }
  }
}

// Screen and refine in one launch (transform length FN, one bin, the
// constants in LDS): each workgroup gates its pairs exactly as
// coarse_mfma_kernel does, then runs the flat filter's per-pair path on its
// own survivors with all 256 threads (one FN = 4096 transform at a time), so
// no survivor list leaves the workgroup and no second launch waits on the
// first.  The number of survivors still reaches g.count (one atomic per
// workgroup that has any, after its refines) for the caller's refinement
// count.  The screen's
// LDS (constants, exponents, fragments) and the refine's (exchange buffer,
// bin keys) share one dynamic allocation; the refine starts after a barrier
// that ends every read of the screen's.  A workgroup's refines run one after
// another, so a workgroup with many survivors finishes late: the split is
// for screens that pass few pairs.
template <bool BF16, int R, bool FULL, int FN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
coarse_refine_kernel(const cf *data, long long data_stride, const cf *ctm, const int *sexp, int nd, int nt, int tpw,
                     int cs, int ce, const uint32_t *consts, CoarseGate g, FlatArgs fa) {
  using S = FlatShape<FN>;
  static_assert(S::WG == 256 && S::SLOTS == 1 && S::TPT == 256, "one transform per 256-thread workgroup");
  extern __shared__ __align__(16) unsigned char smem[];
  uint4 *ctab = reinterpret_cast<uint4 *>(smem);
  int *tsexp = reinterpret_cast<int *>(ctab + 12 * 64);
  uint4 *tfrag = reinterpret_cast<uint4 *>(tsexp + 64);
  __shared__ int2 surv[4 * R * 16];  // the workgroup's survivors (tpw <= 16)
  __shared__ int wave_cnt[4], wg_n;
  const int lane = threadIdx.x & 63;
  int d0, t0, t1;
  coarse_tile<R>(nd, nt, tpw, d0, t0, t1);
  uint64_t fired[R];
  coarse_screen<BF16, R, FULL, true>(data, data_stride, ctm, sexp, nd, nt, tpw, cs, ce, consts, nullptr, g, ctab,
                                     tsexp, tfrag, d0, t0, t1, fired);

  // survivors into the workgroup's list; the other pairs get -1 / 0
  const int w = threadIdx.x >> 6;
  int total = 0;
#pragma unroll
  for (int k = 0; k < R; k++) total += __popcll(fired[k]);
  if (lane == 0) wave_cnt[w] = total;
  __syncthreads();
  if (threadIdx.x == 0) wg_n = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
  int base = 0;
  for (int i = 0; i < w; i++) base += wave_cnt[i];
  const int cnt = t1 - t0;
#pragma unroll
  for (int k = 0; k < R; k++) {
    if (d0 + k >= nd) break;
    const uint64_t f = fired[k];
    const bool mine = lane < cnt && ((f >> lane) & 1ull);
    if (mine) surv[base + __popcll(f & ((1ull << lane) - 1ull))] = int2{d0 + k, t0 + lane};
    base += __popcll(f);
    // one bin per pair: lane j answers template t0 + j
    const long long o = (long long)(d0 + k) * nt + t0 + lane;
    if (lane < cnt && !mine) {
      g.idx[o] = -1;
      g.val[o] = cf{0.f, 0.f};
    }
  }
  __syncthreads();  // the list is complete and the screen's LDS no longer read
  const int ns = wg_n;
  if (ns == 0) return;  // workgroup-uniform

  // the flat kernel's per-pair path (list mode), pair after pair
  cf *xch = reinterpret_cast<cf *>(smem);
  uint64_t *bins_all = reinterpret_cast<uint64_t *>(xch + S::XSTRIDE);
  cf wfirst[15];
  first_twiddles<FN>(wfirst, (int)threadIdx.x, fa.twiddle);
  if (threadIdx.x < 2) bins_all[threadIdx.x] = 0ull;
  __syncthreads();
  for (int i = 0; i < ns; i++) {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    const int2 pr = surv[i];
    cf dv[16], tv[16], v[16];
    load_row_uniform<FN>(dv, uniform_ptr(fa.data + (long long)pr.x * fa.data_stride), t);
    load_row_uniform<FN>(tv, uniform_ptr(fa.tmpl + (long long)pr.y * fa.tmpl_stride), t);
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = cmul_conj(dv[u], tv[u]);
    inverse_fft<FN, true>(v, t, xch, fa.twiddle, wfirst);
    select_and_write<FN, kOneBin, false>(fa, 1, v, t, lane, bins_all + (i & 1), bins_all + ((i + 1) & 1), true, pr.x,
                                         pr.y, 0, 0);
  }
  // the count last: issued before the refines, every workgroup's atomic on
  // the one counter would hold them back until it completed
  if (threadIdx.x == 0) atomicAdd(g.count, ns);
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

// The per-lane constant table of CoarseDft<BF16>, from the FP64-generated
// roots; one wave.
template <bool BF16>
__global__ void coarse_consts_kernel(const cf *w32, const cf *w512, const cf *w16, uint32_t *tab) {
  const int lane = threadIdx.x;
  CoarseDft<BF16> dft;
  dft.load(w32, w512, w16, lane >> 4, lane & 15);
  dft.store_table(tab, lane);
}

}  // namespace

// Roots of unity of orders 32, 512, 16 for the factored DFT, built once.
// Then the per-lane constant tables for fp16 (first) and bf16 operands.
static const uint32_t *coarse_consts(Ctx *c, bool bf16) {
  constexpr int kWords = 64 * CoarseDft<false>::kLaneWords;
  cf *r = (cf *)c->scratch("coarse_roots", (32 + 512 + 16) * sizeof(cf));
  uint32_t *tab = (uint32_t *)c->scratch("coarse_consts", 2 * kWords * sizeof(uint32_t));
  if (!r || !tab) return nullptr;
  if (!c->coarse_roots_ready) {
    hipLaunchKernelGGL(unit_roots_kernel, dim3(1), dim3(64), 0, c->stream, r, 32);
    hipLaunchKernelGGL(unit_roots_kernel, dim3(2), dim3(256), 0, c->stream, r + 32, 512);
    hipLaunchKernelGGL(unit_roots_kernel, dim3(1), dim3(64), 0, c->stream, r + 544, 16);
    hipLaunchKernelGGL(coarse_consts_kernel<false>, dim3(1), dim3(64), 0, c->stream, r, r + 32, r + 544, tab);
    hipLaunchKernelGGL(coarse_consts_kernel<true>, dim3(1), dim3(64), 0, c->stream, r, r + 32, r + 544,
                       tab + kWords);
    if (!hip_ok(hipGetLastError(), "coarse_consts_kernel")) return nullptr;
    c->coarse_roots_ready = true;
  }
  return tab + (bf16 ? kWords : 0);
}

int coarse_mfma_prepare(Ctx *c) {
  if (c->coarse_band != kBand) return 0;
  cf *ctm = (cf *)c->scratch("coarse_mfma_tmpl", (size_t)c->nt * kBand * sizeof(cf));
  int *sexp = (int *)c->scratch("coarse_mfma_exp", (size_t)c->nt * sizeof(int));
  if (!ctm || !sexp) return -1;
  hipLaunchKernelGGL(coarse_mfma_prep_kernel, dim3((unsigned)c->nt), dim3(64), 0, c->stream, c->coarse_tmpl, ctm,
                     sexp);
  if (!hip_ok(hipGetLastError(), "coarse_mfma_prep_kernel")) return -1;
  c->coarse_ctm = ctm;
  c->coarse_sexp = sexp;
  return 0;
}

int launch_coarse_mfma(Ctx *c, const cf *data, long long data_stride, int t0, int nd, int nt, int cs, int ce,
                       bool bf16, float *cmax2, const CoarseGate *gate) {
  if (c->coarse_band != kBand || !c->coarse_ctm) return fail("the MFMA coarse screen is built for band 512");
  const uint32_t *consts = coarse_consts(c, bf16);
  if (!consts) return -1;
  const int tpw = std::min(64, std::max(1, c->coarse_tpw));
  const int rows = c->coarse_rpw == 1 ? 1 : 2;
  dim3 grid((unsigned)((nt + tpw - 1) / tpw), (unsigned)((nd + 4 * rows - 1) / (4 * rows)));
  if (grid.y > 65535) return fail("too many data rows per launch (%d)", nd);
  const cf *ctm = c->coarse_ctm + (size_t)t0 * kBand;
  const int *sexp = c->coarse_sexp + t0;
  const bool full = cs == 0 && ce == kBand;  // whole coarse window: no per-lag test
  using Kern = decltype(&coarse_mfma_kernel<true, 1, true, false>);
  // [bf16][rows - 1][full][lds]
  static const Kern table[2][2][2][2] = {
      {{{coarse_mfma_kernel<false, 1, false, false>, coarse_mfma_kernel<false, 1, false, true>},
        {coarse_mfma_kernel<false, 1, true, false>, coarse_mfma_kernel<false, 1, true, true>}},
       {{coarse_mfma_kernel<false, 2, false, false>, coarse_mfma_kernel<false, 2, false, true>},
        {coarse_mfma_kernel<false, 2, true, false>, coarse_mfma_kernel<false, 2, true, true>}}},
      {{{coarse_mfma_kernel<true, 1, false, false>, coarse_mfma_kernel<true, 1, false, true>},
        {coarse_mfma_kernel<true, 1, true, false>, coarse_mfma_kernel<true, 1, true, true>}},
       {{coarse_mfma_kernel<true, 2, false, false>, coarse_mfma_kernel<true, 2, false, true>},
        {coarse_mfma_kernel<true, 2, true, false>, coarse_mfma_kernel<true, 2, true, true>}}}};
  // LDS staging holds at most 16 templates (64 KiB); wider tiles keep the
  // constants and fragments in registers
  const bool lds = c->coarse_lds && tpw <= 16;
  const Kern kern = table[bf16 ? 1 : 0][rows - 1][full ? 1 : 0][lds ? 1 : 0];
  const size_t smem = lds ? (size_t)tpw * 4096 : 0;
  if (lds && !lds_limit((const void *)kern, c->dev, smem)) return -1;
  Ctx::ProfScope ps(c, "coarse");
  const CoarseGate g = gate ? *gate : CoarseGate{};
  hipLaunchKernelGGL(kern, grid, dim3(256), smem, c->stream, data, data_stride, ctm, sexp, nd, nt, tpw, cs, ce,
                     consts, cmax2, g);
  return hip_ok(hipGetLastError(), "coarse_mfma_kernel") ? 0 : -1;
}

bool coarse_refine_fits(const Ctx *c, size_t n, size_t nb) {
  return c->coarse_fuse && n == 4096 && nb == 1 && c->coarse_band == kBand && c->coarse_lds && c->coarse_tpw >= 1 &&
         c->coarse_tpw <= 16;
}

int launch_coarse_refine(Ctx *c, const cf *data, long long data_stride, int t0, int nd, int nt, int cs, int ce,
                         bool bf16, const CoarseGate &g, const FlatRun &rr) {
  constexpr int FN = 4096;
  if (c->coarse_band != kBand || !c->coarse_ctm) return fail("the MFMA coarse screen is built for band 512");
  if (!coarse_refine_fits(c, FN, 1)) return fail("the one-launch screen and refine does not apply here");
  const uint32_t *consts = coarse_consts(c, bf16);
  if (!consts) return -1;
  FlatArgs fa;
  if (flat_fill_args(c, FN, rr, fa) < 0) return -1;
  if (fa.nb != 1) return fail("the one-launch screen and refine writes one bin per pair");
  const int tpw = c->coarse_tpw;  // <= 16: the survivor list holds 4 R tpw pairs
  const int rows = c->coarse_rpw == 1 ? 1 : 2;
  dim3 grid((unsigned)((nt + tpw - 1) / tpw), (unsigned)((nd + 4 * rows - 1) / (4 * rows)));
  if (grid.y > 65535) return fail("too many data rows per launch (%d)", nd);
  const cf *ctm = c->coarse_ctm + (size_t)t0 * kBand;
  const int *sexp = c->coarse_sexp + t0;
  const bool full = cs == 0 && ce == kBand;
  using Kern = decltype(&coarse_refine_kernel<true, 1, true, FN>);
  // [bf16][rows - 1][full]
  static const Kern table[2][2][2] = {
      {{coarse_refine_kernel<false, 1, false, FN>, coarse_refine_kernel<false, 1, true, FN>},
       {coarse_refine_kernel<false, 2, false, FN>, coarse_refine_kernel<false, 2, true, FN>}},
      {{coarse_refine_kernel<true, 1, false, FN>, coarse_refine_kernel<true, 1, true, FN>},
       {coarse_refine_kernel<true, 2, false, FN>, coarse_refine_kernel<true, 2, true, FN>}}};
  const Kern kern = table[bf16 ? 1 : 0][rows - 1][full ? 1 : 0];
  // the screen's constants, exponents and fragments, or the refine's
  // exchange buffer and two one-bin key slots, whichever is larger
  const size_t screen = 12 * 64 * sizeof(uint4) + 64 * sizeof(int) + (size_t)tpw * 4096;
  const size_t refine = (size_t)FlatShape<FN>::XSTRIDE * sizeof(cf) + 2 * sizeof(uint64_t);
  const size_t smem = std::max(screen, refine);
  if (!lds_limit((const void *)kern, c->dev, smem)) return -1;
  Ctx::ProfScope ps(c, "coarse");
  hipLaunchKernelGGL(kern, grid, dim3(256), smem, c->stream, data, data_stride, ctm, sexp, nd, nt, tpw, cs, ce,
                     consts, g, fa);
  return hip_ok(hipGetLastError(), "coarse_refine_kernel") ? 0 : -1;
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
