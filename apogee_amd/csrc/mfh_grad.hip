// Gradients of reported peak values with respect to the data and template
// spectra, from the stored peak indices alone (no correlation series).
//
// A reported peak of pair (d, t) at lag k has the value
//     z = sum_f D_d[f] conj(T_t[f]) w^(f k),   w = exp(+2 pi i / n),
// so, with g = dL/dRe z + i dL/dIm z the incoming gradient of a real loss
// (PyTorch's convention for complex tensors),
//     dL/dD_d[f] = sum_{t, bins} g       T_t[f] w^(-f k)
//     dL/dT_t[f] = sum_{d, bins} conj(g) D_d[f] w^(+f k)
// Bins that reported nothing (index -1) hold the constant 0 and contribute
// nothing.  Each output element is one thread's ordered sum (no atomics), so
// the result is deterministic.  w^m, m = (f k) mod n in exact integer
// arithmetic, is the product of two FP64-generated roots, w^(256 (m >> 8)) and
// w^(m & 255), read from a 4 KiB copy of both tables in LDS: the indices are
// scattered across lanes, which the vector cache serves poorly.
//
// The same table gives the peak value itself at a stored lag (a direct sum
// over f per peak, one wave each), which is how a caller recomputes z at the
// argmax without keeping the series.
//
// Reference behaviour: the forward value is the flat filter's
// (src/balanced-inl.h:876-969); the reference has no derivative.
#include <hip/hip_runtime.h>

#include "mfh_dft.h"
#include "mfh_internal.h"
#include "mfhip.h"

namespace mfh {

namespace {

// Two-level root table of length n: hi[h] = w^(256 h) for h < max(1, n/256),
// then lo[l] = w^l for l < 256.
__global__ void circle_tables_kernel(cf *out, int n) {
  const int nh = n >= 256 ? n / 256 : 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nh + 256) {
    const long long e = i < nh ? 256ll * i : (long long)(i - nh);
    double sn, cs;
    sincospi(2.0 * (double)(e % n) / (double)n, &sn, &cs);
    out[i] = cf{(float)cs, (float)sn};
  }
}

// The block's LDS copy of the two tables; returns (hi, lo) in *lds.
__device__ __forceinline__ void stage_circle(const cf *tables, int n, cf *lds) {
  const int nh = n >= 256 ? n / 256 : 1;
  for (int i = threadIdx.x; i < nh + 256; i += blockDim.x) lds[i] = tables[i];
  __syncthreads();
}

__device__ __forceinline__ cf circle(const cf *lds, int n, uint32_t m) {
  const int nh = n >= 256 ? n / 256 : 1;
  return cmul(lds[m >> 8], lds[nh + (m & 255u)]);
}

// dL/dT (TEMPL = true) or dL/dD: one workgroup per (row, 64 F consecutive f);
// the row is blockIdx.y, the other side's rows are summed over.  The row's
// (lag, gradient) list is staged through LDS in chunks, so the inner loop
// reads LDS instead of waiting on one global load per term.  The eight
// waves take the other side's rows in turn (row j to wave j mod 8), so each
// reads its row's F spectrum values once for all of that row's bins; a
// fixed-order sum of the eight partial sums keeps the result deterministic.
// Each lane covers F consecutive f: w^(f k) for the first comes from the
// tables (one scattered lookup), the rest by F - 1 products with w^k (one
// lookup at a uniform address per term), so scattered LDS reads -- whose
// bank conflicts bound this kernel -- drop F-fold.
template <bool TEMPL, int F>
__global__ void __launch_bounds__(512) peak_grad_kernel(const cf *__restrict__ other, long long other_stride,
                                                        int nother, const int64_t *__restrict__ idx,
                                                        const cf *__restrict__ gval, int nd, int nt, int nb,
                                                        const cf *__restrict__ tables, int n, cf *__restrict__ out,
                                                        long long out_stride) {
  constexpr int CH = 1024;  // (lag, gradient) entries per LDS chunk
  constexpr int W = 8;      // waves
  __shared__ cf roots[512];
  __shared__ uint32_t sk[CH];
  __shared__ cf sg[CH];
  __shared__ cf part[W][64 * F];
  stage_circle(tables, n, roots);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int fb = blockIdx.x * 64 * F;  // n is a multiple of 64 F
  const int f0 = fb + lane * F;
  const int row = blockIdx.y;
  const uint32_t mask = (uint32_t)n - 1u;
  float ax[F], ay[F];
#pragma unroll
  for (int q = 0; q < F; q++) ax[q] = ay[q] = 0.f;
  const long long total = (long long)nother * nb;
  for (long long e0 = 0; e0 < total; e0 += CH) {
    const int cnt = (int)(total - e0 < CH ? total - e0 : CH);
    __syncthreads();  // the previous chunk is consumed
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      const long long e = e0 + i;
      const int j = (int)(e / nb), b = (int)(e % nb);
      const long long pair = TEMPL ? (long long)j * nt + row : (long long)row * nt + j;
      const long long k = idx[pair * nb + b];
      sk[i] = k < 0 ? 0xffffffffu : (uint32_t)k;
      sg[i] = gval[pair * nb + b];
    }
    __syncthreads();
    // this wave's rows of the chunk: j = w (mod W)
    const long long j0 = e0 / nb, j1 = (e0 + cnt - 1) / nb;
    for (long long j = j0 + ((w - j0 % W) + W) % W; j <= j1; j += W) {
      const int ib = (int)(j * nb - e0 > 0 ? j * nb - e0 : 0);
      const int ie = (int)((j + 1) * nb - e0 < cnt ? (j + 1) * nb - e0 : cnt);
      cf x[F];
      const cf *xr = other + j * other_stride + f0;
#pragma unroll
      for (int q = 0; q < F; q++) x[q] = xr[q];
      for (int i = ib; i < ie; i++) {
        const uint32_t k = sk[i];
        if (k == 0xffffffffu) continue;  // uniform: one term for the whole wave
        const cf g = sg[i];
        cf wk = circle(roots, n, ((uint32_t)f0 * k) & mask);
        const cf step = circle(roots, n, k & mask);
#pragma unroll
        for (int q = 0; q < F; q++) {
          if (q > 0) wk = cmul(wk, step);
          // templates: conj(g) x w ; data: g x conj(w)
          cf xw;
          if (TEMPL) {
            xw = cmul(x[q], wk);
            ax[q] += g.x * xw.x + g.y * xw.y;
            ay[q] += g.x * xw.y - g.y * xw.x;
          } else {
            xw = cmul_conj(x[q], wk);
            ax[q] += g.x * xw.x - g.y * xw.y;
            ay[q] += g.x * xw.y + g.y * xw.x;
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < F; q++) part[w][lane * F + q] = cf{ax[q], ay[q]};
  __syncthreads();
  for (int c = threadIdx.x; c < 64 * F; c += blockDim.x) {
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int v = 0; v < W; v++) {
      sx += part[v][c].x;
      sy += part[v][c].y;
    }
    out[(long long)row * out_stride + fb + c] = cf{sx, sy};
  }
}

// z at stored lags: one wave per (pair, bin), FP32 throughout (FP64 on the
// device is kept to table generation and band-fraction sums): each lane sums
// its n/64 terms, then a fixed-order butterfly adds the lanes.
__global__ void __launch_bounds__(256) peak_value_kernel(const cf *__restrict__ D, long long dstride,
                                                         const cf *__restrict__ T, long long tstride, int nt,
                                                         int nb, long long npeaks, const int64_t *__restrict__ idx,
                                                         const cf *__restrict__ tables, int n, cf *__restrict__ out) {
  __shared__ cf roots[512];
  stage_circle(tables, n, roots);
  const long long p = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= npeaks) return;  // uniform per wave, after the only barrier
  const long long k = idx[p];
  const long long pair = p / nb;
  const int d = (int)(pair / nt), t = (int)(pair % nt);
  float sx = 0.f, sy = 0.f;
  if (k >= 0) {
    const uint32_t mask = (uint32_t)n - 1u;
    const cf *dp = D + (long long)d * dstride;
    const cf *tp = T + (long long)t * tstride;
    for (int f = lane; f < n; f += 64) {
      const cf pr = cmul(cmul_conj(dp[f], tp[f]), circle(roots, n, ((uint32_t)f * (uint32_t)k) & mask));
      sx += pr.x;
      sy += pr.y;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off, 64);
    sy += __shfl_xor(sy, off, 64);
  }
  if (lane == 0) out[p] = k >= 0 ? cf{sx, sy} : cf{0.f, 0.f};
}

}  // namespace

static const cf *unit_circle(Ctx *c) {
  char name[32];
  snprintf(name, sizeof name, "circle%zu", c->n);
  auto it = c->bufs.find(name);
  if (it != c->bufs.end() && it->second.p) return (const cf *)it->second.p;
  const int nh = c->n >= 256 ? (int)(c->n / 256) : 1;
  cf *r = (cf *)c->scratch(name, (size_t)(nh + 256) * sizeof(cf));
  if (!r) return nullptr;
  hipLaunchKernelGGL(circle_tables_kernel, dim3((unsigned)((nh + 256 + 255) / 256)), dim3(256), 0, c->stream, r,
                     (int)c->n);
  if (!hip_ok(hipGetLastError(), "circle_tables_kernel")) return nullptr;
  return r;
}

}  // namespace mfh

using mfh::cf;
using mfh::Ctx;
using mfh::fail;
using mfh::hip_ok;

extern "C" int mfh_peak_grad_async(mfh_ctx *ctx, const void *dev_data_c64, long long data_stride, int nd,
                                   const void *dev_tmpl_c64, long long tmpl_stride, int nt, size_t nb,
                                   const int64_t *dev_idx, const void *dev_gval_c64, void *dev_grad_data_c64,
                                   void *dev_grad_tmpl_c64) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (nd < 1 || nt < 1 || nb < 1) return fail("empty peak array (%d x %d x %zu)", nd, nt, nb);
  if (!dev_data_c64 || !dev_tmpl_c64 || !dev_idx || !dev_gval_c64) return fail("null input pointer");
  if (data_stride < (long long)c->n || tmpl_stride < (long long)c->n) return fail("row stride below n");
  if (nd > 65535 || nt > 65535) return fail("at most 65535 data and template rows per call");
  if (c->n > 65536 || c->n < 64) return fail("peak gradients are built for 64 <= n <= 65536");
  hipSetDevice(c->dev);
  const cf *roots = mfh::unit_circle(c);
  if (!roots) return -1;
  const int n = (int)c->n;
  const dim3 block(512);
  // four frequencies per lane where n allows (n >= 256), else one
  const bool f4 = n >= 256;
  const int fcols = n / (f4 ? 256 : 64);
  auto *ktmpl = f4 ? &mfh::peak_grad_kernel<true, 4> : &mfh::peak_grad_kernel<true, 1>;
  auto *kdata = f4 ? &mfh::peak_grad_kernel<false, 4> : &mfh::peak_grad_kernel<false, 1>;
  if (dev_grad_tmpl_c64) {
    hipLaunchKernelGGL(ktmpl, dim3(fcols, nt), block,
                       0, c->stream, (const cf *)dev_data_c64, data_stride, nd, dev_idx, (const cf *)dev_gval_c64, nd,
                       nt, (int)nb, roots, n, (cf *)dev_grad_tmpl_c64, (long long)n);
    if (!hip_ok(hipGetLastError(), "peak_grad_kernel")) return -1;
  }
  if (dev_grad_data_c64) {
    hipLaunchKernelGGL(kdata, dim3(fcols, nd), block,
                       0, c->stream, (const cf *)dev_tmpl_c64, tmpl_stride, nt, dev_idx, (const cf *)dev_gval_c64, nd,
                       nt, (int)nb, roots, n, (cf *)dev_grad_data_c64, (long long)n);
    if (!hip_ok(hipGetLastError(), "peak_grad_kernel")) return -1;
  }
  return 0;
}

extern "C" int mfh_peak_values_async(mfh_ctx *ctx, const void *dev_data_c64, long long data_stride, int nd,
                                     const void *dev_tmpl_c64, long long tmpl_stride, int nt, size_t nb,
                                     const int64_t *dev_idx, void *dev_out_c64) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (nd < 1 || nt < 1 || nb < 1) return fail("empty peak array (%d x %d x %zu)", nd, nt, nb);
  if (!dev_data_c64 || !dev_tmpl_c64 || !dev_idx || !dev_out_c64) return fail("null pointer");
  if (data_stride < (long long)c->n || tmpl_stride < (long long)c->n) return fail("row stride below n");
  if (c->n > 65536) return fail("peak values are built for n <= 65536");
  hipSetDevice(c->dev);
  const cf *roots = mfh::unit_circle(c);
  if (!roots) return -1;
  const long long npeaks = (long long)nd * nt * (long long)nb;
  hipLaunchKernelGGL(mfh::peak_value_kernel, dim3((unsigned)((npeaks + 3) / 4)), dim3(256), 0, c->stream,
                     (const cf *)dev_data_c64, data_stride, (const cf *)dev_tmpl_c64, tmpl_stride, nt, (int)nb, npeaks,
                     dev_idx, roots, (int)c->n, (cf *)dev_out_c64);
  return hip_ok(hipGetLastError(), "peak_value_kernel") ? 0 : -1;
}
