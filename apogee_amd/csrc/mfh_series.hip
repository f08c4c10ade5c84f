// run_series on the device: gather each block of a time series, zero-pad,
// scale by 1/n, forward FFT, and filter the block spectra -- no spectrum
// crosses the bus.
//
// Reference behaviour (paths relative to the reference checkout):
//   block spectrum FFT(series[s:s+n])/n, zero-padded,
//   1/n applied before the transform (exact for n = 2^k)   src/matchfilt.c:437-490
//   equal-window runs filtered together, equal bin counts   src/matchfilt.c:449-488
//   GPU precedent (seriesForward via conj(IFFT(conj x)))    src/gpu/series_forward.slang:9-33
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mfh_flat.h"
#include "mfh_internal.h"
#include "mfhip.h"

namespace mfh {

namespace {

// One transform slot per block: thread t loads x[t + u*TPT] (conjugated and
// scaled by 1/N), runs the inverse FFT and writes conj(z) -- the forward
// transform -- in natural order.
template <int N>
__global__ void __launch_bounds__(FlatShape<N>::WG) __attribute__((amdgpu_waves_per_eu(4)))
series_forward_kernel(const cf *series, long long nseries, const BlockStarts starts, int nblocks, const cf *tw,
                      cf *out) {
  using S = FlatShape<N>;
  constexpr int TPT = S::TPT;
  extern __shared__ __align__(16) unsigned char smem[];
  cf *xch_all = reinterpret_cast<cf *>(smem);
  const int slot = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int b = blockIdx.x * S::SLOTS + slot;
  const bool valid = b < nblocks;
  const long long s0 = valid ? starts[b] : 0;
  const float inv = 1.0f / (float)N;
  cf v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const long long k = s0 + t + u * TPT;
    const cf x = (valid && k < nseries) ? series[k] : cf{0.f, 0.f};
    v[u] = cf{x.x * inv, -(x.y * inv)};
  }
  inverse_fft<N>(v, t, xch_all + slot * S::XSTRIDE, tw);
  if (valid) {
    cf *o = out + (long long)b * N;
#pragma unroll
    for (int u = 0; u < 16; u++) o[t + u * TPT] = cf{v[u].x, -v[u].y};
  }
}

template <int N>
int launch_fwd(Ctx *c, const cf *series, size_t nseries, const BlockStarts &starts, int nblocks, cf *out) {
  using S = FlatShape<N>;
  const cf *tw = c->twiddles(N);
  if (!tw) return -1;
  const size_t lds = (size_t)S::SLOTS * S::XSTRIDE * sizeof(cf);
  if (!lds_limit((const void *)series_forward_kernel<N>, c->dev, lds)) return -1;
  Ctx::ProfScope ps(c, "series_forward");
  hipLaunchKernelGGL(series_forward_kernel<N>, dim3((unsigned)((nblocks + S::SLOTS - 1) / S::SLOTS)), dim3(S::WG),
                     lds, c->stream, series, (long long)nseries, starts, nblocks, tw, out);
  return hip_ok(hipGetLastError(), "series_forward_kernel") ? 0 : -1;
}

// Four-step forward transform for n = N1 x 256 (the flat filter's split,
// f = f1 + N1 f2, k = k2 + 256 k1): columns take conj(x)/n along f2 through a
// 256-point inverse FFT and the twiddle w_n^(f1 k2) into scratch Y[k2][f1];
// rows run N1-point inverse FFTs over f1 and store conj(z) at k2 + 256 k1.
template <int N1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
series_big_cols(const cf *series, long long nseries, const BlockStarts starts, int b0, const cf *tw256,
                const cf *rootH, const cf *rootL, cf *Y) {
  constexpr int N = N1 * 256;
  extern __shared__ __align__(16) unsigned char smem[];
  const int s = threadIdx.x & 15;
  const int t = threadIdx.x >> 4;
  cf *xch = reinterpret_cast<cf *>(smem) + s * kColStride;
  const int b = blockIdx.y;
  const long long s0 = starts[b0 + b];
  const int f1 = blockIdx.x * 16 + s;
  const float inv = 1.0f / (float)N;
  cf v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const long long k = s0 + f1 + N1 * (t + 16 * u);
    const cf x = k < nseries ? series[k] : cf{0.f, 0.f};
    v[u] = cf{x.x * inv, -(x.y * inv)};
  }
  inverse_fft<256>(v, t, xch, tw256);
  cf *Yb = Y + (long long)b * N;
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int k2 = t + 16 * u;
    const int m = f1 * k2;
    Yb[(long long)k2 * N1 + f1] = cmul(v[u], cmul(rootH[m >> 8], rootL[m & 255]));
  }
}

template <int N1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
series_big_rows(const cf *Y, const cf *twN1, cf *out) {
  constexpr int N = N1 * 256;
  using S = FlatShape<N1>;
  constexpr int TPT = S::TPT;
  extern __shared__ __align__(16) unsigned char smem[];
  const int slot = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  cf *xch = reinterpret_cast<cf *>(smem) + slot * S::XSTRIDE;
  const int b = blockIdx.y;
  const int k2 = blockIdx.x * S::SLOTS + slot;
  const cf *Yb = Y + (long long)b * N;
  cf v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) v[u] = Yb[(long long)k2 * N1 + t + TPT * u];
  inverse_fft<N1>(v, t, xch, twN1);
  cf *ob = out + (long long)b * N;
#pragma unroll
  for (int u = 0; u < 16; u++) ob[k2 + 256 * (t + TPT * u)] = cf{v[u].x, -v[u].y};
}

template <int N1>
int launch_fwd_big(Ctx *c, const cf *series, size_t nseries, const BlockStarts &starts, int nblocks, cf *out) {
  constexpr int N = N1 * 256;
  using S = FlatShape<N1>;
  const cf *tw256 = c->twiddles(256);
  const cf *twN1 = c->twiddles(N1);
  const cf *roots = big_roots(c, N1);
  if (!tw256 || !twN1 || !roots) return -1;
  // scratch for at most 256 MiB of blocks at a time (and 65535 per launch)
  const int chunk = (int)std::max<long long>(1, std::min<long long>({(long long)nblocks, 65535ll,
                                                                     (256ll << 20) / ((long long)N * 8)}));
  cf *Y = (cf *)c->scratch("series_big_y", (size_t)chunk * N * sizeof(cf));
  if (!Y) return -1;
  const size_t lds_cols = 16 * kColStride * sizeof(cf);
  const size_t lds_rows = (size_t)S::SLOTS * S::XSTRIDE * sizeof(cf);
  Ctx::ProfScope ps(c, "series_forward");
  for (int b0 = 0; b0 < nblocks; b0 += chunk) {
    const int nb = std::min(chunk, nblocks - b0);
    hipLaunchKernelGGL(series_big_cols<N1>, dim3(N1 / 16, nb), dim3(256), lds_cols, c->stream, series,
                       (long long)nseries, starts, b0, tw256, roots, roots + N1, Y);
    hipLaunchKernelGGL(series_big_rows<N1>, dim3(256 / S::SLOTS, nb), dim3(256), lds_rows, c->stream, Y, twN1,
                       out + (long long)b0 * N);
  }
  return hip_ok(hipGetLastError(), "series_big") ? 0 : -1;
}

}  // namespace

int launch_series_forward(Ctx *c, const cf *series, size_t nseries, const BlockStarts &starts, int nblocks,
                          cf *out) {
  switch (c->n) {
    case 64: return launch_fwd<64>(c, series, nseries, starts, nblocks, out);
    case 128: return launch_fwd<128>(c, series, nseries, starts, nblocks, out);
    case 256: return launch_fwd<256>(c, series, nseries, starts, nblocks, out);
    case 512: return launch_fwd<512>(c, series, nseries, starts, nblocks, out);
    case 1024: return launch_fwd<1024>(c, series, nseries, starts, nblocks, out);
    case 2048: return launch_fwd<2048>(c, series, nseries, starts, nblocks, out);
    case 4096: return launch_fwd<4096>(c, series, nseries, starts, nblocks, out);
    case 8192: return launch_fwd<8192>(c, series, nseries, starts, nblocks, out);
    case 16384: return launch_fwd<16384>(c, series, nseries, starts, nblocks, out);
    case 32768: return launch_fwd_big<128>(c, series, nseries, starts, nblocks, out);
    case 65536: return launch_fwd_big<256>(c, series, nseries, starts, nblocks, out);
    case 131072: return launch_fwd_big<512>(c, series, nseries, starts, nblocks, out);
    case 262144: return launch_fwd_big<1024>(c, series, nseries, starts, nblocks, out);
    case 524288: return launch_fwd_big<2048>(c, series, nseries, starts, nblocks, out);
    case 1048576: return launch_fwd_big<4096>(c, series, nseries, starts, nblocks, out);
    default: return fail("transform length %zu is not supported", c->n);
  }
}

}  // namespace mfh

using mfh::cf;
using mfh::Ctx;
using mfh::fail;
using mfh::hip_ok;

namespace {

// The blocks' start offsets, clamped to nseries, as the forward kernels take
// them: in the kernel arguments up to BlockStarts::kInline blocks, else copied
// from `host` (which must outlive the queued copy) into a device table.
bool block_starts(Ctx *c, const size_t *start, size_t nseries, int nblocks, mfh::BlockStarts &bs,
                  std::vector<long long> &host) {
  const auto clamp = [&](int b) { return start[b] < nseries ? (long long)start[b] : (long long)nseries; };
  if (nblocks <= mfh::BlockStarts::kInline) {
    bs.dev = nullptr;
    for (int b = 0; b < nblocks; b++) bs.inl[b] = clamp(b);
    return true;
  }
  long long *dstart = (long long *)c->scratch("series_starts", (size_t)nblocks * sizeof(long long));
  if (!dstart) return false;
  host.resize(nblocks);
  for (int b = 0; b < nblocks; b++) host[b] = clamp(b);
  bs.dev = dstart;
  return hip_ok(hipMemcpyAsync(dstart, host.data(), nblocks * sizeof(long long), hipMemcpyHostToDevice, c->stream),
                "copy starts");
}

// Shared body of the flat and hierarchical run_series entry points.
long run_series(Ctx *c, bool hier, const void *series_c64, size_t nseries, int series_on_device, const size_t *start,
                const size_t *win_start, const size_t *win_end, int nblocks, int t0, int nt, size_t binsize,
                float threshold, float coarse_thr, int64_t *idx, void *val_c64, void *rec, int32_t *counts,
                int64_t *n_refined) {
  if (nblocks < 1) return fail("run_series needs at least one block");
  if (!series_c64 || !start || !win_start || !win_end || (!rec && (!idx || !val_c64))) return fail("null argument");
  if (t0 < 0 || nt < 1 || t0 + nt > c->nt) return fail("templates sub-range out of bounds");
  if (binsize < 1) return fail("binsize must be >= 1");
  const size_t n = c->n;
  const size_t nb = mfh_nbins(n, binsize, win_start[0], win_end[0]);
  if (nb == 0) return fail("every block must have a nonempty search window");
  for (int b = 0; b < nblocks; b++) {
    const size_t bb = mfh_nbins(n, binsize, win_start[b], win_end[b]);
    if (bb == 0) return fail("every block must have a nonempty search window");
    if (bb != nb) return fail("every block's window must give the same bin count");
  }
  hipSetDevice(c->dev);
  cf *spec = (cf *)c->scratch("series_spec", (size_t)nblocks * n * sizeof(cf));
  const size_t cells = (size_t)nblocks * nt * nb;
  mfh::PeakBuf pb;
  if (!spec || !mfh::peak_buffer(c, cells, pb)) return -1;
  int64_t *didx = pb.idx;
  cf *dval = pb.val;
  const cf *dser;
  if (series_on_device) {
    dser = (const cf *)series_c64;
  } else {
    cf *buf = (cf *)c->scratch("series_in", (nseries ? nseries : 1) * sizeof(cf));
    if (!buf) return -1;
    if (nseries && !hip_ok(hipMemcpyAsync(buf, series_c64, nseries * sizeof(cf), hipMemcpyHostToDevice, c->stream),
                           "copy series"))
      return -1;
    dser = buf;
  }
  std::vector<long long> st;
  // `st` must outlive its copy: the call ends with a stream synchronisation,
  // and an early return waits for the stream (this guard, declared after
  // `st`, is destroyed before it)
  struct WaitOnReturn {
    hipStream_t s;
    bool armed = true;
    ~WaitOnReturn() {
      if (armed) hipStreamSynchronize(s);
    }
  } wait_on_return{c->stream};
  mfh::BlockStarts bs;
  if (!block_starts(c, start, nseries, nblocks, bs, st)) return -1;
  if (mfh::launch_series_forward(c, dser, nseries, bs, nblocks, spec) < 0) return -1;
  long long refined_total = 0;
  std::vector<int> refined;
  // consecutive blocks with equal windows share one launch
  for (int b0 = 0; b0 < nblocks;) {
    int g = 1;
    while (b0 + g < nblocks && win_start[b0 + g] == win_start[b0] && win_end[b0 + g] == win_end[b0]) g++;
    const size_t ws = win_start[b0], we = std::min(win_end[b0], n);
    int64_t *oi = didx + (size_t)b0 * nt * nb;
    cf *ov = dval + (size_t)b0 * nt * nb;
    if (hier) {
      refined.push_back(0);
      // survivor counts are read from the device list counter per group
      int cnt = 0;
      if (mfh::hier_run(c, 0, g, t0, nt, binsize, threshold, ws, we, coarse_thr, oi, ov, &cnt,
                        spec + (size_t)b0 * n) < 0 ||
          !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return -1;
      refined_total += cnt;
    } else {
      mfh::FlatRun r;
      r.data = spec + (size_t)b0 * n;
      r.tmpl = c->tmpl + (size_t)t0 * n;
      r.data_stride = (long long)n;
      r.tmpl_stride = (long long)n;
      r.nd = g;
      r.nt = nt;
      r.binsize = binsize;
      r.threshold = threshold;
      r.ws = ws;
      r.we = we;
      r.out_idx = oi;
      r.out_val = ov;
      if (mfh::launch_flat(c, n, r) < 0) return -1;
    }
    b0 += g;
  }
  const long total = mfh::read_peaks(c, pb, cells, nb, idx, val_c64, counts, rec);  // waits for the stream
  if (total < 0) return -1;
  wait_on_return.armed = false;
  c->prof_collect();
  if (n_refined) *n_refined = refined_total;
  return total;
}

}  // namespace

extern "C" long mfh_run_series(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                               const size_t *start, const size_t *win_start, const size_t *win_end, int nblocks,
                               int t0, int nt, size_t binsize, float threshold, int64_t *idx, void *val_c64,
                               int32_t *counts) {
  if (!ctx) return fail("null context");
  return run_series(reinterpret_cast<Ctx *>(ctx), false, series_c64, nseries, series_on_device, start, win_start,
                    win_end, nblocks, t0, nt, binsize, threshold, 0.f, idx, val_c64, nullptr, counts, nullptr);
}

extern "C" long mfh_run_series_rec(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                                   const size_t *start, const size_t *win_start, const size_t *win_end,
                                   int nblocks, int t0, int nt, size_t binsize, float threshold, void *rec,
                                   int32_t *counts) {
  if (!ctx) return fail("null context");
  if (!rec) return fail("null argument");
  return run_series(reinterpret_cast<Ctx *>(ctx), false, series_c64, nseries, series_on_device, start, win_start,
                    win_end, nblocks, t0, nt, binsize, threshold, 0.f, nullptr, nullptr, rec, counts, nullptr);
}

extern "C" long mfh_hier_run_series(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                                    const size_t *start, const size_t *win_start, const size_t *win_end,
                                    int nblocks, int t0, int nt, size_t binsize, float threshold, float coarse_thr,
                                    int64_t *idx, void *val_c64, int32_t *counts, int64_t *n_refined) {
  if (!ctx) return fail("null context");
  return run_series(reinterpret_cast<Ctx *>(ctx), true, series_c64, nseries, series_on_device, start, win_start,
                    win_end, nblocks, t0, nt, binsize, threshold, coarse_thr, idx, val_c64, nullptr, counts,
                    n_refined);
}

extern "C" long mfh_hier_run_series_rec(mfh_ctx *ctx, const void *series_c64, size_t nseries,
                                        int series_on_device, const size_t *start, const size_t *win_start,
                                        const size_t *win_end, int nblocks, int t0, int nt, size_t binsize,
                                        float threshold, float coarse_thr, void *rec, int32_t *counts,
                                        int64_t *n_refined) {
  if (!ctx) return fail("null context");
  if (!rec) return fail("null argument");
  return run_series(reinterpret_cast<Ctx *>(ctx), true, series_c64, nseries, series_on_device, start, win_start,
                    win_end, nblocks, t0, nt, binsize, threshold, coarse_thr, nullptr, nullptr, rec, counts,
                    n_refined);
}

// Block spectra into data rows: the forward half of run_series on its own.
extern "C" int mfh_series_forward(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                                  const size_t *start, int nblocks, int row0) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (nblocks < 1) return fail("series_forward needs at least one block");
  if (!series_c64 || !start) return fail("null argument");
  if (row0 < 0 || row0 + nblocks > c->nd) return fail("data rows [%d, %d) out of bounds", row0, row0 + nblocks);
  hipSetDevice(c->dev);
  const cf *dser;
  if (series_on_device) {
    dser = (const cf *)series_c64;
  } else {
    cf *buf = (cf *)c->scratch("series_in", (nseries ? nseries : 1) * sizeof(cf));
    if (!buf) return -1;
    if (nseries && !hip_ok(hipMemcpyAsync(buf, series_c64, nseries * sizeof(cf), hipMemcpyHostToDevice, c->stream),
                           "copy series"))
      return -1;
    dser = buf;
  }
  std::vector<long long> st;
  mfh::BlockStarts bs;
  // every return below waits for the stream first: `st` must outlive its copy
  const bool ok = block_starts(c, start, nseries, nblocks, bs, st) && c->own_bank(false) >= 0 &&
                  mfh::launch_series_forward(c, dser, nseries, bs, nblocks, c->data + (size_t)row0 * c->n) >= 0;
  const bool synced = hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return ok && synced ? 0 : -1;
}
