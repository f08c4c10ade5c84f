// Host launcher for the fused flat kernel: per-length instantiation, LDS
// sizing, window splitting and the data-tile choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mfh_flat.h"
#include "mfh_internal.h"

namespace mfh {

namespace {

// LDS a workgroup may claim on gfx950 (160 KiB per CU).
constexpr size_t kLdsLimit = 160 * 1024;

template <int N>
size_t flat_lds_bytes(size_t nb) {
  using S = FlatShape<N>;
  return (size_t)S::SLOTS * S::XSTRIDE * sizeof(cf) + 2 * (size_t)S::SLOTS * nb * sizeof(uint64_t);
}

template <int N>
size_t flat_max_bins() {
  using S = FlatShape<N>;
  const size_t fixed = (size_t)S::SLOTS * S::XSTRIDE * sizeof(cf);
  return (kLdsLimit - fixed) / (2 * (size_t)S::SLOTS * sizeof(uint64_t));
}

template <int N, int MODE, bool LIST>
int launch_mode(Ctx *c, const FlatArgs &a, dim3 grid, size_t lds) {
  using S = FlatShape<N>;
  const void *fn = (const void *)flat_kernel<N, MODE, LIST>;
  if (!lds_limit(fn, c->dev, lds)) return -1;
  hipLaunchKernelGGL((flat_kernel<N, MODE, LIST>), grid, dim3(S::WG), lds, c->stream, a);
  return hip_ok(hipGetLastError(), "flat_kernel launch") ? 0 : -1;
}

// The kernel arguments of a launch over bins [bin0, bin0 + nb) of r (the
// lag window [ws, we) of those bins; nb_total bins per output row).
template <int N>
int fill_args(Ctx *c, const FlatRun &r, size_t ws, size_t we, size_t nb_total, size_t bin0, size_t nb, FlatArgs &a) {
  a.data = r.data;
  a.tmpl = r.tmpl;
  a.twiddle = c->twiddles(N);
  if (!a.twiddle) return -1;
  a.data_stride = r.data_stride;
  a.tmpl_stride = r.tmpl_stride;
  a.nd = r.nd;
  a.nt = r.nt;
  a.ws = (uint32_t)ws;
  a.we = (uint32_t)we;
  a.binsize = (uint32_t)r.binsize;
  a.nb = (uint32_t)nb;
  a.ostride = (uint32_t)nb_total;
  const bool pow2 = (r.binsize & (r.binsize - 1)) == 0;
  a.bshift = pow2 ? ilog2((int)r.binsize) : -1;
  a.inv_bs = 1.0f / (float)r.binsize;
  const float thr = r.threshold;
  a.seed = thr > 0.f ? (float)((double)thr * (double)thr) : 0.f;
  // Output rows are [nd][nt][nb_total]; a window piece writes bins
  // [bin0, bin0 + nb) of each row.
  a.out_idx = r.out_idx + bin0;
  a.out_val = r.out_val + bin0;
  a.list = r.list;
  a.list_count = r.list_count;
  a.list_ooff = r.list_ooff;
  a.zero = bin0 == 0 ? r.zero : nullptr;
  a.hint = bin0 == 0 ? r.hint : nullptr;
  int dpw = c->tune_dpw;
  if (dpw <= 0) dpw = std::min(r.nd, 4);
  a.dpw = dpw;
  return 0;
}

template <int N>
int launch_n(Ctx *c, const FlatRun &r, size_t ws, size_t we, size_t nb_total, size_t bin0, size_t nb) {
  using S = FlatShape<N>;
  FlatArgs a;
  if (fill_args<N>(c, r, ws, we, nb_total, bin0, nb, a) < 0) return -1;
  const bool list = r.list != nullptr;
  const int dpw = a.dpw;
  dim3 grid;
  if (list) {
    // enough workgroups to fill the device; each drains the list grid-stride.
    // With a hint (the previous call's list length), about two pairs per
    // workgroup and at least 512 workgroups: an emptier list leaves most of
    // a full grid with nothing to do, and each idle workgroup still costs
    // its dispatch (any grid drains any list, so a stale hint is harmless)
    int want = (r.list_cap + S::SLOTS - 1) / S::SLOTS;
    if (r.hint) {
      const int last = *(volatile int *)r.hint;
      want = std::min(want, std::max(512, (std::max(last, 0) / S::SLOTS + 1) / 2));
    }
    grid = dim3((unsigned)std::max(1, std::min(want, c->refine_grid_cap)), 1);
  } else {
    grid = dim3((unsigned)((r.nt + S::SLOTS - 1) / S::SLOTS), (unsigned)((r.nd + dpw - 1) / dpw));
    if (grid.y > 65535) return fail("too many data rows per launch (%d)", r.nd);
  }
  const size_t lds = flat_lds_bytes<N>(nb);
  int mode = kGeneral;
  if (nb == 1) mode = kOneBin;
  else if (ws % 64 == 0 && (we % 64 == 0) && r.binsize % 64 == 0) mode = kAligned;
  Ctx::ProfScope ps(c, list ? r.list_prof_name : r.prof_name);
  if (list) {
    if (mode == kOneBin) return launch_mode<N, kOneBin, true>(c, a, grid, lds);
    if (mode == kAligned) return launch_mode<N, kAligned, true>(c, a, grid, lds);
    return launch_mode<N, kGeneral, true>(c, a, grid, lds);
  }
  if (mode == kOneBin) return launch_mode<N, kOneBin, false>(c, a, grid, lds);
  if (mode == kAligned) return launch_mode<N, kAligned, false>(c, a, grid, lds);
  return launch_mode<N, kGeneral, false>(c, a, grid, lds);
}

template <int N>
int launch_len(Ctx *c, const FlatRun &r) {
  const size_t nb = (r.we - r.ws + r.binsize - 1) / r.binsize;
  const size_t maxb = flat_max_bins<N>();
  if (nb <= maxb) return launch_n<N>(c, r, r.ws, r.we, nb, 0, nb);
  // Too many bins for one LDS bin table: split the window into pieces of
  // whole bins.  Each piece recomputes the transform; bins are independent.
  for (size_t b0 = 0; b0 < nb; b0 += maxb) {
    const size_t cnt = std::min(maxb, nb - b0);
    const size_t ws = r.ws + b0 * r.binsize;
    const size_t we = std::min(r.we, ws + cnt * r.binsize);
    if (launch_n<N>(c, r, ws, we, nb, b0, cnt) < 0) return -1;
  }
  return 0;
}

// Rows r = 1, 4 of exp(+2 pi i r k / (16 ns)), k < ns, at [row][k],
// generated in FP64.
__global__ void twiddle_pass_kernel(cf *out, int ns) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns * kTwiddleRows) {
    const int r = i < ns ? 1 : 4, k = i % ns;
    double s, c;
    sincospi(2.0 * (double)r * (double)k / (double)(16 * ns), &s, &c);
    out[i] = cf{(float)c, (float)s};
  }
}

template <int N>
int fill_twiddles(Ctx *c, cf *tw) {
  using L = TwiddleLayout<N>;
  using S = FlatShape<N>;
  int ns = L::NS0;
  for (int p = 0; p < S::NPASS16; p++, ns *= 16) {
    if (ns == 1) continue;
    const int cnt = ns * kTwiddleRows;
    hipLaunchKernelGGL(twiddle_pass_kernel, dim3((cnt + 255) / 256), dim3(256), 0, c->stream,
                       tw + L::offset(p), ns);
  }
  return hip_ok(hipGetLastError(), "twiddle launch") ? 0 : -1;
}

}  // namespace

size_t flat_twiddle_count(size_t n) {
  switch (n) {
    case 64: return TwiddleLayout<64>::total();
    case 128: return TwiddleLayout<128>::total();
    case 256: return TwiddleLayout<256>::total();
    case 512: return TwiddleLayout<512>::total();
    case 1024: return TwiddleLayout<1024>::total();
    case 2048: return TwiddleLayout<2048>::total();
    case 4096: return TwiddleLayout<4096>::total();
    case 8192: return TwiddleLayout<8192>::total();
    case 16384: return TwiddleLayout<16384>::total();
    default: return 0;
  }
}

int flat_fill_twiddles(Ctx *c, size_t n, cf *tw) {
  switch (n) {
    case 64: return fill_twiddles<64>(c, tw);
    case 128: return fill_twiddles<128>(c, tw);
    case 256: return fill_twiddles<256>(c, tw);
    case 512: return fill_twiddles<512>(c, tw);
    case 1024: return fill_twiddles<1024>(c, tw);
    case 2048: return fill_twiddles<2048>(c, tw);
    case 4096: return fill_twiddles<4096>(c, tw);
    case 8192: return fill_twiddles<8192>(c, tw);
    case 16384: return fill_twiddles<16384>(c, tw);
    default: return fail("transform length %zu is not supported", n);
  }
}

int flat_fill_args(Ctx *c, size_t n, const FlatRun &r, FlatArgs &a) {
  if (n != 4096) return fail("flat arguments are built here for n = 4096 only (got %zu)", n);
  if (r.binsize < 1) return fail("binsize must be >= 1");
  const size_t we = std::min(r.we, n);
  if (r.ws >= we) return fail("empty window (%zu, %zu)", r.ws, r.we);
  const size_t nb = (we - r.ws + r.binsize - 1) / r.binsize;
  if (nb > flat_max_bins<4096>()) return fail("too many bins for one launch (%zu)", nb);
  return fill_args<4096>(c, r, r.ws, we, nb, 0, nb, a);
}

bool flat_supported(size_t n) {
  switch (n) {
    case 64: case 128: case 256: case 512: case 1024: case 2048: case 4096: case 8192: case 16384:
    case 32768: case 65536: case 131072: case 262144: case 524288: case 1048576:
      return true;
    default:
      return false;
  }
}

int launch_flat(Ctx *c, size_t n, const FlatRun &r) {
  if (r.binsize < 1) return fail("binsize must be >= 1");
  size_t we = std::min(r.we, n);
  if (r.ws >= we) return fail("empty window (%zu, %zu)", r.ws, r.we);
  FlatRun q = r;
  q.we = we;
  hipSetDevice(c->dev);
  switch (n) {
    case 64: return launch_len<64>(c, q);
    case 128: return launch_len<128>(c, q);
    case 256: return launch_len<256>(c, q);
    case 512: return launch_len<512>(c, q);
    case 1024: return launch_len<1024>(c, q);
    case 2048: return launch_len<2048>(c, q);
    case 4096: return launch_len<4096>(c, q);
    case 8192: return launch_len<8192>(c, q);
    case 16384: return launch_len<16384>(c, q);
    case 32768:
    case 65536:
    case 131072:
    case 262144:
    case 524288:
    case 1048576: return launch_big(c, n, q);
    default: return fail("transform length %zu is not supported", n);
  }
}

}  // namespace mfh
