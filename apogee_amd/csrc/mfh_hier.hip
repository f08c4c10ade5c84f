// Hierarchical matched filter on the device: coarse band screen, gate with
// survivor compaction, and a list-driven refine -- one stream, no host round
// trip between the stages.
//
// Reference behaviour (paths relative to the reference checkout):
//   coarse templates T[:m]/sqrt(f), f in FP64     src/hmf.c:118-156,
//                                                 python/matchedfilter/__init__.py:1284-1316
//   coarse window, gate, refine, dismissal -1/0   src/hmf.c:222-328
//   GPU precedent: coarse pass, compactPairs,
//   indirect refine                               src/gpu/tierb.slang:986-1034,
//                                                 python/matchedfilter/_vkcompute.py:650-850
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "mfh_dft.h"
#include "mfh_flat.h"
#include "mfh_internal.h"
#include "mfhip.h"

namespace mfh {

namespace {

// One workgroup per template: FP64 band fraction and the scaled coarse
// template ct[k] = (float)(T[k] * scale), k < band, scale = f > 0 ? 1/sqrt(f)
// : 0, with f = sum_{k<band}|T|^2 / sum |T|^2, or the bank-wide `ref_f` when
// it is >= 0 (a reference profile was set).
__global__ void __launch_bounds__(256) coarse_template_kernel(const cf *tmpl, long long n, int band, double ref_f,
                                                              cf *ct) {
  const int t = blockIdx.x;
  const cf *T = tmpl + (long long)t * n;
  __shared__ double part[2][256];
  double f = ref_f;
  if (ref_f < 0.0) {
    double low = 0.0, tot = 0.0;
    for (long long k = threadIdx.x; k < n; k += 256) {
      const double re = T[k].x, im = T[k].y;
      const double p = re * re + im * im;
      tot += p;
      if (k < band) low += p;
    }
    part[0][threadIdx.x] = low;
    part[1][threadIdx.x] = tot;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        part[0][threadIdx.x] += part[0][threadIdx.x + s];
        part[1][threadIdx.x] += part[1][threadIdx.x + s];
      }
      __syncthreads();
    }
    f = part[1][0] > 0.0 ? part[0][0] / part[1][0] : 0.0;
  }
  const double scale = f > 0.0 ? 1.0 / sqrt(f) : 0.0;
  for (int k = threadIdx.x; k < band; k += 256)
    ct[(long long)t * band + k] = cf{(float)((double)T[k].x * scale), (float)((double)T[k].y * scale)};
}

// Gate one pair on its coarse peak: fire iff a peak was reported (|z|^2 above
// thr^2) and sqrt(|z|^2) >= thr, exactly the CPU engine's two tests.  Fired
// pairs are appended to the refine list; the rest get -1 / 0 in every bin.
// One atomic per wave: lanes that fire take consecutive slots after the
// wave's base, found from the ballot.
__global__ void gate_kernel(const int64_t *cidx, const cf *cval, int nd, int nt, float thr, int nb,
                            int64_t *out_idx, cf *out_val, int2 *list, int *count, int *other) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) *other = 0;  // the next call's counter
  const bool inside = p < (long long)nd * nt;
  bool fire = false;
  if (inside && cidx[p] >= 0) {
    const cf z = cval[p];
    const float m2 = __builtin_fmaf(z.x, z.x, z.y * z.y);
    fire = sqrtf(m2) >= thr;
  }
  const unsigned long long ballot = __ballot(fire);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (ballot) {
    const int leader = __ffsll((long long)ballot) - 1;
    if (lane == leader) base = atomicAdd(count, __popcll(ballot));
    base = __shfl(base, leader, 64);
  }
  if (!inside) return;
  if (fire) {
    const int slot = base + __popcll(ballot & ((1ull << lane) - 1ull));
    list[slot] = int2{(int)(p / nt), (int)(p % nt)};
  } else {
    for (int b = 0; b < nb; b++) {
      out_idx[p * nb + b] = -1;
      out_val[p * nb + b] = cf{0.f, 0.f};
    }
  }
}

// Gate and refine in one launch after the FP32 screen (refine length FN, one
// transform per 256-thread workgroup): each workgroup takes a tile of
// kGateRows data rows x kGateTemplates templates, decides their gates in
// parallel by gate_kernel's rule from the screen's results, writes -1 / 0
// into every bin of the dismissed pairs, clears the next call's counter,
// runs the flat kernel's per-pair path on each survivor in turn, and adds
// its survivors to this call's counter (one atomic per workgroup that has
// any).  Replaces gate_kernel, the survivor list and the list-driven refine
// launch.  A dense gate gives every workgroup at most 16 refines, whose rows
// and templates are each read four times in a row.
constexpr int kGateRows = 4, kGateTemplates = 4;
template <int FN, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
gate_refine_kernel(const int64_t *cidx, const cf *cval, int nd, int nt, float thr, int *count, int *other,
                   FlatArgs fa) {
  using S = FlatShape<FN>;
  static_assert(S::WG == 256 && S::SLOTS == 1 && S::TPT == 256, "one transform per 256-thread workgroup");
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int2 surv[kGateRows * kGateTemplates];
  __shared__ int ns_sh;
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *other = 0;  // the next call's counter
  const int nb = (int)fa.nb;
  if (threadIdx.x < 64) {
    const int d = blockIdx.y * kGateRows + lane / kGateTemplates;
    const int tt = blockIdx.x * kGateTemplates + lane % kGateTemplates;
    const bool inside = lane < kGateRows * kGateTemplates && d < nd && tt < nt;
    const long long p = (long long)d * nt + tt;
    bool fire = false;
    if (inside && cidx[p] >= 0) {
      const cf z = cval[p];
      const float m2 = __builtin_fmaf(z.x, z.x, z.y * z.y);
      fire = sqrtf(m2) >= thr;
    }
    const unsigned long long ballot = __ballot(fire);
    if (fire) surv[__popcll(ballot & ((1ull << lane) - 1ull))] = int2{d, tt};
    if (inside && !fire) {
      for (int b = 0; b < nb; b++) {
        fa.out_idx[p * fa.ostride + b] = -1;
        fa.out_val[p * fa.ostride + b] = cf{0.f, 0.f};
      }
    }
    if (lane == 0) ns_sh = __popcll(ballot);
  }
  __syncthreads();
  const int ns = ns_sh;
  if (ns == 0) return;  // workgroup-uniform

  cf *xch = reinterpret_cast<cf *>(smem);
  uint64_t *bins_all = reinterpret_cast<uint64_t *>(xch + S::XSTRIDE);
  cf wfirst[15];
  first_twiddles<FN>(wfirst, (int)threadIdx.x, fa.twiddle);
  for (int i = threadIdx.x; i < 2 * nb; i += 256) bins_all[i] = 0ull;
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
    select_and_write<FN, MODE, false>(fa, nb, v, t, lane, bins_all + (i & 1) * nb, bins_all + ((i + 1) & 1) * nb,
                                      true, pr.x, pr.y, 0, 0);
  }
  // the count last: issued at the start, every workgroup's atomic on the one
  // counter would hold its refines back until it completed
  if (threadIdx.x == 0) atomicAdd(count, ns);
}

// Clears `n` ints (n <= 64), one workgroup: a kernel rather than a memset so
// that a captured sequence holds no memset node.
__global__ void zero_ints_kernel(int *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

}  // namespace

int hier_prepare(Ctx *c) {
  if (!c->coarse_dirty) return 0;
  const int band = c->coarse_band;
  cf *ct = (cf *)c->scratch("coarse_tmpl", (size_t)c->nt * band * sizeof(cf));
  if (!ct) return -1;
  hipLaunchKernelGGL(coarse_template_kernel, dim3(c->nt), dim3(256), 0, c->stream, c->tmpl, (long long)c->n, band,
                     c->coarse_ref_f, ct);
  if (!hip_ok(hipGetLastError(), "coarse_template_kernel")) return -1;
  c->coarse_tmpl = ct;
  if (coarse_mfma_prepare(c) < 0) return -1;
  c->coarse_dirty = false;
  return 0;
}

// The two refine-list counters.  Consecutive calls alternate between them
// (c->hier_parity): a call's coarse stage appends to its own counter, which
// the previous call cleared, and clears the other one for the next call --
// stream order puts that after every read of it -- so no launch of its own
// clears a counter.  Both start at zero when first allocated (always on an
// eager call: a graph is captured only on a repeated call).
int *hier_counter(Ctx *c, int parity) {
  auto it = c->bufs.find("refine_counters");
  const bool fresh = it == c->bufs.end() || !it->second.p;
  int *cnt = (int *)c->scratch("refine_counters", 2 * sizeof(int));
  if (!cnt) return nullptr;
  if (fresh) {
    if (c->capturing) {
      fail("refine counters allocated during graph capture");
      return nullptr;
    }
    if (!hip_ok(hipMemsetAsync(cnt, 0, 2 * sizeof(int), c->stream), "hipMemsetAsync")) return nullptr;
  }
  return cnt + (parity & 1);
}

// The refine's launch description (everything but the list fields): the
// run's data rows and templates, its window, bins and threshold, its outputs.
static FlatRun refine_base(Ctx *c, const cf *drows, int t0, int nd, int nt, size_t binsize, float threshold,
                           size_t ws, size_t we, int64_t *idx, cf *val) {
  FlatRun rr;
  rr.data = drows;
  rr.tmpl = c->tmpl + (size_t)t0 * c->n;
  rr.data_stride = (long long)c->n;
  rr.tmpl_stride = (long long)c->n;
  rr.nd = nd;
  rr.nt = nt;
  rr.binsize = binsize;
  rr.threshold = threshold;
  rr.ws = ws;
  rr.we = we;
  rr.out_idx = idx;
  rr.out_val = val;
  return rr;
}

// After the FP32 screen: gate and refine in one launch (tuning key
// coarse_fuse), 4 x 4 pairs per workgroup, at n = 4096 with one bin or with the
// window and bins on 64-lag boundaries (the general bin form of the kernel
// would spill registers; it keeps the separate launches).
static bool gate_refine_fits(const Ctx *c, size_t n, size_t nb, size_t ws, size_t we, size_t binsize) {
  return c->coarse_fuse && n == 4096 && (nb == 1 || (ws % 64 == 0 && we % 64 == 0 && binsize % 64 == 0));
}

static int launch_gate_refine(Ctx *c, const int64_t *cidx, const cf *cval, int nd, int nt, float thr, int *count,
                              int *other, const FlatRun &rr) {
  constexpr int FN = 4096;
  FlatArgs fa;
  if (flat_fill_args(c, FN, rr, fa) < 0) return -1;
  const int nb = (int)fa.nb;
  if (!gate_refine_fits(c, FN, nb, fa.ws, fa.we, fa.binsize))
    return fail("the one-launch gate and refine does not apply here");
  using Kern = decltype(&gate_refine_kernel<FN, kOneBin>);
  const Kern kern = nb == 1 ? gate_refine_kernel<FN, kOneBin> : gate_refine_kernel<FN, kAligned>;
  const dim3 grid((unsigned)((nt + kGateTemplates - 1) / kGateTemplates), (unsigned)((nd + kGateRows - 1) / kGateRows));
  if (grid.y > 65535) return fail("too many data rows per launch (%d)", nd);
  const size_t smem = (size_t)FlatShape<FN>::XSTRIDE * sizeof(cf) + 2 * (size_t)nb * sizeof(uint64_t);
  if (!lds_limit((const void *)kern, c->dev, smem)) return -1;
  Ctx::ProfScope ps(c, "refine");
  hipLaunchKernelGGL(kern, grid, dim3(256), smem, c->stream, cidx, cval, nd, nt, thr, count, other, fa);
  return hip_ok(hipGetLastError(), "gate_refine_kernel") ? 0 : -1;
}

int hier_run(Ctx *c, int d0, int nd, int t0, int nt, size_t binsize, float threshold, size_t ws, size_t we,
             float coarse_thr, int64_t *idx, cf *val, int *count_out, const cf *data_rows, const HierScratch *hs) {
  if (c->coarse_band <= 0) return fail("no coarse band: call mfh_hier_set_band first");
  if (!(coarse_thr >= 0.f) || !std::isfinite(coarse_thr))
    return fail("coarse threshold must be finite and nonnegative");
  if (!data_rows && (d0 < 0 || d0 + nd > c->nd)) return fail("data/templates sub-range out of bounds");
  if (t0 < 0 || nd < 1 || nt < 1 || t0 + nt > c->nt) return fail("data/templates sub-range out of bounds");
  if (binsize < 1) return fail("binsize must be >= 1");
  if (we > c->n) we = c->n;
  if (ws >= we) return fail("empty window (%zu, %zu)", ws, we);
  hipSetDevice(c->dev);
  if (hier_prepare(c) < 0) return -1;
  const size_t n = c->n, m = (size_t)c->coarse_band;
  const size_t nb = (we - ws + binsize - 1) / binsize;
  const size_t pairs = (size_t)nd * nt;
  // coarse window on the decimated grid (src/hmf.c:233-239)
  const size_t R = n / m;
  size_t cs = ws / R, ce = std::min(m, (we + R - 1) / R);
  if (cs > 0) cs--;
  int64_t *cidx;
  cf *cval;
  int2 *list;
  int *count, *other;
  if (hs) {
    // caller-supplied buffers: this call's counter starts at zero here, and
    // the screen's clearing of "the next call's counter" lands on a spare word
    cidx = hs->cidx;
    cval = hs->cval;
    list = hs->list;
    count = hs->count;
    other = hs->spare;
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, c->stream, count, 1);
    if (!hip_ok(hipGetLastError(), "zero_ints_kernel")) return -1;
  } else {
    cidx = (int64_t *)c->scratch("coarse_idx", pairs * sizeof(int64_t));
    cval = (cf *)c->scratch("coarse_val", pairs * sizeof(cf));
    list = (int2 *)c->scratch("refine_list", pairs * sizeof(int2));
    count = hier_counter(c, c->hier_parity);
    other = hier_counter(c, c->hier_parity ^ 1);
  }
  if (!cidx || !cval || !list || !count || !other) return -1;

  const cf *drows = data_rows ? data_rows : c->data + (size_t)d0 * n;
  const bool mfma = c->coarse_prec > 0 && m == 512;
  if (mfma) {
    // the matrix-core screen gates each pair itself and appends survivors
    // to this call's counter, which the previous call's screen cleared; it
    // clears the other one for the next call
    CoarseGate g;
    if (coarse_thr > 0.f) {
      const double w = (double)coarse_thr / (1.0 + (double)c->coarse_delta);
      g.thr2 = std::nextafter((float)(w * w), 0.f);
    }
    g.nb = (int)nb;
    g.idx = idx;
    g.val = val;
    g.list = list;
    g.count = count;
    g.zero = other;
    if (coarse_refine_fits(c, n, nb)) {
      // one launch: each workgroup of the screen refines its own survivors
      // (the count still goes to this call's counter; no list, no second
      // launch)
      const FlatRun rr = refine_base(c, drows, t0, nd, nt, binsize, threshold, ws, we, idx, val);
      if (launch_coarse_refine(c, drows, (long long)n, t0, nd, nt, (int)cs, (int)ce, c->coarse_prec == 2, g, rr) < 0)
        return -1;
      if (!c->capturing && !hs) c->hier_parity ^= 1;
      if (count_out &&
          !hip_ok(hipMemcpyAsync(count_out, count, sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy count"))
        return -1;
      return 0;
    }
    if (launch_coarse_mfma(c, drows, (long long)n, t0, nd, nt, (int)cs, (int)ce, c->coarse_prec == 2, nullptr, &g) <
        0)
      return -1;
  }
  FlatRun cr;
  cr.data = drows;
  cr.tmpl = c->coarse_tmpl + (size_t)t0 * m;
  cr.data_stride = (long long)n;
  cr.tmpl_stride = (long long)m;
  cr.nd = nd;
  cr.nt = nt;
  cr.binsize = ce - cs;
  cr.threshold = coarse_thr;
  cr.ws = cs;
  cr.we = ce;
  cr.out_idx = cidx;
  cr.out_val = cval;
  cr.prof_name = "coarse";
  cr.zero = count;  // the coarse launch clears the survivor counter
  if (!mfma) {
    if (launch_flat(c, m, cr) < 0) return -1;
    if (gate_refine_fits(c, n, nb, ws, we, binsize)) {
      // the gate and the refine in one launch; the screen above cleared this
      // call's counter, the launch clears the next call's
      const FlatRun rr = refine_base(c, drows, t0, nd, nt, binsize, threshold, ws, we, idx, val);
      if (launch_gate_refine(c, cidx, cval, nd, nt, coarse_thr, count, other, rr) < 0) return -1;
      if (!c->capturing && !hs) c->hier_parity ^= 1;
      if (count_out &&
          !hip_ok(hipMemcpyAsync(count_out, count, sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy count"))
        return -1;
      return 0;
    }
    Ctx::ProfScope ps(c, "gate");
    hipLaunchKernelGGL(gate_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, c->stream, cidx, cval, nd,
                       nt, coarse_thr, (int)nb, idx, val, list, count, other);
    if (!hip_ok(hipGetLastError(), "gate_kernel")) return -1;
  }
  // the screen has cleared the other counter: the next call uses it (a
  // capture flips the parity when its graph runs; caller-supplied counters
  // leave the context's pair untouched)
  if (!c->capturing && !hs) c->hier_parity ^= 1;

  FlatRun rr = refine_base(c, drows, t0, nd, nt, binsize, threshold, ws, we, idx, val);
  rr.list = list;
  rr.list_count = count;
  rr.list_cap = (int)pairs;
  if (!hs && !c->no_alloc && !c->capturing && !c->refine_hint) {
    // one pinned word per context, mapped for the device (default flags);
    // starts at the full list so the first call sizes as before
    if (!hip_ok(hipHostMalloc((void **)&c->refine_hint, sizeof(int), hipHostMallocDefault), "hipHostMalloc"))
      return -1;
    *c->refine_hint = (int)pairs;
  }
  rr.hint = hs ? nullptr : c->refine_hint;
  if (launch_flat(c, n, rr) < 0) return -1;
  if (count_out &&
      !hip_ok(hipMemcpyAsync(count_out, count, sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy count"))
    return -1;
  return 0;
}

}  // namespace mfh

using mfh::Ctx;
using mfh::fail;
using mfh::hip_ok;

extern "C" int mfh_hier_set_band(mfh_ctx *ctx, size_t band, double ref_fraction) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (band < 64 || band >= c->n || (band & (band - 1)) || !mfh::flat_supported(band))
    return fail("band must be a power of two, >= 64 and < n (got %zu)", band);
  if (ref_fraction > 1.0 || !(ref_fraction == ref_fraction)) return fail("reference band fraction out of range");
  c->coarse_band = (int)band;
  c->coarse_ref_f = ref_fraction < 0.0 ? -1.0 : ref_fraction;
  c->coarse_dirty = true;
  return 0;
}

extern "C" int mfh_hier_set_coarse_templates(mfh_ctx *ctx, size_t band, const void *ct0_c64, int on_device) {
  if (!ctx || !ct0_c64) return fail("null argument");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (band < 64 || band >= c->n || (band & (band - 1)) || !mfh::flat_supported(band))
    return fail("band must be a power of two, >= 64 and < n (got %zu)", band);
  hipSetDevice(c->dev);
  const size_t bytes = (size_t)c->nt * band * sizeof(mfh::cf);
  mfh::cf *ct = (mfh::cf *)c->scratch("coarse_tmpl", bytes);
  if (!ct) return -1;
  if (!hip_ok(hipMemcpyAsync(ct, ct0_c64, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                             c->stream),
              "copy coarse templates"))
    return -1;
  c->coarse_band = (int)band;
  // a later template change rebuilds from the spectra, each template by its
  // own band fraction (the reference's rule when no profile is set)
  c->coarse_ref_f = -1.0;
  c->coarse_tmpl = ct;
  c->coarse_dirty = false;
  if (mfh::coarse_mfma_prepare(c) < 0) return -1;
  return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ? 0 : -1;
}

extern "C" int mfh_hier_set_coarse_precision(mfh_ctx *ctx, int precision, float delta) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (precision < 0 || precision > 2) return fail("precision must be 0 (fp32), 1 (fp16) or 2 (bf16)");
  if (!(delta >= 0.f) || delta >= 1.f) return fail("delta must be in [0, 1)");
  c->coarse_prec = precision;
  c->coarse_delta = delta;
  return 0;
}

extern "C" int mfh_hier_coarse_max(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t ws, size_t we,
                                   int precision, float *host_max) {
  if (!ctx || !host_max) return fail("null argument");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (c->coarse_band <= 0) return fail("no coarse band: call mfh_hier_set_band first");
  if (d0 < 0 || t0 < 0 || nd < 1 || nt < 1 || d0 + nd > c->nd || t0 + nt > c->nt)
    return fail("data/templates sub-range out of bounds");
  if (precision < 0 || precision > 2) return fail("precision must be 0 (fp32), 1 (fp16) or 2 (bf16)");
  if (we > c->n) we = c->n;
  if (ws >= we) return fail("empty window (%zu, %zu)", ws, we);
  hipSetDevice(c->dev);
  if (mfh::hier_prepare(c) < 0) return -1;
  const size_t n = c->n, m = (size_t)c->coarse_band, pairs = (size_t)nd * nt;
  const size_t R = n / m;
  size_t cs = ws / R, ce = std::min(m, (we + R - 1) / R);
  if (cs > 0) cs--;
  float *cmax2 = (float *)c->scratch("coarse_max", pairs * sizeof(float));
  if (!cmax2) return -1;
  std::vector<float> m2(pairs);
  if (precision > 0) {
    if (mfh::launch_coarse_mfma(c, c->data + (size_t)d0 * n, (long long)n, t0, nd, nt,
                                (int)cs, (int)ce, precision == 2, cmax2) < 0 ||
        !hip_ok(hipMemcpyAsync(m2.data(), cmax2, pairs * sizeof(float), hipMemcpyDeviceToHost, c->stream), "copy") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
      return -1;
    for (size_t p = 0; p < pairs; p++) host_max[p] = sqrtf(m2[p]);
    return 0;
  }
  int64_t *cidx = (int64_t *)c->scratch("coarse_idx", pairs * sizeof(int64_t));
  mfh::cf *cval = (mfh::cf *)c->scratch("coarse_val", pairs * sizeof(mfh::cf));
  if (!cidx || !cval) return -1;
  mfh::FlatRun cr;
  cr.data = c->data + (size_t)d0 * n;
  cr.tmpl = c->coarse_tmpl + (size_t)t0 * m;
  cr.data_stride = (long long)n;
  cr.tmpl_stride = (long long)m;
  cr.nd = nd;
  cr.nt = nt;
  cr.binsize = ce - cs;
  cr.threshold = 0.f;
  cr.ws = cs;
  cr.we = ce;
  cr.out_idx = cidx;
  cr.out_val = cval;
  cr.prof_name = "coarse";
  std::vector<mfh::cf> v(pairs);
  if (mfh::launch_flat(c, m, cr) < 0 ||
      !hip_ok(hipMemcpyAsync(v.data(), cval, pairs * sizeof(mfh::cf), hipMemcpyDeviceToHost, c->stream), "copy") ||
      !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
    return -1;
  for (size_t p = 0; p < pairs; p++) host_max[p] = sqrtf(__builtin_fmaf(v[p].x, v[p].x, v[p].y * v[p].y));
  return 0;
}

namespace mfh {

static long long fbits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return (long long)u;
}
static long long dbits(double x) {
  long long u;
  memcpy(&u, &x, 8);
  return u;
}

// hier_run through the context's graph cache: the first call with a set of
// launch parameters runs eagerly, the second identical call is captured into
// a HIP graph, and later identical calls replay it with one launch.
int hier_run_cached(Ctx *c, int d0, int nd, int t0, int nt, size_t binsize, float threshold, size_t ws, size_t we,
                    float coarse_thr, int64_t *idx, cf *val, int *count_out) {
  if (!c->use_graphs || c->prof_on)
    return hier_run(c, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, idx, val, count_out, nullptr);
  hipSetDevice(c->dev);
  if (c->coarse_band > 0 && hier_prepare(c) < 0) return -1;
  std::vector<long long> key = {d0, nd, t0, nt, (long long)binsize, fbits(threshold), (long long)ws, (long long)we,
                                fbits(coarse_thr), (long long)(intptr_t)idx, (long long)(intptr_t)val,
                                c->coarse_band, dbits(c->coarse_ref_f), c->coarse_prec, fbits(c->coarse_delta),
                                c->refine_grid_cap, (long long)(intptr_t)c->coarse_tmpl, c->coarse_tpw, c->coarse_rpw, c->coarse_lds,
                                c->coarse_fuse,
                                (long long)(intptr_t)c->data, (long long)(intptr_t)c->tmpl};
  const int par = c->hier_parity;
  if (key == c->graph_key && c->hier_graph[par]) {
    if (!hip_ok(hipGraphLaunch(c->hier_graph[par], c->stream), "hipGraphLaunch")) return -1;
    c->hier_parity ^= 1;
  } else if (key == c->graph_key || key == c->last_key) {
    // a repeated call: capture the graph of this counter parity
    if (key != c->graph_key) {
      c->drop_hier_graphs();
      c->graph_key = key;
    }
    hipGraph_t g = nullptr;
    if (!hip_ok(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return -1;
    c->capturing = true;
    const int rc = hier_run(c, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, idx, val, nullptr, nullptr);
    c->capturing = false;
    const bool ended = hip_ok(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
    if (rc < 0 || !ended) {
      if (g) hipGraphDestroy(g);
      return -1;
    }
    hipGraphExec_t ex = nullptr;
    const bool inst = hip_ok(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    hipGraphDestroy(g);
    if (!inst) return -1;
    c->hier_graph[par] = ex;
    if (!hip_ok(hipGraphLaunch(ex, c->stream), "hipGraphLaunch")) return -1;
    c->hier_parity ^= 1;
  } else {
    if (hier_run(c, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, idx, val, nullptr, nullptr) < 0)
      return -1;
    c->last_key = key;
  }
  if (count_out) {
    int *count = hier_counter(c, par);
    if (!count || !hip_ok(hipMemcpyAsync(count_out, count, sizeof(int), hipMemcpyDeviceToHost, c->stream),
                          "copy count"))
      return -1;
  }
  return 0;
}

}  // namespace mfh

extern "C" int mfh_hier_peaks_async(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold,
                                    size_t ws, size_t we, float coarse_thr, int64_t *dev_idx, void *dev_val_c64,
                                    int32_t *host_refined) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  return mfh::hier_run_cached(c, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, dev_idx,
                              (mfh::cf *)dev_val_c64, host_refined);
}

static long hier_peaks_host(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold, size_t ws,
                            size_t we, float coarse_thr, int64_t *idx, void *val_c64, void *rec, int32_t *counts,
                            int64_t *n_refined) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!rec && (!idx || !val_c64)) return fail("null output pointer");
  const size_t nb = mfh_nbins(c->n, binsize, ws, we);
  if (binsize < 1) return fail("binsize must be >= 1");
  if (nb == 0) return fail("empty window (%zu, %zu)", ws, we);
  if (nd < 1 || nt < 1) return fail("data/templates sub-range out of bounds");
  const size_t cells = (size_t)nd * (size_t)nt * nb;
  hipSetDevice(c->dev);
  mfh::PeakBuf pb;
  if (!mfh::peak_buffer(c, cells, pb)) return -1;
  int refined = 0;
  if (mfh::hier_run_cached(c, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, pb.idx, pb.val, &refined) < 0)
    return -1;
  const long total = mfh::read_peaks(c, pb, cells, nb, idx, val_c64, counts, rec);
  if (total < 0) return -1;
  c->prof_collect();
  if (n_refined) *n_refined = refined;
  return total;
}

extern "C" long mfh_hier_peaks(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold,
                               size_t ws, size_t we, float coarse_thr, int64_t *idx, void *val_c64, int32_t *counts,
                               int64_t *n_refined) {
  return hier_peaks_host(ctx, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, idx, val_c64, nullptr, counts,
                         n_refined);
}

extern "C" long mfh_hier_peaks_rec(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold,
                                   size_t ws, size_t we, float coarse_thr, void *rec, int32_t *counts,
                                   int64_t *n_refined) {
  if (!rec) return fail("null output pointer");
  return hier_peaks_host(ctx, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, nullptr, nullptr, rec, counts,
                         n_refined);
}

// ---------------------------------------------------------------------------
// Allocation-free hierarchical run.  The per-call buffers (coarse outputs,
// survivor list, counter) live in one file-scope device array, `scratch`, of
// fixed size; each device holds its own copy, whose address is looked up once
// per device with hipGetSymbolAddress (the array's name in host code is not a
// device address).  Every buffer the context itself needs (coarse bank,
// twiddle and MFMA tables) must already exist: with Ctx::no_alloc set, any
// attempt to create one fails the call instead of allocating.
static constexpr size_t kNoAllocMaxPairs = (size_t)1 << 18;
static constexpr size_t kNoAllocHead = 256;  // counter words, one 256-byte line
static constexpr size_t kNoAllocBytes =
    kNoAllocHead + kNoAllocMaxPairs * (sizeof(int64_t) + sizeof(mfh::cf) + sizeof(int2));  // 6 MiB + 256 B

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mutex;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of its copy of `scratch`

// The calling thread's current device must be `dev`.
unsigned char *device_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void *p = nullptr;
  if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return nullptr;
  if (!p) {
    fail("hipGetSymbolAddress returned a null address for the scratch array");
    return nullptr;
  }
  g_scratch_ptr[dev] = (unsigned char *)p;
  return (unsigned char *)p;
}

}  // namespace

extern "C" int mfh_hier_peaks_noalloc(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize,
                                      float threshold, size_t ws, size_t we, float coarse_thr, int64_t *dev_idx,
                                      void *dev_val_c64) {
  if (!ctx) return fail("null context");
  if (!dev_idx || !dev_val_c64) return fail("null output pointer");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (nd < 1 || nt < 1) return fail("data/templates sub-range out of bounds");
  const size_t pairs = (size_t)nd * (size_t)nt;
  if (pairs > kNoAllocMaxPairs)
    return fail("%zu pairs exceed the allocation-free limit of %zu pairs per call", pairs, kNoAllocMaxPairs);
  if (!hip_ok(hipSetDevice(c->dev), "hipSetDevice")) return -1;
  unsigned char *base = device_scratch(c->dev);
  if (!base) return -1;
  mfh::HierScratch hs;
  hs.count = (int *)base;
  hs.spare = (int *)base + 1;
  hs.cidx = (int64_t *)(base + kNoAllocHead);
  hs.cval = (mfh::cf *)(base + kNoAllocHead + pairs * sizeof(int64_t));
  hs.list = (int2 *)(base + kNoAllocHead + pairs * (sizeof(int64_t) + sizeof(mfh::cf)));
  c->no_alloc = true;
  const int rc = mfh::hier_run(c, d0, nd, t0, nt, binsize, threshold, ws, we, coarse_thr, dev_idx,
                               (mfh::cf *)dev_val_c64, nullptr, nullptr, &hs);
  c->no_alloc = false;
  return rc;
}
