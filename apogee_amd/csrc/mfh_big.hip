// Flat filter for n = 32768 and 65536: a four-step split across two kernels.
//
// A length-n transform no longer fits one workgroup's LDS (256/512 KiB), so
// n = N1 x 256 (N1 = 128 or 256) with f = f1 + N1 f2 and k = k2 + 256 k1:
//   Y[f1][k2] = w_n^(f1 k2) sum_f2 P[f1 + N1 f2] w_256^(f2 k2)     kernel A
//   z[k2 + 256 k1] = sum_f1 Y[f1][k2] w_N1^(f1 k1)                 kernel B
// Kernel A forms the product D*conj(T) on load, runs 256-point inverse FFTs
// over 16 columns per workgroup and writes Y k2-major to scratch (each
// 16-lane group writes 128 contiguous bytes).  Kernel B gives each pair one
// workgroup that walks all k2 in groups, runs N1-point inverse FFTs, and keeps
// a per-bin (key, value) table in LDS: keys rise by ds_max_u64, and after
// each group the lane holding the bin's key writes its value, so the table
// always pairs the largest |z|^2 (lowest lag among equals) with its own z.
// Only the final per-bin result leaves the chip.
//
// Reference behaviour: the same selection rule as the flat kernel
// (src/balanced-inl.h:876-969 in the reference checkout); the reference's GPU
// carried these lengths in one workgroup with R = 32/64 points per thread
// (src/gpu/tierb.slang:13-22), which wave64 register budgets do not allow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mfh_flat.h"
#include "mfh_internal.h"

namespace mfh {

namespace {

struct BigArgs {
  const cf *data, *tmpl;
  long long data_stride, tmpl_stride;
  int nd, nt;           // tiled mode: pair p = p0 + blockIdx -> (d, t) = (p % nd, p / nd)
  long long p0;         // first pair (tiled) or list position (list) of this chunk
  const int2 *list;     // list mode
  const int *list_count;
  const long long *list_ooff;  // list mode: output element offset per pair
  cf *scratch;          // [chunk pair][256][N1]
  const cf *tw256, *twN1;  // Stockham tables of the two lengths
  const cf *rootH, *rootL; // w_N1^h (h < N1) and w_n^l (l < 256)
  int nsplit;               // rows workgroups per pair (gridDim.y of the rows kernel)
  uint64_t *pkey;           // nsplit > 1: per-part bin tables [pair][part][nb]
  cf *pval;
  int64_t *out_idx;
  cf *out_val;
  uint32_t ws, we, binsize, nb, ostride;
  int bshift;
  float inv_bs, seed;
};

__device__ __forceinline__ bool big_pair(const BigArgs &a, long long p, int &d, int &t) {
  if (a.list) {
    if (p >= *a.list_count) return false;
    const int2 pr = a.list[p];
    d = pr.x;
    t = pr.y;
    return true;
  }
  if (p >= (long long)a.nd * a.nt) return false;
  // template-major: the nd pairs of one template are neighbours, so its
  // spectrum is re-read from L2 rather than HBM (2.5% at n = 65536, 4 x 128)
  t = (int)(p / a.nd);
  d = (int)(p % a.nd);
  return true;
}

// w_n^m for 0 <= m < n from the two root tables (global, a few KiB, cached).
__device__ __forceinline__ cf big_root(const BigArgs &a, int m) { return cmul(a.rootH[m >> 8], a.rootL[m & 255]); }

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// One (pair, bin) cell of a chunk whose rows launch left per-part tables:
// the largest key over the parts (keys are distinct per lag, so exactly one
// part holds it) and that part's value.  The keys are read first, all
// independent loads, then the one value.
__device__ __forceinline__ void finish_cell(const BigArgs &a, long long np, long long i) {
  const int nb = (int)a.nb;
  if (i >= np * nb) return;
  const long long pl = i / nb;
  const int b = (int)(i % nb);
  int d = 0, tt = 0;
  if (!big_pair(a, a.p0 + pl, d, tt)) return;
  const uint64_t *keys = a.pkey + pl * a.nsplit * (long long)nb + b;
  uint64_t best = 0ull;
  int bw = 0;
#pragma unroll 8
  for (int w = 0; w < a.nsplit; w++) {
    const uint64_t key = keys[(long long)w * nb];
    bw = key > best ? w : bw;
    best = key > best ? key : best;
  }
  const long long o = a.list_ooff ? a.list_ooff[a.p0 + pl] : ((long long)d * a.nt + tt) * a.ostride;
  a.out_idx[o + b] = best ? (int64_t)(~(uint32_t)(best & 0xffffffffull)) : -1;
  a.out_val[o + b] = best ? a.pval[(pl * a.nsplit + bw) * (long long)nb + b] : cf{0.f, 0.f};
}

// Columns: 16 columns f1 of one pair per workgroup; lane = 16 * t + s, so the
// 16 lanes of a row group touch 16 consecutive columns (128-byte accesses).
// The pair's rows and its Y block are wave-uniform, so loads and stores go
// through buffer descriptors: one vector offset per thread and the per-u
// stride as a scalar offset, with no vector address arithmetic per point.
template <int N1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) big_cols_kernel(BigArgs a) {
  constexpr int N = N1 * 256;
  extern __shared__ __align__(16) unsigned char smem[];
  const int s = threadIdx.x & 15;
  const int t = threadIdx.x >> 4;
  cf *xch = reinterpret_cast<cf *>(smem) + s * kColStride;
  // The closing twiddle of output k2 = t + 16 u of column f1 splits as
  // w_n^(f1 k2) = w_n^(f1 t) * w_n^(16 f1 u): the first factor is one per
  // thread, the second one per (column, u), 256 per workgroup, kept in LDS
  // as bt[17 s + u] (rows of 17: the 16 columns a wave reads in one step
  // fall in distinct banks, its 4 row groups read the same words).
  cf *bt = reinterpret_cast<cf *>(smem) + 16 * kColStride;
  const long long pl = blockIdx.y;  // pair within the chunk
  int d = 0, tt = 0;
  const bool valid = big_pair(a, a.p0 + pl, d, tt);  // uniform across the workgroup
  if (!valid) return;                                 // no barrier has run yet
  const int f1 = blockIdx.x * 16 + s;
  bt[17 * s + t] = big_root(a, 16 * f1 * t);  // u = t; 16 f1 t < n
  const cf wt = big_root(a, f1 * t);
  const cf *D = uniform_ptr(a.data + (long long)d * a.data_stride);
  const cf *T = uniform_ptr(a.tmpl + (long long)tt * a.tmpl_stride);
  const __amdgpu_buffer_rsrc_t rd =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(D), (short)0, N * (int)sizeof(cf), 0x00020000);
  const __amdgpu_buffer_rsrc_t rt =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(T), (short)0, N * (int)sizeof(cf), 0x00020000);
  const int voff = (f1 + N1 * t) * (int)sizeof(cf);  // element f1 + N1 (t + 16 u) at voff + u * 16 N1 * 8
  cf v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int so = u * 16 * N1 * (int)sizeof(cf);
    const cf x = __builtin_bit_cast(cf, __builtin_amdgcn_raw_buffer_load_b64(rd, voff, so, 0));
    const cf y = __builtin_bit_cast(cf, __builtin_amdgcn_raw_buffer_load_b64(rt, voff, so, 0));
    v[u] = cmul_conj(x, y);
  }
  inverse_fft<256>(v, t, xch, a.tw256);  // its barriers order the bt stores before the reads
  cf *Y = const_cast<cf *>(uniform_ptr(a.scratch + pl * N));
  const __amdgpu_buffer_rsrc_t ry = __builtin_amdgcn_make_buffer_rsrc(Y, (short)0, N * (int)sizeof(cf), 0x00020000);
  const int yoff = (t * N1 + f1) * (int)sizeof(cf);  // Y[k2][f1], k2 = t + 16 u
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const cf w = cmul(wt, bt[17 * s + u]);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, cmul(v[u], w)), ry, yoff,
                                          u * 16 * N1 * (int)sizeof(cf), 0);
  }
}

// Rows: nsplit workgroups per pair (workgroup = pair * nsplit + part), each
// walking its share of the 256 rows k2, SLOTS rows at a time.  With one part
// the table is the result; with several, each part leaves its (key, value)
// table, and finish_cell keeps the largest key per bin: in the next chunk's
// rows launch, or in big_rows_finish_kernel after the last chunk.
template <int N1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
big_rows_kernel(BigArgs a, BigArgs fin, long long fin_np, int fin_blocks) {
  // the last fin_blocks workgroups finish the previous chunk (its tables
  // were completed by the previous launch): one thread per (pair, bin)
  const int row_blocks = (int)gridDim.x - fin_blocks;
  if ((int)blockIdx.x >= row_blocks) {
    finish_cell(fin, fin_np, (long long)(blockIdx.x - row_blocks) * 256 + threadIdx.x);
    return;
  }
  using S = FlatShape<N1>;
  constexpr int TPT = S::TPT;
  constexpr int SLOTS = S::SLOTS;
  extern __shared__ __align__(16) unsigned char smem[];
  cf *xch_all = reinterpret_cast<cf *>(smem);
  uint64_t *bkey = reinterpret_cast<uint64_t *>(xch_all + SLOTS * S::XSTRIDE);
  cf *bval = reinterpret_cast<cf *>(bkey + a.nb);
  const int slot = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  cf *xch = xch_all + slot * S::XSTRIDE;
  const long long pl = blockIdx.x / a.nsplit;  // pair within the chunk
  const int part = (int)(blockIdx.x % a.nsplit);
  int d = 0, tt = 0;
  if (!big_pair(a, a.p0 + pl, d, tt)) return;  // uniform; before any barrier
  const int nb = (int)a.nb;
  for (int b = threadIdx.x; b < nb; b += 256) bkey[b] = 0ull;
  __syncthreads();
  const cf *Y = a.scratch + pl * (long long)(N1 * 256);
  constexpr int G = 256 / SLOTS;  // row groups of the pair
  const int gbeg = part * G / a.nsplit, gend = (part + 1) * G / a.nsplit;
  const bool whole = a.ws == 0 && a.we == (uint32_t)(N1 * 256);
  const bool full8 = whole && a.bshift == ilog2(N1 * 256 / 8);
  FlatArgs fa;  // for bin_of / make_key
  fa.ws = a.ws;
  fa.we = a.we;
  fa.binsize = a.binsize;
  fa.bshift = a.bshift;
  fa.inv_bs = a.inv_bs;
  // the first twiddled pass's twiddles depend on t only: derived once for
  // all the part's row groups where the registers allow (N1 <= 512)
  constexpr bool HOIST = N1 <= 512;
  cf wf[HOIST ? 15 : 1];
  if constexpr (HOIST) first_twiddles<N1>(wf, t, a.twN1);
  for (int g = gbeg; g < gend; g++) {
    const int k2 = g * SLOTS + slot;
    cf v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = Y[(long long)k2 * N1 + t + TPT * u];
    inverse_fft<N1, HOIST>(v, t, xch, a.twN1, wf);
    // thread t holds z[k2 + 256 (t + TPT u)]; lags rise with u.  A wave none
    // of whose window lags has |z|^2 above the seed raises no key and records
    // no value, so it skips both steps below (wave-uniform; nonnegative
    // floats order like their bits, and a NaN's bits, above all of them,
    // send the wave through the steps, which never select it)
    uint32_t mloc = 0;
    if (whole) {
#pragma unroll
      for (int u = 0; u < 16; u++) mloc = max(mloc, __float_as_uint(__builtin_fmaf(v[u].x, v[u].x, v[u].y * v[u].y)));
    } else {
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const uint32_t k = (uint32_t)(k2 + 256 * (t + TPT * u));
        const uint32_t m = __float_as_uint(__builtin_fmaf(v[u].x, v[u].x, v[u].y * v[u].y));
        mloc = (k >= a.ws && k < a.we) ? max(mloc, m) : mloc;
      }
    }
    const bool hot = wave_umax(mloc) > __float_as_uint(a.seed);
    if (!hot) {
    } else if (full8) {
      // n/8 bins over the whole transform: every wave's 64 lanes put u and
      // u + 1 = 2 rho, 2 rho + 1 in bin rho (their lags span < 256 TPT)
      const int lane = threadIdx.x & 63;
      const int k2w = __builtin_amdgcn_readfirstlane(k2);
      runs8_max_to_bins<256 * TPT>(v, (uint32_t)(k2 + 256 * t), (uint32_t)k2w, lane, a.seed, a.bshift, bkey);
    } else {
    uint64_t acc = 0ull;
    int acc_bin = -1;
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const uint32_t k = (uint32_t)(k2 + 256 * (t + TPT * u));
      if (k >= a.ws && k < a.we) {
        const uint64_t key = make_key(v[u], k, a.seed);
        const int b = (int)bin_of(k, fa);
        if (b != acc_bin) {
          if (acc) atomicMax((unsigned long long *)&bkey[acc_bin], (unsigned long long)acc);
          acc = 0ull;
          acc_bin = b;
        }
        acc = key > acc ? key : acc;
      }
    }
    if (acc) atomicMax((unsigned long long *)&bkey[acc_bin], (unsigned long long)acc);
    }
    __syncthreads();
    // the holder of each bin's key records its value
    if (hot) {
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const uint32_t k = (uint32_t)(k2 + 256 * (t + TPT * u));
        if (k >= a.ws && k < a.we) {
          const uint64_t key = make_key(v[u], k, a.seed);
          if (key) {
            const uint32_t b = bin_of(k, fa);
            if (bkey[b] == key) bval[b] = v[u];
          }
        }
      }
    }
    __syncthreads();  // values settled before the next group raises keys; xch reusable
  }
  if (a.nsplit > 1) {
    const long long po = (pl * a.nsplit + part) * (long long)nb;
    for (int b = threadIdx.x; b < nb; b += 256) {
      const uint64_t key = bkey[b];
      a.pkey[po + b] = key;
      a.pval[po + b] = key ? bval[b] : cf{0.f, 0.f};
    }
    return;
  }
  const long long o = a.list_ooff ? a.list_ooff[a.p0 + pl] : ((long long)d * a.nt + tt) * a.ostride;
  int64_t *oidx = a.out_idx + o;
  cf *oval = a.out_val + o;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const uint64_t key = bkey[b];
    if (key) {
      oidx[b] = (int64_t)(~(uint32_t)(key & 0xffffffffull));
      oval[b] = bval[b];
    } else {
      oidx[b] = -1;
      oval[b] = cf{0.f, 0.f};
    }
  }
}

// The last chunk's finish (every other chunk's runs inside the next rows
// launch): one thread per (pair, bin) cell.
__global__ void __launch_bounds__(256) big_rows_finish_kernel(BigArgs a, long long np) {
  finish_cell(a, np, (long long)blockIdx.x * 256 + threadIdx.x);
}

__global__ void roots_kernel(cf *out, int L, int period) {
  // out[k] = exp(+2 pi i k / period), k < L
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < L) {
    double sn, cs;
    sincospi(2.0 * (double)k / (double)period, &sn, &cs);
    out[k] = cf{(float)cs, (float)sn};
  }
}

constexpr size_t kLds = 160 * 1024;

template <int N1>
size_t rows_max_bins() {
  using S = FlatShape<N1>;
  return (kLds - (size_t)S::SLOTS * S::XSTRIDE * sizeof(cf)) / (sizeof(uint64_t) + sizeof(cf));
}

template <int N1>
int launch_big_n1(Ctx *c, const FlatRun &r) {
  constexpr int N = N1 * 256;
  using S = FlatShape<N1>;
  BigArgs a;
  a.data = r.data;
  a.tmpl = r.tmpl;
  a.data_stride = r.data_stride;
  a.tmpl_stride = r.tmpl_stride;
  a.nd = r.nd;
  a.nt = r.nt;
  a.list = r.list;
  a.list_count = r.list_count;
  a.list_ooff = r.list_ooff;
  a.tw256 = c->twiddles(256);
  a.twN1 = c->twiddles(N1);
  const cf *roots = big_roots(c, N1);
  if (!a.tw256 || !a.twN1 || !roots) return -1;
  a.rootH = roots;
  a.rootL = roots + N1;
  const long long total = r.list ? (long long)r.list_cap : (long long)r.nd * r.nt;
  // scratch in chunks of pairs: at most c->big_chunk_bytes of intermediate data
  const long long chunk = std::max<long long>(1, std::min<long long>(total, c->big_chunk_bytes / ((long long)N * 8)));
  cf *scr = (cf *)c->scratch("big_scratch", (size_t)chunk * N * sizeof(cf));
  if (!scr) return -1;
  a.scratch = scr;
  const size_t lds_cols = (16 * kColStride + 16 * 17) * sizeof(cf);
  const float thr = r.threshold;
  a.seed = thr > 0.f ? (float)((double)thr * (double)thr) : 0.f;
  a.binsize = (uint32_t)r.binsize;
  const bool pow2 = (r.binsize & (r.binsize - 1)) == 0;
  a.bshift = pow2 ? ilog2((int)r.binsize) : -1;
  a.inv_bs = 1.0f / (float)r.binsize;
  const size_t nb_total = (r.we - r.ws + r.binsize - 1) / r.binsize;
  const size_t maxb = rows_max_bins<N1>();
  if (!lds_limit((const void *)big_cols_kernel<N1>, c->dev, lds_cols) ||
      !lds_limit((const void *)big_rows_kernel<N1>, c->dev, kLds))
    return -1;
  // rows workgroups per pair: enough for about 1024 workgroups per launch (at
  // least 8, at most one per row group), as long as the per-part tables stay
  // within 64 MiB
  constexpr int G = 256 / S::SLOTS;
  const size_t nb_piece = std::min(maxb, nb_total);
  int nsplit = (int)std::min<long long>(G, std::max<long long>(8, (1024 + chunk - 1) / chunk));
  nsplit = 1 << ilog2(nsplit);  // a power of two: parts split the row groups evenly
  while (nsplit > 1 && (size_t)chunk * nsplit * nb_piece * 16 > (64ull << 20)) nsplit /= 2;
  a.nsplit = nsplit;
  a.pkey = nullptr;
  a.pval = nullptr;
  // part tables in two halves, by chunk parity: a rows launch writes one
  // while its finishing blocks read the previous chunk's
  const size_t part_cells = (size_t)chunk * nsplit * nb_piece;
  uint64_t *pkey2 = nullptr;
  cf *pval2 = nullptr;
  if (nsplit > 1) {
    pkey2 = (uint64_t *)c->scratch("big_part_key", 2 * part_cells * sizeof(uint64_t));
    pval2 = (cf *)c->scratch("big_part_val", 2 * part_cells * sizeof(cf));
    if (!pkey2 || !pval2) return -1;
  }
  // the finish of a chunk runs in the next rows launch when the window is
  // one piece (one rows launch per chunk); else after each rows launch
  const bool fold = nsplit > 1 && nb_total <= maxb;
  BigArgs fin = a;
  long long fin_np = 0;
  int fin_blocks = 0;
  long long ci = 0;  // chunk index
  Ctx::ProfScope ps(c, r.list ? r.list_prof_name : r.prof_name);
  for (long long p0 = 0; p0 < total; p0 += chunk, ci++) {
    const long long np = std::min(chunk, total - p0);
    a.p0 = p0;
    if (nsplit > 1) {
      a.pkey = pkey2 + (ci & 1) * part_cells;
      a.pval = pval2 + (ci & 1) * part_cells;
    }
    hipLaunchKernelGGL(big_cols_kernel<N1>, dim3(N1 / 16, (unsigned)np), dim3(256), lds_cols, c->stream, a);
    if (!hip_ok(hipGetLastError(), "big_cols_kernel")) return -1;
    for (size_t b0 = 0; b0 < nb_total; b0 += maxb) {
      const size_t cnt = std::min(maxb, nb_total - b0);
      a.ws = (uint32_t)(r.ws + b0 * r.binsize);
      a.we = (uint32_t)std::min(r.we, r.ws + (b0 + cnt) * r.binsize);
      a.nb = (uint32_t)cnt;
      a.ostride = (uint32_t)nb_total;
      a.out_idx = r.out_idx + b0;
      a.out_val = r.out_val + b0;
      const size_t lds = (size_t)S::SLOTS * S::XSTRIDE * sizeof(cf) + cnt * (sizeof(uint64_t) + sizeof(cf));
      const long long row_blocks = np * nsplit;
      hipLaunchKernelGGL(big_rows_kernel<N1>, dim3((unsigned)(row_blocks + fin_blocks)), dim3(256), lds, c->stream, a,
                         fin, fin_np, fin_blocks);
      if (!hip_ok(hipGetLastError(), "big_rows_kernel")) return -1;
      fin_blocks = 0;
      if (nsplit > 1) {
        const long long cells = np * (long long)cnt;
        if (fold) {  // finished by the next rows launch, or after the loop
          fin = a;
          fin_np = np;
          fin_blocks = (int)((cells + 255) / 256);
        } else {
          hipLaunchKernelGGL(big_rows_finish_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream,
                             a, np);
          if (!hip_ok(hipGetLastError(), "big_rows_finish_kernel")) return -1;
        }
      }
    }
  }
  if (fin_blocks) {
    hipLaunchKernelGGL(big_rows_finish_kernel, dim3((unsigned)fin_blocks), dim3(256), 0, c->stream, fin, fin_np);
    if (!hip_ok(hipGetLastError(), "big_rows_finish_kernel")) return -1;
  }
  return 0;
}

}  // namespace

const cf *big_roots(Ctx *c, int n1) {
  const int n = n1 * 256;
  char rname[32];
  snprintf(rname, sizeof rname, "bigroots%d", n);
  cf *roots = (cf *)c->scratch(rname, (n1 + 256) * sizeof(cf));
  if (!roots) return nullptr;
  if (!c->big_roots_ready[ilog2(n1)]) {
    hipLaunchKernelGGL(roots_kernel, dim3((n1 + 255) / 256), dim3(256), 0, c->stream, roots, n1, n1);
    hipLaunchKernelGGL(roots_kernel, dim3(1), dim3(256), 0, c->stream, roots + n1, 256, n);
    if (!hip_ok(hipGetLastError(), "roots_kernel")) return nullptr;
    c->big_roots_ready[ilog2(n1)] = true;
  }
  return roots;
}

int launch_big(Ctx *c, size_t n, const FlatRun &r) {
  switch (n) {
    case 32768: return launch_big_n1<128>(c, r);
    case 65536: return launch_big_n1<256>(c, r);
    case 131072: return launch_big_n1<512>(c, r);
    case 262144: return launch_big_n1<1024>(c, r);
    case 524288: return launch_big_n1<2048>(c, r);
    case 1048576: return launch_big_n1<4096>(c, r);
    default: return fail("transform length %zu is not a four-step length", n);
  }
}

}  // namespace mfh
