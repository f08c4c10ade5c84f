// Fused flat matched filter for gfx950: D*conj(T) -> inverse FFT -> per-bin
// |z| argmax, with the correlation series never leaving registers and LDS.
//
// Reference behaviour (paths relative to the reference checkout):
//   product / transform / scan   src/balanced-inl.h:1102-1113, :876-969
//   selection rule               src/balanced-inl.h:886, :944, :956
//   GPU precedent (one transform per workgroup, LDS exchange, one elected
//   writer per bin)              src/gpu/tierb.slang:591-771
//
// Layout.  A transform of length N is carried by TPT = N/16 threads holding
// 16 points each: thread t owns elements t + u*TPT, u = 0..15.  The inverse
// FFT is a Stockham autosort over radix passes [2^b] + [16]*L (N = 2^b 16^L),
// every pass a register DFT preceded by twiddles from an FP64-generated table
// and followed by a padded LDS exchange; after the last pass thread t holds
// z[t + u*TPT].  A workgroup carries SLOTS = WG/TPT transforms side by side;
// each slot keeps one template spectrum in registers and walks a tile of data
// rows, so every template is read from HBM once per tile and every data row
// is reused from L2 across the template bank.
//
// Selection.  Each lag's key is (bits(|z|^2) << 32) | ~k with
// |z|^2 = fma(re, re, im*im) in FP32; only keys with |z|^2 > seed
// (seed = threshold^2, or 0) enter.  The per-bin maximum of the keys is the
// largest |z|^2 with the LOWEST lag among equals, independent of the order in
// which lanes and waves contribute: per-thread runs and wave64 shuffles reduce
// first and LDS ds_max_u64 finishes.  The lane that owns the winning lag then
// writes index and value, so the value always belongs to the reported lag.
#pragma once
#include "mfh_dft.h"
#include "mfh_wave.h"

namespace mfh {

struct FlatArgs {
  const cf *data;      // [rows][N] natural-order spectra (already offset to the first run row)
  const cf *tmpl;      // [rows][N] (already offset to the first run template)
  const cf *twiddle;   // per-pass tables, TwiddleLayout<N>
  int64_t *out_idx;    // [nd][nt][nb]
  cf *out_val;         // [nd][nt][nb]
  long long data_stride, tmpl_stride;  // elements between rows
  int nd, nt;          // run extent
  int dpw;             // data rows per workgroup (tiled mode)
  const int2 *list;    // list mode: (data row, template row) per pair
  const int *list_count;  // list mode: number of pairs in the list
  const long long *list_ooff;  // list mode: output element offset per pair (null: row-major d*nt + t)
  uint32_t ws, we;     // lag window [ws, we)
  uint32_t binsize, nb;  // nb: bins of this launch
  uint32_t ostride;      // output elements between pairs (>= nb)
  int bshift;          // log2(binsize) when a power of two, else -1
  float inv_bs;        // 1 / binsize
  float seed;          // threshold^2 (FP32), or 0
  int *zero;           // if set, workgroup 0 stores 0 there (a counter for the next kernel)
  int *hint;           // list mode, if set: workgroup 0 stores the list length there (host-mapped)
};

template <int N>
struct FlatShape {
  static constexpr int R = 16;
  static constexpr int TPT = N / 16;
  static constexpr int WG = TPT >= 256 ? TPT : 256;
  static constexpr int SLOTS = WG / TPT;
  static constexpr int LOGN = ilog2(N);
  static constexpr int FIRST = 1 << (LOGN % 4);  // radix of the first pass (1: none)
  static constexpr int NPASS16 = LOGN / 4;
  static constexpr int XCH = N + N / 16;          // padded exchange elements per slot
  // distance between slots' exchange buffers: one element more when several
  // slots share a wave (TPT < 64), so that equal offsets of neighbouring
  // slots fall in different LDS banks
  static constexpr int XSTRIDE = TPT < 64 ? XCH + 1 : XCH;
};

// Four-step column transforms: 16 lanes of a row group serve 16 adjacent
// columns, each column a 256-point transform with its own exchange buffer.
// A buffer stride of 273 elements (546 dwords, 34 mod 64) puts the 16 lanes'
// equal offsets in 16 distinct bank pairs; 272 would fold them onto two.
// (An even stride, 274, measured 2.8x the bank-conflict cycles of 273 in the
// columns kernel: the 16 columns are what must spread over the banks.)
constexpr int kColStride = FlatShape<256>::XCH + 1;

__device__ __forceinline__ uint32_t lds_pad(uint32_t i) { return i + (i >> 4); }

// The 16 points t + u*TPT of a row whose address is uniform across the wave,
// through a buffer descriptor: one vector offset (8 t) for the thread and the
// per-u offset 8 u TPT as a scalar, so the strided loads need no vector
// address arithmetic.  The descriptor covers exactly the row's n elements.
template <int N>
__device__ __forceinline__ void load_row_uniform(cf *dst, const cf *row, int t) {
  constexpr int TPT = N / 16;
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<cf *>(row), (short)0, N * (int)sizeof(cf), 0x00020000);
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, t * (int)sizeof(cf), u * TPT * (int)sizeof(cf), 0);
    dst[u] = __builtin_bit_cast(cf, x);
  }
}

// A wave-uniform row pointer, made provably uniform (scalar registers).
__device__ __forceinline__ const cf *uniform_ptr(const cf *p) {
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  return (const cf *)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ uint32_t bin_of(uint32_t k, const FlatArgs &a) {
  const uint32_t x = k - a.ws;
  if (a.bshift >= 0) return x >> a.bshift;
  uint32_t q = (uint32_t)((float)x * a.inv_bs);
  if (q * a.binsize > x) q -= 1;
  else if ((q + 1) * a.binsize <= x) q += 1;
  return q;
}

__device__ __forceinline__ uint64_t make_key(cf z, uint32_t k, float seed) {
  const float m2 = __builtin_fmaf(z.x, z.x, z.y * z.y);
  return m2 > seed ? ((uint64_t)__float_as_uint(m2) << 32) | (uint64_t)(~k) : 0ull;
}

// Per-bin maxima of a wave's keys when the lag window is the whole transform
// and bins are exactly 2 TPT lags wide: thread t's lags t + u*TPT, u < 16,
// form 8 runs of two consecutive u, and each run lies in one bin for all 64
// lanes of the wave.  Instead of one 64-lane reduction per run (6 exchanges
// each), a transposing butterfly halves every lane's list of run maxima per
// exchange step -- at xor distance 32 the upper lanes keep the upper half of
// the runs, at 16 the upper quarter of theirs, at 8 one run -- and three
// plain steps finish it: 10 exchanges instead of 48.  One lane per run then
// raises the bin's key in LDS.
template <int STEP>
__device__ __forceinline__ uint64_t pair_run_max(const cf *v, uint32_t lag0, int r, float seed) {
  const uint64_t a = make_key(v[2 * r], lag0 + (uint32_t)(2 * r * STEP), seed);
  const uint64_t b = make_key(v[2 * r + 1], lag0 + (uint32_t)((2 * r + 1) * STEP), seed);
  return a > b ? a : b;
}

// The transposing exchanges of the butterfly below, on 64-bit keys as two
// 32-bit halves.  v_permlane32_swap on (lo, hi) gives lanes 0-31 the lo of
// lane + 32 and lanes 32-63 the hi of lane - 32, with the own value in the
// other result; v_permlane16_swap does the same between even and odd rows of
// 16 lanes.  So the maximum of the two results is, in the lower half, lo
// over the pair of lanes and, in the upper half, hi over it: one exchange
// step with no selects and no LDS.  Like every cross-lane step here, these
// read other lanes' registers whether or not those lanes are active, so they
// are called only where the whole wave is (wave-uniform control flow).
__device__ __forceinline__ uint64_t max64(uint64_t a, uint64_t b) { return a > b ? a : b; }

__device__ __forceinline__ uint64_t join64(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

__device__ __forceinline__ uint64_t swap32_max(uint64_t lo, uint64_t hi) {
  const auto a = __builtin_amdgcn_permlane32_swap((uint32_t)lo, (uint32_t)hi, false, false);
  const auto b = __builtin_amdgcn_permlane32_swap((uint32_t)(lo >> 32), (uint32_t)(hi >> 32), false, false);
  return max64(join64(a[0], b[0]), join64(a[1], b[1]));
}

__device__ __forceinline__ uint64_t swap16_max(uint64_t lo, uint64_t hi) {
  const auto a = __builtin_amdgcn_permlane16_swap((uint32_t)lo, (uint32_t)hi, false, false);
  const auto b = __builtin_amdgcn_permlane16_swap((uint32_t)(lo >> 32), (uint32_t)(hi >> 32), false, false);
  return max64(join64(a[0], b[0]), join64(a[1], b[1]));
}

// The key of another lane of the same row of 16 through DPP.
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xf, 0xf, false);
  return join64(lo, hi);
}

// Max of a 64-bit key over the wave, every lane receiving it: DPP within
// rows of 16 lanes, then the row and half-wave swaps (call only with the whole
// wave active).
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  v = max64(v, dpp64<0xB1>(v));   // quad_perm(1, 0, 3, 2)
  v = max64(v, dpp64<0x4E>(v));   // quad_perm(2, 3, 0, 1)
  v = max64(v, dpp64<0x141>(v));  // row_half_mirror
  v = max64(v, dpp64<0x140>(v));  // row_mirror
  v = swap16_max(v, v);
  return swap32_max(v, v);
}

// R = 8 runs of two u each, streamed so that at most four partial maxima
// are live at once.  Lane lags are lag0 + u*STEP; lane 0's lag0 is wave_lag0,
// every lane's lag0 lies in [wave_lag0, wave_lag0 + STEP), and bins are 2 STEP
// wide starting at lag 0, so run rho of every lane lies in bin
// (wave_lag0 + 2 rho STEP) >> bshift.  Lane pairs: l and l ^ 32, l and l ^ 16,
// then l and its mirror 15 - l within the row (the other half of the row, as
// l ^ 8 would be), then three maxima within groups of 8 lanes.
template <int STEP>
__device__ __forceinline__ void runs8_max_to_bins(const cf *v, uint32_t lag0, uint32_t wave_lag0, int lane,
                                                  float seed, int bshift, uint64_t *bins) {
  const bool up1 = (lane & 32) != 0, up2 = (lane & 16) != 0, up3 = (lane & 8) != 0;
  uint64_t q[4];
#pragma unroll
  for (int i = 0; i < 4; i++)
    q[i] = swap32_max(pair_run_max<STEP>(v, lag0, i, seed), pair_run_max<STEP>(v, lag0, i + 4, seed));
  const uint64_t r0 = swap16_max(q[0], q[2]);
  const uint64_t r1 = swap16_max(q[1], q[3]);
  uint64_t w = max64(up3 ? r1 : r0, dpp64<0x140>(up3 ? r0 : r1));  // row_mirror
  w = max64(w, dpp64<0x141>(w));                                   // row_half_mirror
  w = max64(w, dpp64<0x4E>(w));                                    // quad_perm(2, 3, 0, 1)
  w = max64(w, dpp64<0xB1>(w));                                    // quad_perm(1, 0, 3, 2)
  const int rho = (up1 ? 4 : 0) + (up2 ? 2 : 0) + (up3 ? 1 : 0);
  if ((lane & 7) == 0 && w) {
    const uint32_t b = (wave_lag0 + (uint32_t)(rho * 2 * STEP)) >> bshift;
    atomicMax((unsigned long long *)&bins[b], (unsigned long long)w);
  }
}

// Whole-transform window with bins of exactly 2 TPT lags (n/8 bins); other
// shapes take the per-run reductions.  Returns false when the shape does not
// qualify.
template <int TPT>
__device__ __forceinline__ bool full_window_bins(const cf *v, int t, int lane, float seed, int bshift,
                                                 uint64_t *bins) {
  constexpr int LT = ilog2(TPT);
  if (bshift - LT != 1) return false;
  runs8_max_to_bins<TPT>(v, (uint32_t)t, (uint32_t)(t - lane), lane, seed, bshift, bins);
  return true;
}

// One bin over the whole transform, TPT >= 64 (a slot spans whole waves):
// the wave's key (bits(m2) << 32) | ~lag of its largest m2 > seed, lowest lag
// first -- the key the per-sample scan builds -- from 32-bit operations only:
// each lane's largest m2 on the bit patterns (nonnegative floats order like
// their bits) and its lowest u, the wave maximum, then the lowest lag among
// the lanes that hold it.  Returns false when the wave holds a NaN (bits
// above +inf): the caller then runs the per-sample scan, which never selects
// one.
//
// WINDOWED: the window [ws, we) need not be the whole transform, but both
// ends are multiples of 64 lags.  A wave's 64 lags at one u are then
// consecutive and all inside or all outside the window (a wave-uniform
// test), and the lags outside count as |z|^2 = 0, which never beats a seed
// >= 0 -- the per-sample scan's result, without its per-sample tests.
template <int TPT, bool WINDOWED = false>
__device__ __forceinline__ bool one_bin_wave_key(const cf *v, int t, float seed, uint64_t &wkey, uint32_t ws = 0,
                                                 uint32_t we = 0) {
  uint32_t mb[16];
#pragma unroll
  for (int u = 0; u < 16; u++) mb[u] = __float_as_uint(__builtin_fmaf(v[u].x, v[u].x, v[u].y * v[u].y));
  if constexpr (WINDOWED) {
    static_assert(TPT % 64 == 0, "a wave's lags at one u are 64 consecutive lags");
    const uint32_t kw = (uint32_t)__builtin_amdgcn_readfirstlane(t) & ~63u;  // the wave's first lag at u = 0
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const uint32_t k = kw + (uint32_t)(u * TPT);
      if (!(k >= ws && k < we)) mb[u] = 0u;  // wave-uniform
    }
  }
  uint32_t m = mb[0];
#pragma unroll
  for (int u = 1; u < 16; u++) m = max(m, mb[u]);
  const uint32_t mw = wave_umax(m);
  if (mw > 0x7f800000u) return false;  // wave-uniform
  // a wave whose maximum does not beat the seed raises no key, and skips
  // finding its lag (wave-uniform): under a threshold most waves stop here
  if (mw <= __float_as_uint(seed)) {
    wkey = 0ull;
    return true;
  }
  int bu = 15;
#pragma unroll
  for (int u = 14; u >= 0; u--) bu = mb[u] == m ? u : bu;
  const uint32_t kw = wave_umin(m == mw ? (uint32_t)(t + bu * TPT) : 0xffffffffu);
  wkey = ((uint64_t)mw << 32) | (uint64_t)(~kw);
  return true;
}

// Twiddle tables, one per twiddled radix-16 pass, holding only the powers
// r = 1 and r = 4 of each butterfly's twiddle, laid out [row][k]:
//   tw[pass_offset(p) + row*NS + k] = exp(+2 pi i r k / (16 NS)), r = 1, 4,
// k < NS.  For each row consecutive lanes read consecutive elements, so each
// of the 2 loads is one contiguous run; the other 13 powers are products of
// these (at most 4 roundings deep, ~2e-7 relative), 52 FMA-class
// instructions.  Small tables stay in the vector cache while the spectra
// stream through it: on gfx950 at n = 4096 two rows measured 2.8% faster per
// launch than four (w, w^2, w^4, w^8) and four 7% faster than eight.
constexpr int kTwiddleRows = 2;

template <int N>
struct TwiddleLayout {
  using S = FlatShape<N>;
  static constexpr int NS0 = S::FIRST > 1 ? S::FIRST : 1;
  // radix-16 passes with NS = NS0 * 16^p, p = 0..NPASS16-1
  static constexpr int offset(int p) {
    int off = 0, ns = NS0;
    for (int i = 0; i < p; i++) {
      if (ns > 1) off += ns * kTwiddleRows;
      ns *= 16;
    }
    return off;
  }
  static constexpr int total() { return offset(S::NPASS16); }
};

// Twiddles of one radix-16 pass for butterfly j = t (one butterfly per
// thread): w[r-1] = exp(+2 pi i r (t mod NS) / (16 NS)), r = 1..15.  Issued
// ahead of the barrier that precedes the pass, so their latency hides there.
template <int N, int NS>
__device__ __forceinline__ void load_twiddles16(cf *w, int t, const cf *tw) {
  const cf *p = tw + (t % NS);
  const cf w1 = p[0], w4 = p[NS];
  const cf w2 = cmul(w1, w1), w8 = cmul(w4, w4);
  w[0] = w1;
  w[1] = w2;
  w[3] = w4;
  w[7] = w8;
  w[2] = cmul(w1, w2);
  w[4] = cmul(w1, w4);
  w[5] = cmul(w2, w4);
  w[6] = cmul(w[2], w4);
#pragma unroll
  for (int r = 9; r < 16; r++) w[r - 1] = cmul(w[r - 9], w8);
}

// One Stockham pass of radix R: butterflies j = t + q*TPT (q < 16/R) read
// x[j + r*N/R] = registers v[q + r*Q], twiddle by exp(+2 pi i r (j mod NS) /
// (R NS)) and run the R-point DFT in place.  `w` holds the radix-16
// twiddles of this thread's butterfly when NS > 1 (then R = 16, Q = 1).
template <int N, int R, int NS>
__device__ __forceinline__ void stockham_pass(cf *v, const cf *w) {
  constexpr int Q = 16 / R;  // butterflies per thread
  static_assert(NS == 1 || R == 16, "twiddled passes are radix 16");
#pragma unroll
  for (int q = 0; q < Q; q++) {
    cf b[R];
#pragma unroll
    for (int r = 0; r < R; r++) b[r] = v[q + r * Q];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) b[r] = cmul(b[r], w[r - 1]);
    }
    dftR<R>(b);
#pragma unroll
    for (int r = 0; r < R; r++) v[q + r * Q] = b[r];
  }
}

// Store the pass outputs y[(j/NS) NS R + (j mod NS) + r NS] to the padded LDS
// buffer (index i lives at i + i/16: conflict-free for both the strided
// stores and the unit-stride loads).  Base address + compile-time offsets
// whenever the padding is separable.
template <int N, int R, int NS>
__device__ __forceinline__ void stockham_store(const cf *v, int t, cf *xch) {
  constexpr int TPT = N / 16;
  constexpr int Q = 16 / R;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int j = t + q * TPT;
    const int base = (j / NS) * NS * R + (j % NS);
    if constexpr (NS % 16 == 0) {
      cf *p = xch + lds_pad(base);
#pragma unroll
      for (int r = 0; r < R; r++) p[r * (NS + NS / 16)] = v[q + r * Q];
    } else if constexpr (NS == 1 && R == 16) {
      cf *p = xch + 17 * j;
#pragma unroll
      for (int r = 0; r < R; r++) p[r] = v[q + r * Q];
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) xch[lds_pad(base + r * NS)] = v[q + r * Q];
    }
  }
}

template <int N>
__device__ __forceinline__ void stockham_load(cf *v, int t, const cf *xch) {
  constexpr int TPT = N / 16;
  if constexpr (TPT % 16 == 0) {
    const cf *p = xch + lds_pad(t);
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = p[u * (TPT + TPT / 16)];
  } else {
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = xch[lds_pad(t + u * TPT)];
  }
}

// Full inverse FFT of the slot's N points held in v; returns with thread t
// holding z[t + u*TPT].  Each exchange is store -> barrier -> load, and the
// barrier that protects the buffer from the next store sits just before that
// store, after the pass's arithmetic, so waves do not idle between load and
// compute.  The caller must barrier before the next call's first store.
//
// With HOISTED, `wfirst` holds the twiddles of the first twiddled radix-16
// pass (first_twiddles below), derived once by the caller: they depend on
// the thread only, so a kernel that runs many pairs per thread derives them
// once instead of per pair.  The values, and so the results, are bit for bit
// those of the per-pair derivation.
template <int N, bool HOISTED = false>
__device__ __forceinline__ void inverse_fft(cf *v, int t, cf *xch, const cf *tw, const cf *wfirst = nullptr) {
  using S = FlatShape<N>;
  using L = TwiddleLayout<N>;
  constexpr int F = S::FIRST;
  constexpr int NS0 = L::NS0;
  cf w[15];
  if constexpr (F > 1) {
    stockham_pass<N, F, 1>(v, w);
    stockham_store<N, F, 1>(v, t, xch);
    if constexpr (!HOISTED) load_twiddles16<N, NS0>(w, t, tw + L::offset(0));
    __syncthreads();
    stockham_load<N>(v, t, xch);
  }
  stockham_pass<N, 16, NS0>(v, (HOISTED && F > 1) ? wfirst : w);
  if constexpr (S::NPASS16 >= 2) {
    if constexpr (F > 1) __syncthreads();
    stockham_store<N, 16, NS0>(v, t, xch);
    if constexpr (!(HOISTED && F == 1)) load_twiddles16<N, NS0 * 16>(w, t, tw + L::offset(1));
    __syncthreads();
    stockham_load<N>(v, t, xch);
    stockham_pass<N, 16, NS0 * 16>(v, (HOISTED && F == 1) ? wfirst : w);
  }
  if constexpr (S::NPASS16 >= 3) {
    __syncthreads();
    stockham_store<N, 16, NS0 * 16>(v, t, xch);
    load_twiddles16<N, NS0 * 256>(w, t, tw + L::offset(2));
    __syncthreads();
    stockham_load<N>(v, t, xch);
    stockham_pass<N, 16, NS0 * 256>(v, w);
  }
  static_assert(S::NPASS16 >= 1 && S::NPASS16 <= 3, "N range");
}

// The twiddles of the first twiddled radix-16 pass of a length-N transform
// for thread t (the pass after the small-radix one, or the second radix-16
// pass when N is a power of 16), as inverse_fft<N, true> takes them.
template <int N>
__device__ __forceinline__ void first_twiddles(cf *w, int t, const cf *tw) {
  using S = FlatShape<N>;
  using L = TwiddleLayout<N>;
  if constexpr (S::FIRST > 1) {
    load_twiddles16<N, L::NS0>(w, t, tw + L::offset(0));
  } else {
    static_assert(S::NPASS16 >= 2, "a power of 16 has a second radix-16 pass");
    load_twiddles16<N, 16>(w, t, tw + L::offset(1));
  }
}

// Bin-table modes of the flat kernel.
enum BinMode : int {
  kOneBin = 0,   // a single bin over the whole window
  kAligned = 1,  // window and binsize multiples of 64: every bin is wave-uniform
  kGeneral = 2,  // anything else
};

// The end of one pair's work in the flat kernel, after its inverse FFT has
// left z[t + u*TPT] in v: the per-bin keys of the lags in the window, the
// workgroup's bin maxima in LDS, and the writes of each bin's winning lag and
// value (-1 / 0 for bins nothing reached).  Every thread of the workgroup
// calls it (it holds a barrier); `valid` says whether this slot's pair is
// real, (d, tt) are its data row and template, nb = a.nb, and its output row starts at
// list_ooff[dd + slot] (list mode with offsets) or (d nt + tt) ostride.
// Clears `bins_next` for the next pair.
template <int N, int MODE, bool LIST>
__device__ __forceinline__ void select_and_write(const FlatArgs &a, int nb, const cf *v, int t, int lane,
                                                 uint64_t *bins, uint64_t *bins_next, bool valid, int d, int tt,
                                                 int dd, int slot) {
  constexpr int TPT = FlatShape<N>::TPT;
  if constexpr (MODE == kOneBin) {
    // ---- one bin: a wave key, one LDS max; the winning lag's owner writes ----
    uint64_t wkey = 0ull;
    bool have = false;
    if constexpr (TPT >= 64) {
      if (a.ws == 0 && a.we == (uint32_t)N) have = one_bin_wave_key<TPT>(v, t, a.seed, wkey);
      else if ((a.ws & 63u) == 0 && (a.we & 63u) == 0)
        have = one_bin_wave_key<TPT, true>(v, t, a.seed, wkey, a.ws, a.we);
    }
    if (!have) {
      // per-sample scan: strict, so the lowest lag keeps a tie
      float bm = a.seed;
      int bu = -1;
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const uint32_t k = (uint32_t)(t + u * TPT);
        const float m2 = __builtin_fmaf(v[u].x, v[u].x, v[u].y * v[u].y);
        if (k >= a.ws && k < a.we && m2 > bm) {
          bm = m2;
          bu = u;
        }
      }
      const uint64_t key =
          bu >= 0 ? ((uint64_t)__float_as_uint(bm) << 32) | (uint64_t)(~(uint32_t)(t + bu * TPT)) : 0ull;
      if constexpr (TPT >= 64) wkey = wave_max_u64(key);
      else wkey = key;  // several slots per wave: every lane raises its own key
    }
    if (TPT >= 64 ? (lane == 0 && wkey) : wkey != 0ull)
      atomicMax((unsigned long long *)&bins[0], (unsigned long long)wkey);
    __syncthreads();
    if (valid) {
      const long long o = (LIST && a.list_ooff) ? a.list_ooff[dd + slot] : ((long long)d * a.nt + tt) * a.ostride;
      const uint64_t win = bins[0];
      const uint32_t kwin = ~(uint32_t)win;
      if (win && (int)(kwin % TPT) == t) {
        // lag kwin = t + uw * TPT is this thread's sample uw
        const int uw = (int)(kwin / TPT);
        // a select per sample on the bit patterns, so that v stays in
        // registers (an indexed read would put the array in scratch)
        uint32_t zx = 0u, zy = 0u;
#pragma unroll
        for (int u = 0; u < 16; u++) {
          zx |= u == uw ? __float_as_uint(v[u].x) : 0u;
          zy |= u == uw ? __float_as_uint(v[u].y) : 0u;
        }
        const cf z = {__uint_as_float(zx), __uint_as_float(zy)};
        a.out_idx[o] = (int64_t)kwin;
        a.out_val[o] = z;
      }
      if (t == 0 && win == 0ull) {
        a.out_idx[o] = -1;
        a.out_val[o] = cf{0.f, 0.f};
      }
    }
  } else {
    // ---- several bins: per-thread runs, wave reductions, LDS max ----------
    // whole-transform window with power-of-two bins of >= 64 lags: the
    // transposing butterfly (wave-uniform test)
    const bool full = MODE == kAligned && a.ws == 0 && a.we == (uint32_t)N && a.bshift >= 6;
    if constexpr (TPT >= 64) {
      if (full && full_window_bins<TPT>(v, t, lane, a.seed, a.bshift, bins)) {
      } else {
        uint64_t acc = 0ull;
        int acc_bin = -1;
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const uint32_t k = (uint32_t)(t + u * TPT);
          const uint64_t key = (k >= a.ws && k < a.we) ? make_key(v[u], k, a.seed) : 0ull;
          bool uni;
          uint32_t kf = 0;
          if constexpr (MODE == kAligned) {
            kf = (uint32_t)__builtin_amdgcn_readfirstlane((int)(k - lane));
            uni = kf >= a.ws && kf < a.we;
          } else {
            kf = (uint32_t)__builtin_amdgcn_readfirstlane((int)(k - lane));
            uni = kf >= a.ws && kf + 63 < a.we && bin_of(kf, a) == bin_of(kf + 63, a);
          }
          if (uni) {
            const int b = (int)bin_of(kf, a);
            if (b != acc_bin) {
              if (acc_bin >= 0) {
                const uint64_t w = wave_max_u64(acc);
                if (lane == 0 && w) atomicMax((unsigned long long *)&bins[acc_bin], (unsigned long long)w);
              }
              acc = 0ull;
              acc_bin = b;
            }
            acc = key > acc ? key : acc;
          } else if (key) {
            atomicMax((unsigned long long *)&bins[bin_of(k, a)], (unsigned long long)key);
          }
        }
        if (acc_bin >= 0) {
          const uint64_t w = wave_max_u64(acc);
          if (lane == 0 && w) atomicMax((unsigned long long *)&bins[acc_bin], (unsigned long long)w);
        }
      }
    } else {
      // several transforms share a wave: per-thread runs, then LDS atomics
      uint64_t acc = 0ull;
      int acc_bin = -1;
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const uint32_t k = (uint32_t)(t + u * TPT);
        if (k >= a.ws && k < a.we) {
          const uint64_t key = make_key(v[u], k, a.seed);
          const int b = (int)bin_of(k, a);
          if (b != acc_bin) {
            if (acc) atomicMax((unsigned long long *)&bins[acc_bin], (unsigned long long)acc);
            acc = 0ull;
            acc_bin = b;
          }
          acc = key > acc ? key : acc;
        }
      }
      if (acc) atomicMax((unsigned long long *)&bins[acc_bin], (unsigned long long)acc);
    }
    __syncthreads();

    // ---- the owner of each winning lag writes it; empty bins get -1/0 ----
    if (valid) {
      const long long o = (LIST && a.list_ooff) ? a.list_ooff[dd + slot] : ((long long)d * a.nt + tt) * a.ostride;
      int64_t *oidx = a.out_idx + o;
      cf *oval = a.out_val + o;
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const uint32_t k = (uint32_t)(t + u * TPT);
        if (k >= a.ws && k < a.we) {
          const uint64_t key = make_key(v[u], k, a.seed);
          if (key) {
            const uint32_t b = bin_of(k, a);
            if (bins[b] == key) {
              oidx[b] = (int64_t)k;
              oval[b] = v[u];
            }
          }
        }
      }
      for (int b = t; b < nb; b += TPT) {
        if (bins[b] == 0ull) {
          oidx[b] = -1;
          oval[b] = cf{0.f, 0.f};
        }
      }
    }
  }
  // the other buffer was last read before this pair's first barrier
  for (int b = t; b < nb; b += TPT) bins_next[b] = 0ull;
}

template <int N, int MODE, bool LIST>
__global__ void __launch_bounds__(FlatShape<N>::WG) __attribute__((amdgpu_waves_per_eu(4)))
flat_kernel(FlatArgs a) {
  using S = FlatShape<N>;
  constexpr int TPT = S::TPT;
  extern __shared__ __align__(16) unsigned char smem[];
  cf *xch_all = reinterpret_cast<cf *>(smem);
  uint64_t *bins_all = reinterpret_cast<uint64_t *>(xch_all + S::SLOTS * S::XSTRIDE);

  const int slot = threadIdx.x / TPT;
  const int t0 = threadIdx.x % TPT;
  const int lane = threadIdx.x & 63;
  cf *xch = xch_all + slot * S::XSTRIDE;
  const int nb = (int)a.nb;

  // Tiled mode: this slot's template stays in registers while the workgroup
  // walks a tile of data rows.  List mode: (data, template) pairs come from a
  // device list whose length another kernel wrote, drained grid-stride.
  int tt = blockIdx.x * S::SLOTS + slot;
  bool valid = tt < a.nt;
  int dbeg = blockIdx.y * a.dpw;
  int dend = min(a.nd, dbeg + a.dpw);
  if constexpr (LIST) {
    dbeg = blockIdx.x * S::SLOTS;
    dend = *a.list_count;
  }
  if (a.zero && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *a.zero = 0;
  if constexpr (LIST) {
    // the list length, for the host's sizing of the next refine launch
    if (a.hint && blockIdx.x == 0 && threadIdx.x == 0) *a.hint = dend;
    // workgroup-uniform: the list is shorter than this slot range (with few
    // survivors most of a refine grid leaves here, before any other load)
    if (dbeg >= dend) return;
  }
  // the first twiddled pass's twiddles, once for every pair this thread runs
  cf wfirst[15];
  first_twiddles<N>(wfirst, t0, a.twiddle);
  // both bin buffers start empty
  for (int i = threadIdx.x; i < 2 * S::SLOTS * nb; i += S::WG) bins_all[i] = 0ull;
  __syncthreads();

  const int dstep = LIST ? (int)gridDim.x * S::SLOTS : 1;
  // list mode: each pair's list entry is read one trip ahead, so the load
  // that the row addresses depend on is in flight during the previous pair
  int2 pr_next = int2{0, 0};
  if constexpr (LIST) pr_next = dbeg + slot < dend ? a.list[dbeg + slot] : int2{0, 0};
  for (int dd = dbeg, it = 0; dd < dend; dd += dstep, it++) {
    // Everything derived from the thread index (bins, window flags, LDS and
    // twiddle addresses) is loop-invariant; recomputing it per pair is far
    // cheaper than the registers the hoisted copies would pin.
    int t = t0;
    asm volatile("" : "+v"(t));
    uint64_t *bins = bins_all + ((it & 1) * S::SLOTS + slot) * nb;
    uint64_t *bins_next = bins_all + (((it + 1) & 1) * S::SLOTS + slot) * nb;
    int d = dd;
    cf v[16];
    if constexpr (LIST) {
      const int li = dd + slot;
      valid = li < dend;
      const int2 pr = pr_next;  // a.list[li] when valid
      pr_next = li + dstep < dend ? a.list[li + dstep] : int2{0, 0};
      d = pr.x;
      tt = pr.y;
      const cf *dp = a.data + (long long)d * a.data_stride;
      const cf *tp = a.tmpl + (long long)tt * a.tmpl_stride;
      if constexpr (TPT >= 64) {
        // one slot per wave: the pair, and so both rows, are wave-uniform
        cf dv[16], tv[16];
        load_row_uniform<N>(dv, uniform_ptr(dp), t);
        load_row_uniform<N>(tv, uniform_ptr(tp), t);
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = cmul_conj(dv[u], tv[u]);
      } else {
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = cmul_conj(dp[t + u * TPT], tp[t + u * TPT]);
      }
    } else {
      // the slot's template is read again for every data row of the tile
      // (from the caches after the first): holding it in registers would
      // cost the 32 VGPRs the hoisted twiddles use.  A slot past the bank
      // (valid false) runs template 0 and writes nothing.
      const cf *dp = a.data + (long long)d * a.data_stride;  // uniform: d is the tile's row
      const cf *tp = a.tmpl + (long long)(valid ? tt : 0) * a.tmpl_stride;
      cf dv[16], tv[16];
      load_row_uniform<N>(dv, dp, t);
      if constexpr (TPT >= 64) {
        load_row_uniform<N>(tv, uniform_ptr(tp), t);
      } else {
#pragma unroll
        for (int u = 0; u < 16; u++) tv[u] = tp[t + u * TPT];
      }
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = cmul_conj(dv[u], tv[u]);
    }
    inverse_fft<N, true>(v, t, xch, a.twiddle, wfirst);
    select_and_write<N, MODE, LIST>(a, nb, v, t, lane, bins, bins_next, valid, d, tt, dd, slot);
  }
}

}  // namespace mfh
