// Uneven batched refine: one persistent launch drains a device work queue of
// (data spectrum, template spectrum) pairs whose transform lengths differ.
//
// Work queue.  A plan turns requests (one data spectrum, c templates, a
// length n, binsize, threshold and window each) into items, one per pair,
// grouped into tiles of equal length: a tile is as many pairs as one
// 1024-lane workgroup carries side by side (1024 / (n/16) for n <= 16384, one
// pair for the four-step lengths).  Tiles are ordered longest first, so the
// last tiles to be taken are the shortest and the tail of the launch stays
// short.  Every workgroup takes the next tile with one atomic on a device
// counter, runs it and repeats until the counter passes the tile count, so
// each workgroup reaches the exit.  One workgroup of 1024 lanes per CU.
//
// Arithmetic.  A pair of length n <= 16384 runs the same product,
// inverse_fft<n> and twiddle tables as the flat kernel; the four-step lengths
// run the same columns/rows split, roots and tables as the two-kernel path,
// here inside one workgroup with a per-workgroup scratch.  Selection is the
// flat kernel's exact key rule, so indices and values equal the per-request
// flat filter bit for bit.
//
// Reference behaviour: each pair is one `ap_mf_run` / `Context.peaks` call of
// the flat filter (src/matchfilt.c:349-420, python/matchedfilter/_vkcompute.py:962);
// the list-driven GPU refine of python/matchedfilter/_vkcompute.py:534-850 is
// the single-length precedent.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mfh_flat.h"
#include "mfh_internal.h"
#include "mfhip.h"

namespace mfh {

namespace {

struct QItem {
  long long d_off, t_off, o_off;  // element offsets: data, template (pool), outputs
  uint32_t ws, we, binsize, nb;   // window, bin width, bins of this item
  float seed, inv_bs;             // threshold^2 (FP32) or 0; 1 / binsize
  int bshift, pad;                // log2(binsize) for powers of two, else -1
};

struct QTile {
  int log2n, first, count, pad;  // length class, first item, items (<= slots)
};

constexpr int kQThreads = 1024;
constexpr int kQMaxLog2 = 16;
// exchange space of a tile: 1024 lanes x 16 points, padded by 1/16, plus one
// element per slot where slots share a wave (largest at n = 64: 256 slots)
template <int N>
constexpr size_t q_xch_bytes() {
  return (size_t)(kQThreads / FlatShape<N>::TPT) * FlatShape<N>::XSTRIDE * sizeof(cf);
}
constexpr size_t kQXch = q_xch_bytes<64>();
static_assert(q_xch_bytes<128>() <= kQXch && q_xch_bytes<4096>() <= kQXch && q_xch_bytes<16384>() <= kQXch, "");
static_assert((size_t)64 * kColStride * sizeof(cf) <= kQXch, "four-step columns");
static_assert(q_xch_bytes<128>() <= kQXch && q_xch_bytes<256>() <= kQXch, "four-step rows");
// per-bin (key, value) table of a four-step pair, after the exchange space;
// with the 16-byte tile slot the total is the 160 KiB a workgroup may hold
constexpr int kQBigBins = (int)((160 * 1024 - 16 - kQXch) / (sizeof(uint64_t) + sizeof(cf)));
constexpr size_t kQLds = kQXch + (size_t)kQBigBins * (sizeof(uint64_t) + sizeof(cf));

struct QArgs {
  const cf *pool;
  const QItem *items;
  const QTile *tiles;
  int ntiles;
  int *counter;
  const cf *tw[kQMaxLog2 + 1];  // flat twiddle tables by log2 length (and 128, 256)
  const cf *roots[2];           // four-step roots: n = 32768, 65536
  cf *scratch;                  // [gridDim.x][65536] four-step intermediate
  int64_t *out_idx;
  cf *out_val;
};

__device__ __forceinline__ uint32_t q_bin(uint32_t k, const QItem &it) {
  const uint32_t x = k - it.ws;
  if (it.bshift >= 0) return x >> it.bshift;
  uint32_t q = (uint32_t)((float)x * it.inv_bs);
  if (q * it.binsize > x) q -= 1;
  else if ((q + 1) * it.binsize <= x) q += 1;
  return q;
}

// Raise the bin keys of a slot's 16 lags (lag t + u*TPT); wave reductions
// where a wave's 64 lags share one bin, per-lane LDS atomics elsewhere.
template <int TPT>
__device__ __forceinline__ void q_select(const cf *v, int t, int tid, const QItem &it, uint64_t *bkey) {
  const int lane = tid & 63;
  if constexpr (TPT >= 64) {
    // whole-transform window in n/8 bins: the transposing butterfly
    if (it.ws == 0 && it.we == 16u * TPT && it.bshift == ilog2(TPT) + 1) {
      runs8_max_to_bins<TPT>(v, t, lane, it.seed, it.bshift, bkey);
      return;
    }
  }
  uint64_t acc = 0ull;
  int acc_bin = -1;
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const uint32_t k = (uint32_t)(t + u * TPT);
    const bool in = k >= it.ws && k < it.we;
    const uint64_t key = in ? make_key(v[u], k, it.seed) : 0ull;
    if constexpr (TPT >= 64) {
      // one wave = 64 consecutive lags of one slot: a wave-uniform test
      const uint32_t kf = (uint32_t)__builtin_amdgcn_readfirstlane((int)(k - lane));
      const bool uni = kf >= it.ws && kf + 63 < it.we && q_bin(kf, it) == q_bin(kf + 63, it);
      if (uni) {
        const int b = (int)q_bin(kf, it);
        if (b != acc_bin) {
          if (acc_bin >= 0) {
            const uint64_t w = wave_max_u64(acc);
            if (lane == 0 && w) atomicMax((unsigned long long *)&bkey[acc_bin], (unsigned long long)w);
          }
          acc = 0ull;
          acc_bin = b;
        }
        acc = key > acc ? key : acc;
      } else if (key) {
        atomicMax((unsigned long long *)&bkey[q_bin(k, it)], (unsigned long long)key);
      }
    } else {
      if (key) atomicMax((unsigned long long *)&bkey[q_bin(k, it)], (unsigned long long)key);
    }
  }
  if constexpr (TPT >= 64) {
    if (acc_bin >= 0) {
      const uint64_t w = wave_max_u64(acc);
      if (lane == 0 && w) atomicMax((unsigned long long *)&bkey[acc_bin], (unsigned long long)w);
    }
  }
}

// A tile of pairs of length N <= 16384: SLOTS = 1024 / (N/16) transforms side
// by side.  After the transform the slot's exchange space holds its bin keys
// (nb <= N).  Ends after its last barrier; the caller barriers before reuse.
template <int N>
__device__ __forceinline__ void q_flat(const QArgs &a, const QTile &tile, cf *smem, int tid) {
  constexpr int TPT = N / 16;
  constexpr int XCH = FlatShape<N>::XSTRIDE;
  const int slot = tid / TPT;
  const int t = tid % TPT;
  const bool valid = slot < tile.count;
  const QItem it = a.items[tile.first + (valid ? slot : 0)];
  cf *xch = smem + slot * XCH;
  const cf *D = a.pool + it.d_off;
  const cf *T = a.pool + it.t_off;
  cf v[16];
#pragma unroll
  for (int u = 0; u < 16; u++) v[u] = cmul_conj(D[t + u * TPT], T[t + u * TPT]);
  inverse_fft<N>(v, t, xch, a.tw[FlatShape<N>::LOGN]);
  __syncthreads();  // every slot's last exchange read is done
  uint64_t *bkey = reinterpret_cast<uint64_t *>(xch);
  const int nb = (int)it.nb;
  for (int b = t; b < nb; b += TPT) bkey[b] = 0ull;
  __syncthreads();
  q_select<TPT>(v, t, tid, it, bkey);
  __syncthreads();
  if (valid) {
    int64_t *oidx = a.out_idx + it.o_off;
    cf *oval = a.out_val + it.o_off;
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const uint32_t k = (uint32_t)(t + u * TPT);
      if (k >= it.ws && k < it.we) {
        const uint64_t key = make_key(v[u], k, it.seed);
        if (key) {
          const uint32_t b = q_bin(k, it);
          if (bkey[b] == key) {
            oidx[b] = (int64_t)k;
            oval[b] = v[u];
          }
        }
      }
    }
    for (int b = t; b < nb; b += TPT) {
      if (bkey[b] == 0ull) {
        oidx[b] = -1;
        oval[b] = cf{0.f, 0.f};
      }
    }
  }
}

// One four-step pair, n = N1 x 256 (f = f1 + N1 f2, k = k2 + 256 k1), inside
// the workgroup: columns (64 per round: four groups of 16 adjacent f1, each
// column a 256-point transform on 16 lanes) go to this workgroup's scratch,
// then rows (1024 / (N1/16) per round) transform and raise a per-bin
// (key, value) table in LDS; after each round the holder of a bin's key
// records its value.
template <int N1>
__device__ __forceinline__ void q_big(const QArgs &a, const QTile &tile, cf *smem, int tid) {
  constexpr int N = N1 * 256;
  const QItem it = a.items[tile.first];
  const cf *D = a.pool + it.d_off;
  const cf *T = a.pool + it.t_off;
  cf *Y = a.scratch + (long long)blockIdx.x * 65536;
  const cf *rootH = a.roots[N1 == 256];
  const cf *rootL = rootH + N1;
  {
    const int s = tid & 15;
    const int t = (tid >> 4) & 15;
    const int g = tid >> 8;
    cf *xch = smem + (g * 16 + s) * kColStride;
    for (int r = 0; r < N1 / 64; r++) {
      const int f1 = (r * 4 + g) * 16 + s;
      cf v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int f = f1 + N1 * (t + 16 * u);
        v[u] = cmul_conj(D[f], T[f]);
      }
      inverse_fft<256>(v, t, xch, a.tw[8]);
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int k2 = t + 16 * u;
        const int m = f1 * k2;
        const cf w = cmul(rootH[m >> 8], rootL[m & 255]);
        Y[(long long)k2 * N1 + f1] = cmul(v[u], w);
      }
      __syncthreads();  // exchange reusable; after the last round, Y is visible
    }
  }
  constexpr int TPT = N1 / 16;
  constexpr int SLOTS = kQThreads / TPT;
  const int slot = tid / TPT;
  const int t = tid % TPT;
  cf *xch = smem + slot * FlatShape<N1>::XSTRIDE;
  uint64_t *bkey = reinterpret_cast<uint64_t *>(reinterpret_cast<unsigned char *>(smem) + kQXch);
  cf *bval = reinterpret_cast<cf *>(bkey + kQBigBins);
  const int nb = (int)it.nb;
  for (int b = tid; b < nb; b += kQThreads) bkey[b] = 0ull;
  __syncthreads();
  for (int r = 0; r < 256 / SLOTS; r++) {
    const int k2 = r * SLOTS + slot;
    cf v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = Y[(long long)k2 * N1 + t + TPT * u];
    inverse_fft<N1>(v, t, xch, a.tw[ilog2(N1)]);
    uint64_t acc = 0ull;
    int acc_bin = -1;
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const uint32_t k = (uint32_t)(k2 + 256 * (t + TPT * u));
      if (k >= it.ws && k < it.we) {
        const uint64_t key = make_key(v[u], k, it.seed);
        const int b = (int)q_bin(k, it);
        if (b != acc_bin) {
          if (acc) atomicMax((unsigned long long *)&bkey[acc_bin], (unsigned long long)acc);
          acc = 0ull;
          acc_bin = b;
        }
        acc = key > acc ? key : acc;
      }
    }
    if (acc) atomicMax((unsigned long long *)&bkey[acc_bin], (unsigned long long)acc);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const uint32_t k = (uint32_t)(k2 + 256 * (t + TPT * u));
      if (k >= it.ws && k < it.we) {
        const uint64_t key = make_key(v[u], k, it.seed);
        if (key) {
          const uint32_t b = q_bin(k, it);
          if (bkey[b] == key) bval[b] = v[u];
        }
      }
    }
    __syncthreads();  // values settled, exchange reusable
  }
  int64_t *oidx = a.out_idx + it.o_off;
  cf *oval = a.out_val + it.o_off;
  for (int b = tid; b < nb; b += kQThreads) {
    const uint64_t key = bkey[b];
    if (key) {
      oidx[b] = (int64_t)(~(uint32_t)(key & 0xffffffffull));
      oval[b] = bval[b];
    } else {
      oidx[b] = -1;
      oval[b] = cf{0.f, 0.f};
    }
  }
  (void)N;
}

__global__ void __launch_bounds__(kQThreads) queue_kernel(QArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int tile_slot;
  cf *smem = reinterpret_cast<cf *>(smem_raw);
  for (;;) {
    if (threadIdx.x == 0) tile_slot = atomicAdd(a.counter, 1);
    __syncthreads();
    const int ti = tile_slot;
    if (ti >= a.ntiles) break;  // uniform: every lane read the same slot
    const QTile tile = a.tiles[ti];
    // the lane index is made opaque per tile: otherwise every branch's
    // lane-derived addresses are hoisted out of the loop and pin registers
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    switch (tile.log2n) {
      case 6: q_flat<64>(a, tile, smem, tid); break;
      case 7: q_flat<128>(a, tile, smem, tid); break;
      case 8: q_flat<256>(a, tile, smem, tid); break;
      case 9: q_flat<512>(a, tile, smem, tid); break;
      case 10: q_flat<1024>(a, tile, smem, tid); break;
      case 11: q_flat<2048>(a, tile, smem, tid); break;
      case 12: q_flat<4096>(a, tile, smem, tid); break;
      case 13: q_flat<8192>(a, tile, smem, tid); break;
      case 14: q_flat<16384>(a, tile, smem, tid); break;
      case 15: q_big<128>(a, tile, smem, tid); break;
      case 16: q_big<256>(a, tile, smem, tid); break;
      default: break;
    }
    __syncthreads();  // tile done: LDS and tile_slot reusable
  }
}

__global__ void q_zero_kernel(int *counter) {
  if (threadIdx.x == 0) *counter = 0;
}

}  // namespace

struct Queue {
  Ctx c;
  int cus = 0;
  // pool of spectra: caller's device memory, or a copy owned here
  const cf *pool = nullptr;
  size_t pool_elems = 0;
  // plan
  bool planned = false;
  long long plan_gen = 0;
  int nreq = 0;
  int ntiles = 0;
  long long nout = 0;
  long long npairs = 0;
  int max_log2 = 0;
  std::vector<long long> req_out;  // output offset per request
  std::vector<int> req_nb, req_nc;
  int grid_per_cu = 1;
};

}  // namespace mfh

using mfh::cf;
using mfh::fail;
using mfh::hip_ok;
using mfh::Queue;

extern "C" mfh_queue *mfh_queue_create(int device) {
  int count = mfh_device_count();
  if (device < 0 || device >= count) {
    fail("no HIP device with index %d (%d visible)", device, count);
    return nullptr;
  }
  if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return nullptr;
  Queue *q = new Queue;
  q->c.dev = device;
  hipDeviceProp_t prop;
  if (!hip_ok(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties") ||
      !hip_ok(hipStreamCreateWithFlags(&q->c.stream, hipStreamNonBlocking), "hipStreamCreate")) {
    delete q;
    return nullptr;
  }
  q->cus = prop.multiProcessorCount;
  static bool attr_set = false;
  if (!attr_set) {
    if (!hip_ok(hipFuncSetAttribute((const void *)mfh::queue_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)mfh::kQLds),
                "hipFuncSetAttribute")) {
      delete q;
      return nullptr;
    }
    attr_set = true;
  }
  return reinterpret_cast<mfh_queue *>(q);
}

extern "C" void mfh_queue_destroy(mfh_queue *queue) { delete reinterpret_cast<Queue *>(queue); }

extern "C" int mfh_queue_set_pool(mfh_queue *queue, const void *spectra_c64, size_t elems, int on_device) {
  if (!queue || (!spectra_c64 && elems)) return fail("null argument");
  Queue *q = reinterpret_cast<Queue *>(queue);
  hipSetDevice(q->c.dev);
  q->planned = false;  // offsets were checked against the previous pool
  if (on_device) {
    q->pool = (const cf *)spectra_c64;
  } else {
    cf *p = (cf *)q->c.scratch("pool", elems * sizeof(cf));
    if (!p) return -1;
    if (!hip_ok(hipMemcpyAsync(p, spectra_c64, elems * sizeof(cf), hipMemcpyHostToDevice, q->c.stream),
                "copy pool") ||
        !hip_ok(hipStreamSynchronize(q->c.stream), "hipStreamSynchronize"))
      return -1;
    q->pool = p;
  }
  q->pool_elems = elems;
  return 0;
}

extern "C" long long mfh_queue_plan(mfh_queue *queue, int nreq, const int64_t *n, const int64_t *data_off,
                                    const int64_t *tmpl_off, const int32_t *ncand, const int64_t *binsize,
                                    const float *threshold, const int64_t *ws, const int64_t *we,
                                    int64_t *out_off) {
  using namespace mfh;
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  q->planned = false;
  if (nreq < 0) return fail("negative request count");
  if (nreq && (!n || !data_off || !tmpl_off || !ncand || !binsize || !threshold || !ws || !we))
    return fail("null request array");
  std::vector<std::vector<QItem>> by_class(kQMaxLog2 + 1);
  q->req_out.assign(nreq, 0);
  q->req_nb.assign(nreq, 0);
  q->req_nc.assign(nreq, 0);
  long long nout = 0, npairs = 0;
  for (int r = 0; r < nreq; r++) {
    const long long len = n[r];
    if (len < 64 || len > 65536 || (len & (len - 1)))
      return fail("request %d: transform length %lld is not supported", r, len);
    if (ncand[r] < 0) return fail("request %d: negative candidate count", r);
    if (binsize[r] < 1) return fail("request %d: binsize must be >= 1", r);
    if (ws[r] < 0 || we[r] > len || ws[r] >= we[r])
      return fail("request %d: window (%lld, %lld) is empty or outside [0, %lld]", r, (long long)ws[r],
                  (long long)we[r], len);
    if (data_off[r] < 0 || (size_t)(data_off[r] + len) > q->pool_elems)
      return fail("request %d: data spectrum outside the pool", r);
    if (tmpl_off[r] < 0 || (size_t)(tmpl_off[r] + (long long)ncand[r] * len) > q->pool_elems)
      return fail("request %d: template spectra outside the pool", r);
    const float thr = threshold[r];  // <= 0 (or NaN): every nonzero sample may report
    const int lg = ilog2((int)len);
    const long long nb = (we[r] - ws[r] + binsize[r] - 1) / binsize[r];
    q->req_out[r] = nout;
    q->req_nb[r] = (int)nb;
    q->req_nc[r] = ncand[r];
    if (out_off) out_off[r] = nout;
    QItem base{};
    base.d_off = data_off[r];
    base.seed = thr > 0.f ? (float)((double)thr * (double)thr) : 0.f;
    base.binsize = (uint32_t)binsize[r];
    const bool pow2 = (binsize[r] & (binsize[r] - 1)) == 0;
    base.bshift = pow2 ? ilog2((int)binsize[r]) : -1;
    base.inv_bs = 1.0f / (float)binsize[r];
    // four-step pairs keep at most kQBigBins bins per item: longer windows
    // are split at bin boundaries into items that share the pair
    const long long piece = lg >= 15 ? kQBigBins : nb;
    for (int j = 0; j < ncand[r]; j++) {
      for (long long b0 = 0; b0 < nb; b0 += piece) {
        QItem it = base;
        it.t_off = tmpl_off[r] + (long long)j * len;
        it.o_off = nout + (long long)j * nb + b0;
        it.ws = (uint32_t)(ws[r] + b0 * binsize[r]);
        it.we = (uint32_t)std::min<long long>(we[r], ws[r] + (b0 + piece) * binsize[r]);
        it.nb = (uint32_t)std::min<long long>(piece, nb - b0);
        by_class[lg].push_back(it);
      }
    }
    nout += (long long)ncand[r] * nb;
    npairs += ncand[r];
  }
  std::vector<QItem> items;
  std::vector<QTile> tiles;
  q->max_log2 = 0;
  for (int lg = kQMaxLog2; lg >= 6; lg--) {  // longest first
    auto &v = by_class[lg];
    if (v.empty()) continue;
    q->max_log2 = std::max(q->max_log2, lg);
    const int slots = lg >= 15 ? 1 : kQThreads / ((1 << lg) / 16);
    for (size_t i = 0; i < v.size(); i += slots) {
      QTile t{};
      t.log2n = lg;
      t.first = (int)(items.size() + i);
      t.count = (int)std::min<size_t>(slots, v.size() - i);
      tiles.push_back(t);
    }
    items.insert(items.end(), v.begin(), v.end());
  }
  if (items.size() > (size_t)INT32_MAX) return fail("too many pairs in one plan");
  hipSetDevice(q->c.dev);
  // tables for every length in the plan
  for (int lg = 6; lg <= kQMaxLog2; lg++) {
    if (by_class[lg].empty()) continue;
    if (lg <= 14) {
      if (!q->c.twiddles((size_t)1 << lg)) return -1;
    } else {
      if (!q->c.twiddles(256) || !q->c.twiddles((size_t)1 << (lg - 8)) || !big_roots(&q->c, 1 << (lg - 8)))
        return -1;
    }
  }
  QItem *ditems = (QItem *)q->c.scratch("q_items", std::max<size_t>(1, items.size()) * sizeof(QItem));
  QTile *dtiles = (QTile *)q->c.scratch("q_tiles", std::max<size_t>(1, tiles.size()) * sizeof(QTile));
  if (!ditems || !dtiles) return -1;
  if ((!items.empty() && !hip_ok(hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(QItem),
                                                hipMemcpyHostToDevice, q->c.stream),
                                 "copy items")) ||
      (!tiles.empty() && !hip_ok(hipMemcpyAsync(dtiles, tiles.data(), tiles.size() * sizeof(QTile),
                                                hipMemcpyHostToDevice, q->c.stream),
                                 "copy tiles")) ||
      !hip_ok(hipStreamSynchronize(q->c.stream), "hipStreamSynchronize"))
    return -1;
  q->nreq = nreq;
  q->ntiles = (int)tiles.size();
  q->nout = nout;
  q->npairs = npairs;
  q->plan_gen++;
  q->planned = true;
  return nout;
}

namespace mfh {

static int queue_launch(Queue *q, int64_t *idx, cf *val) {
  Ctx *c = &q->c;
  int *counter = (int *)c->scratch("q_counter", sizeof(int));
  if (!counter) return -1;
  QArgs a{};
  a.pool = q->pool;
  a.items = (const QItem *)c->bufs["q_items"].p;
  a.tiles = (const QTile *)c->bufs["q_tiles"].p;
  a.ntiles = q->ntiles;
  a.counter = counter;
  for (int lg = 6; lg <= 14; lg++) {
    auto it = c->tw_by_n.find((size_t)1 << lg);
    a.tw[lg] = it == c->tw_by_n.end() ? nullptr : it->second;
  }
  int grid = std::max(1, std::min(q->cus * q->grid_per_cu, q->ntiles));
  if (q->max_log2 >= 15) {
    a.scratch = (cf *)c->scratch("q_scratch", (size_t)grid * 65536 * sizeof(cf));
    if (!a.scratch) return -1;
    for (int k = 0; k < 2; k++) {
      const int n1 = k ? 256 : 128;
      char rname[32];
      snprintf(rname, sizeof rname, "bigroots%d", n1 * 256);
      auto b = c->bufs.find(rname);
      a.roots[k] = b == c->bufs.end() ? nullptr : (const cf *)b->second.p;
    }
  }
  a.out_idx = idx;
  a.out_val = val;
  hipLaunchKernelGGL(q_zero_kernel, dim3(1), dim3(64), 0, c->stream, counter);
  if (q->ntiles == 0) return hip_ok(hipGetLastError(), "q_zero_kernel") ? 0 : -1;
  {
    Ctx::ProfScope ps(c, "queue");
    hipLaunchKernelGGL(queue_kernel, dim3(grid), dim3(kQThreads), kQLds, c->stream, a);
  }
  return hip_ok(hipGetLastError(), "queue_kernel") ? 0 : -1;
}

// queue_launch through a graph cache: captured on the second identical call
// (same plan, pool and outputs), replayed afterwards.
static int queue_run_cached(Queue *q, int64_t *idx, cf *val) {
  Ctx *c = &q->c;
  if (!c->use_graphs || c->prof_on) return queue_launch(q, idx, val);
  std::vector<long long> key = {q->plan_gen, (long long)(intptr_t)q->pool, (long long)(intptr_t)idx,
                                (long long)(intptr_t)val, q->grid_per_cu};
  if (c->graph_exec && key == c->graph_key) {
    return hip_ok(hipGraphLaunch(c->graph_exec, c->stream), "hipGraphLaunch") ? 0 : -1;
  }
  if (key != c->last_key) {
    if (queue_launch(q, idx, val) < 0) return -1;
    c->last_key = key;
    return 0;
  }
  if (c->graph_exec) {
    hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  hipGraph_t g = nullptr;
  if (!hip_ok(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
    return -1;
  c->capturing = true;
  const int rc = queue_launch(q, idx, val);
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
  c->graph_exec = ex;
  c->graph_key = key;
  return hip_ok(hipGraphLaunch(c->graph_exec, c->stream), "hipGraphLaunch") ? 0 : -1;
}

}  // namespace mfh

extern "C" int mfh_queue_run_async(mfh_queue *queue, int64_t *dev_idx, void *dev_val_c64) {
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  if (!q->planned) return fail("no plan: call mfh_queue_plan first");
  if (q->nout && (!dev_idx || !dev_val_c64)) return fail("null output pointer");
  hipSetDevice(q->c.dev);
  // the scratch that the plan's items live in must not move under a graph
  return mfh::queue_run_cached(q, dev_idx, (cf *)dev_val_c64);
}

extern "C" long long mfh_queue_run(mfh_queue *queue, int64_t *idx, void *val_c64, int32_t *counts) {
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  if (!q->planned) return fail("no plan: call mfh_queue_plan first");
  if (q->nout && (!idx || !val_c64)) return fail("null output pointer");
  hipSetDevice(q->c.dev);
  const size_t cells = (size_t)q->nout;
  int64_t *didx = (int64_t *)q->c.scratch("q_out_idx", std::max<size_t>(1, cells) * sizeof(int64_t));
  cf *dval = (cf *)q->c.scratch("q_out_val", std::max<size_t>(1, cells) * sizeof(cf));
  if (!didx || !dval) return -1;
  if (mfh::queue_run_cached(q, didx, dval) < 0) return -1;
  if (cells && (!hip_ok(hipMemcpyAsync(idx, didx, cells * sizeof(int64_t), hipMemcpyDeviceToHost, q->c.stream),
                        "copy idx") ||
                !hip_ok(hipMemcpyAsync(val_c64, dval, cells * sizeof(cf), hipMemcpyDeviceToHost, q->c.stream),
                        "copy val")))
    return -1;
  if (!hip_ok(hipStreamSynchronize(q->c.stream), "hipStreamSynchronize")) return -1;
  q->c.prof_collect();
  long long total = 0;
  long long p = 0;
  for (int r = 0; r < q->nreq; r++) {
    const long long nb = q->req_nb[r];
    for (long long j = 0; j < q->req_nc[r]; j++, p++) {
      int cnt = 0;
      const int64_t *row = idx + q->req_out[r] + j * nb;
      for (long long b = 0; b < nb; b++) cnt += row[b] >= 0;
      if (counts) counts[p] = cnt;
      total += cnt;
    }
  }
  return total;
}

extern "C" int mfh_queue_set_tuning(mfh_queue *queue, const char *key, long value) {
  if (!queue || !key) return fail("null argument");
  Queue *q = reinterpret_cast<Queue *>(queue);
  if (!strcmp(key, "grid_per_cu")) {
    if (value < 1 || value > 64) return fail("grid_per_cu must be in [1, 64]");
    q->grid_per_cu = (int)value;
    return 0;
  }
  if (!strcmp(key, "graphs")) {
    q->c.use_graphs = value != 0;
    return 0;
  }
  return fail("unknown queue tuning key '%s'", key);
}

extern "C" int mfh_queue_sync(mfh_queue *queue) {
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  hipSetDevice(q->c.dev);
  return hip_ok(hipStreamSynchronize(q->c.stream), "hipStreamSynchronize") ? 0 : -1;
}

extern "C" void *mfh_queue_stream(mfh_queue *queue) {
  return queue ? (void *)reinterpret_cast<Queue *>(queue)->c.stream : nullptr;
}

extern "C" int mfh_queue_prof_enable(mfh_queue *queue, int on) {
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  q->c.prof_collect();
  q->c.prof_on = on != 0;
  q->c.prof_totals.clear();
  return 0;
}

extern "C" int mfh_queue_prof_read(mfh_queue *queue, double *total_ms, long *count) {
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  q->c.prof_collect();
  auto it = q->c.prof_totals.find("queue");
  if (total_ms) *total_ms = it == q->c.prof_totals.end() ? 0.0 : it->second.first;
  if (count) *count = it == q->c.prof_totals.end() ? 0 : it->second.second;
  return 0;
}
