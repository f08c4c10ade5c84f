// Uneven batched refine: one persistent launch drains a device work queue of
// (data spectrum, template spectrum) pairs whose transform lengths differ.
//
// Work queue.  A plan turns requests (one data spectrum, c templates, a
// length n, binsize, threshold and window each) into items, one per pair.
// Pairs of length n <= 16384 are grouped into tiles of one length, as many
// pairs as one 1024-lane workgroup carries side by side (1024 / (n/16)).
// Tiles are ordered longest first, so the last tiles to be taken are the
// shortest and the tail of the launch stays short.  Every workgroup takes the
// next tile with one atomic on a device counter, runs it and repeats until
// the counter passes the tile count, so each workgroup reaches the exit.  One
// workgroup of 1024 lanes per CU.
//
// Four-step lengths (32768, 65536) need a round trip through scratch between
// their column and row transforms; inside one resident workgroup that round
// trip is exposed latency (measured: time per pair independent of how many
// workgroups share the device).  Their pairs therefore go, as a device list
// with per-pair pool and output offsets, to the two-kernel four-step path of
// the flat filter (columns of 16 lags per workgroup, then rows), queued just
// before the persistent launch in the same stream and graph.
//
// Arithmetic.  Every pair runs the same product, transforms, twiddle tables
// and exact key rule as the per-request flat filter, so indices and values
// equal it bit for bit.
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
constexpr int kQMaxLog2 = 14;  // longest length of the persistent kernel
// exchange space of a tile: 1024 lanes x 16 points, padded by 1/16, plus one
// element per slot where slots share a wave (largest at n = 64: 256 slots)
template <int N>
constexpr size_t q_xch_bytes() {
  return (size_t)(kQThreads / FlatShape<N>::TPT) * FlatShape<N>::XSTRIDE * sizeof(cf);
}
constexpr size_t kQLds = q_xch_bytes<64>();
static_assert(q_xch_bytes<128>() <= kQLds && q_xch_bytes<4096>() <= kQLds && q_xch_bytes<16384>() <= kQLds, "");

struct QArgs {
  const cf *pool;
  const QItem *items;
  const QTile *tiles;
  int ntiles;
  int *counter;
  const cf *tw[kQMaxLog2 + 1];  // flat twiddle tables by log2 length
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
      runs8_max_to_bins<TPT>(v, (uint32_t)t, (uint32_t)(t - lane), lane, it.seed, it.bshift, bkey);
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
  // four-step pairs, grouped by (length, binsize, threshold, window): each
  // group is one list-mode launch of the two-kernel path
  struct BigGroup {
    int n;
    long long binsize;
    float threshold;
    long long ws, we;
    int first, count;  // entries in the device list
  };
  std::vector<BigGroup> big;
  std::vector<long long> req_out;  // output offset per request
  std::vector<int> req_nb, req_nc;
  int grid_per_cu = 1;
  int grid_cap = 0;  // if > 0: at most this many persistent workgroups
  long prof_runs = 0;  // runs recorded while event timing is on
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
  q->c.use_graphs = true;  // replay and direct launches measured within 1%
  hipDeviceProp_t prop;
  if (!hip_ok(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties") ||
      !hip_ok(hipStreamCreateWithFlags(&q->c.stream, hipStreamNonBlocking), "hipStreamCreate")) {
    delete q;
    return nullptr;
  }
  q->cus = prop.multiProcessorCount;
  if (!mfh::lds_limit((const void *)mfh::queue_kernel, q->c.dev, mfh::kQLds)) {
    delete q;
    return nullptr;
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
  std::vector<Queue::BigGroup> bigs;
  std::vector<std::vector<std::pair<int2, long long>>> big_pairs;
  q->req_out.assign(nreq, 0);
  q->req_nb.assign(nreq, 0);
  q->req_nc.assign(nreq, 0);
  long long nout = 0, npairs = 0;
  for (int r = 0; r < nreq; r++) {
    const long long len = n[r];
    if (len < 64 || len > 1048576 || (len & (len - 1)))
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
    if (lg >= 15) {
      // four-step: one list entry per pair, grouped by launch parameters
      if (data_off[r] > INT32_MAX || tmpl_off[r] + (long long)ncand[r] * len > INT32_MAX)
        return fail("request %d: four-step spectra must start below pool element 2^31", r);
      size_t g = 0;
      for (; g < bigs.size(); g++) {
        const auto &G = bigs[g];
        if (G.n == (int)len && G.binsize == binsize[r] && !memcmp(&G.threshold, &thr, sizeof thr) &&
            G.ws == ws[r] && G.we == we[r])
          break;
      }
      if (g == bigs.size()) {
        bigs.push_back({(int)len, (long long)binsize[r], thr, (long long)ws[r], (long long)we[r], 0, 0});
        big_pairs.emplace_back();
      }
      for (int j = 0; j < ncand[r]; j++)
        big_pairs[g].push_back({int2{(int)data_off[r], (int)(tmpl_off[r] + (long long)j * len)},
                                nout + (long long)j * nb});
    } else {
      QItem base{};
      base.d_off = data_off[r];
      base.seed = thr > 0.f ? (float)((double)thr * (double)thr) : 0.f;
      base.binsize = (uint32_t)binsize[r];
      const bool pow2 = (binsize[r] & (binsize[r] - 1)) == 0;
      base.bshift = pow2 ? ilog2((int)binsize[r]) : -1;
      base.inv_bs = 1.0f / (float)binsize[r];
      base.ws = (uint32_t)ws[r];
      base.we = (uint32_t)we[r];
      base.nb = (uint32_t)nb;
      for (int j = 0; j < ncand[r]; j++) {
        QItem it = base;
        it.t_off = tmpl_off[r] + (long long)j * len;
        it.o_off = nout + (long long)j * nb;
        by_class[lg].push_back(it);
      }
    }
    nout += (long long)ncand[r] * nb;
    npairs += ncand[r];
  }
  std::vector<QItem> items;
  std::vector<QTile> tiles;
  for (int lg = kQMaxLog2; lg >= 6; lg--) {  // longest first
    auto &v = by_class[lg];
    if (v.empty()) continue;
    const int slots = kQThreads / ((1 << lg) / 16);
    for (size_t i = 0; i < v.size(); i += slots) {
      QTile t{};
      t.log2n = lg;
      t.first = (int)(items.size() + i);
      t.count = (int)std::min<size_t>(slots, v.size() - i);
      tiles.push_back(t);
    }
    items.insert(items.end(), v.begin(), v.end());
  }
  std::vector<int2> blist;
  std::vector<long long> booff;
  std::vector<int> bcount;
  for (size_t g = 0; g < bigs.size(); g++) {
    bigs[g].first = (int)blist.size();
    bigs[g].count = (int)big_pairs[g].size();
    bcount.push_back(bigs[g].count);
    for (auto &e : big_pairs[g]) {
      blist.push_back(e.first);
      booff.push_back(e.second);
    }
  }
  if (items.size() > (size_t)INT32_MAX || blist.size() > (size_t)INT32_MAX) return fail("too many pairs in one plan");
  hipSetDevice(q->c.dev);
  // tables for every length in the plan
  for (int lg = 6; lg <= kQMaxLog2; lg++)
    if (!by_class[lg].empty() && !q->c.twiddles((size_t)1 << lg)) return -1;
  QItem *ditems = (QItem *)q->c.scratch("q_items", std::max<size_t>(1, items.size()) * sizeof(QItem));
  QTile *dtiles = (QTile *)q->c.scratch("q_tiles", std::max<size_t>(1, tiles.size()) * sizeof(QTile));
  int2 *dblist = (int2 *)q->c.scratch("q_big_list", std::max<size_t>(1, blist.size()) * sizeof(int2));
  long long *dbooff = (long long *)q->c.scratch("q_big_ooff", std::max<size_t>(1, booff.size()) * sizeof(long long));
  int *dbcount = (int *)q->c.scratch("q_big_count", std::max<size_t>(1, bcount.size()) * sizeof(int));
  if (!ditems || !dtiles || !dblist || !dbooff || !dbcount) return -1;
  auto up = [&](void *dst, const void *src, size_t bytes, const char *what) {
    return bytes == 0 || hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, q->c.stream), what);
  };
  if (!up(ditems, items.data(), items.size() * sizeof(QItem), "copy items") ||
      !up(dtiles, tiles.data(), tiles.size() * sizeof(QTile), "copy tiles") ||
      !up(dblist, blist.data(), blist.size() * sizeof(int2), "copy four-step list") ||
      !up(dbooff, booff.data(), booff.size() * sizeof(long long), "copy four-step offsets") ||
      !up(dbcount, bcount.data(), bcount.size() * sizeof(int), "copy four-step counts") ||
      !hip_ok(hipStreamSynchronize(q->c.stream), "hipStreamSynchronize"))
    return -1;
  q->big = bigs;
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
  if (c->prof_on && !c->capturing) q->prof_runs++;
  // four-step pairs first: list-mode launches of the two-kernel path
  const int2 *blist = (const int2 *)c->bufs["q_big_list"].p;
  const long long *booff = (const long long *)c->bufs["q_big_ooff"].p;
  const int *bcount = (const int *)c->bufs["q_big_count"].p;
  for (size_t g = 0; g < q->big.size(); g++) {
    const Queue::BigGroup &G = q->big[g];
    FlatRun r;
    r.data = q->pool;  // list entries are element offsets into the pool
    r.tmpl = q->pool;
    r.data_stride = 1;
    r.tmpl_stride = 1;
    r.nd = 1;
    r.nt = 1;
    r.binsize = (size_t)G.binsize;
    r.threshold = G.threshold;
    r.ws = (size_t)G.ws;
    r.we = (size_t)G.we;
    r.out_idx = idx;
    r.out_val = (cf *)val;
    r.list = blist + G.first;
    r.list_count = bcount + g;
    r.list_ooff = booff + G.first;
    r.list_cap = G.count;
    r.list_prof_name = "queue_big";
    if (launch_big(c, (size_t)G.n, r) < 0) return -1;
  }
  QArgs a{};
  a.pool = q->pool;
  a.items = (const QItem *)c->bufs["q_items"].p;
  a.tiles = (const QTile *)c->bufs["q_tiles"].p;
  a.ntiles = q->ntiles;
  a.counter = counter;
  for (int lg = 6; lg <= kQMaxLog2; lg++) {
    auto it = c->tw_by_n.find((size_t)1 << lg);
    a.tw[lg] = it == c->tw_by_n.end() ? nullptr : it->second;
  }
  int grid = std::max(1, std::min(q->cus * q->grid_per_cu, q->ntiles));
  if (q->grid_cap > 0) grid = std::min(grid, q->grid_cap);
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
                                (long long)(intptr_t)val, q->grid_per_cu, q->grid_cap, q->c.big_chunk_bytes};
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
  if (!strcmp(key, "grid_cap")) {
    if (value < 0) return fail("grid_cap must be >= 0");
    q->grid_cap = (int)value;
    return 0;
  }
  if (!strcmp(key, "graphs")) {
    q->c.use_graphs = value != 0;
    return 0;
  }
  if (!strcmp(key, "big_chunk_mb")) {
    if (value < 1 || value > 65536) return fail("big_chunk_mb must be in [1, 65536]");
    q->c.big_chunk_bytes = (long long)value << 20;
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
  q->prof_runs = 0;
  return 0;
}

extern "C" int mfh_queue_prof_read(mfh_queue *queue, double *total_ms, long *count) {
  if (!queue) return fail("null queue");
  Queue *q = reinterpret_cast<Queue *>(queue);
  q->c.prof_collect();
  double ms = 0.0;
  for (const char *name : {"queue", "queue_big"}) {
    auto it = q->c.prof_totals.find(name);
    if (it != q->c.prof_totals.end()) ms += it->second.first;
  }
  if (total_ms) *total_ms = ms;
  if (count) *count = q->prof_runs;
  return 0;
}
