// C ABI of the MI355X matched filter: contexts, ingest, dispatch, timing.
// Declarations and the reference interface each replaces: include/mfhip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "mfhip.h"
#include "mfh_internal.h"

using mfh::cf;

namespace {
thread_local std::string g_err;
}

namespace mfh {

int fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  fail("%s: %s", what, hipGetErrorString(e));
  return false;
}

bool lds_limit(const void *fn, int dev, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<const void *, int>, size_t> done;
  std::lock_guard<std::mutex> lock(mu);
  size_t &have = done[{fn, dev}];
  if (bytes <= have) return true;
  if (!hip_ok(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes), "hipFuncSetAttribute"))
    return false;
  have = bytes;
  return true;
}

Ctx::~Ctx() {
  if (dev >= 0) hipSetDevice(dev);
  if (stream) hipStreamSynchronize(stream);
  if (graph_exec) hipGraphExecDestroy(graph_exec);
  drop_hier_graphs();
  for (auto &kv : bufs) hipFree(kv.second.p);
  for (auto &e : prof_events) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  for (auto e : event_pool) hipEventDestroy(e);
  if (order_ev) hipEventDestroy(order_ev);
  if (stream) hipStreamDestroy(stream);
  if (stage) hipHostFree(stage);
  if (refine_hint) hipHostFree(refine_hint);
}

void *Ctx::host_stage(size_t bytes) {
  if (stage && stage_bytes >= bytes) return stage;
  if (stage) {
    hipStreamSynchronize(stream);
    hipHostFree(stage);
    stage = nullptr;
    stage_bytes = 0;
  }
  // coherent: the readback kernel's stores reach host memory without
  // stopping in the GPU's caches
  if (!hip_ok(hipHostMalloc(&stage, bytes ? bytes : 1, hipHostMallocCoherent), "hipHostMalloc")) {
    stage = nullptr;
    return nullptr;
  }
  stage_bytes = bytes;
  return stage;
}

bool peak_buffer(Ctx *c, size_t cells, PeakBuf &b) {
  char *p = (char *)c->scratch("out_peaks", cells * (sizeof(int64_t) + sizeof(cf)));
  if (!p) return false;
  b.idx = (int64_t *)p;
  b.val = (cf *)(p + cells * sizeof(int64_t));
  return true;
}

namespace {

// A call's device results written straight into page-locked host memory in
// the layout the caller asked for: 16-byte {int64 index; complex64 value}
// records, or [indices | values].  A kernel queued behind the filter takes
// the place of a copy-engine transfer, whose start and completion cost more
// than moving these few hundred KiB.
__global__ void __launch_bounds__(256) results_to_host(const int64_t *__restrict__ idx, const cf *__restrict__ val,
                                                       long long cells, int records, char *__restrict__ dst) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= cells) return;
  const long long k = idx[i];
  const cf v = val[i];
  if (records) {
    int4 w;
    w.x = (int)(unsigned)(unsigned long long)k;
    w.y = (int)(k >> 32);
    w.z = __float_as_int(v.x);
    w.w = __float_as_int(v.y);
    reinterpret_cast<int4 *>(dst)[i] = w;
  } else {
    reinterpret_cast<long long *>(dst)[i] = k;
    reinterpret_cast<cf *>(dst + cells * sizeof(long long))[i] = v;
  }
}

}  // namespace

// Bins set (index >= 0) among `cells` indices `step` int64 words apart;
// counts[p] (if set) per pair of nb bins.  Flat loops the compiler
// vectorises: a loop nest over the run-time bin count measured about 10 us
// per 16384 cells on the host.
long count_set(const int64_t *si, size_t step, size_t cells, size_t nb, int32_t *counts) {
  long total = 0;
  for (size_t i = 0; i < cells; i++) total += si[i * step] >= 0;
  if (counts && nb == 1) {
    for (size_t p = 0; p < cells; p++) counts[p] = si[p * step] >= 0;
  } else if (counts) {
    for (size_t p = 0; p < cells / nb; p++) {
      int cnt = 0;
      for (size_t k = 0; k < nb; k++) cnt += si[(p * nb + k) * step] >= 0;
      counts[p] = cnt;
    }
  }
  return total;
}

// Page-locked host buffers handed out by mfh_host_alloc: start -> bytes.
std::mutex host_blocks_mu;
std::map<uintptr_t, size_t> host_blocks;

// Whether [p, p + bytes) lies inside one mfh_host_alloc buffer, which the
// readback kernel may then write directly.
bool host_block(const void *p, size_t bytes) {
  const uintptr_t a = (uintptr_t)p;
  std::lock_guard<std::mutex> lk(host_blocks_mu);
  auto it = host_blocks.upper_bound(a);
  if (it == host_blocks.begin()) return false;
  --it;
  return a - it->first + bytes <= it->second;
}

long read_peaks(Ctx *c, const PeakBuf &b, size_t cells, size_t nb, int64_t *idx, void *val_c64, int32_t *counts,
                void *rec) {
  const size_t bi = cells * sizeof(int64_t), bytes = bi + cells * sizeof(cf);
  if (rec && cells && host_block(rec, cells * 16)) {
    // the records go straight into the caller's page-locked buffer
    hipLaunchKernelGGL(results_to_host, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream, b.idx,
                       b.val, (long long)cells, 1, (char *)rec);
    if (!hip_ok(hipGetLastError(), "results_to_host") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
      return -1;
    return count_set(reinterpret_cast<const int64_t *>(rec), 2, cells, nb, counts);
  }
  if (!rec && cells && idx && (char *)val_c64 == (char *)idx + bi && host_block(idx, bytes)) {
    // [indices | values] straight into the caller's page-locked buffer
    hipLaunchKernelGGL(results_to_host, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream, b.idx,
                       b.val, (long long)cells, 0, (char *)idx);
    if (!hip_ok(hipGetLastError(), "results_to_host") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
      return -1;
    return count_set(idx, 1, cells, nb, counts);
  }
  char *st = (char *)c->host_stage(bytes);
  std::vector<char> pageable;
  const bool staged = st != nullptr;  // the readback kernel fills the staging in the caller's layout
  if (staged) {
    if (cells) {
      hipLaunchKernelGGL(results_to_host, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream, b.idx,
                         b.val, (long long)cells, rec ? 1 : 0, st);
      if (!hip_ok(hipGetLastError(), "results_to_host")) return -1;
    }
    if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return -1;
  } else if (!rec) {
    // no page-locked staging to be had: copy straight into the caller's
    // (pageable) arrays, one copy each
    (void)hipGetLastError();
    if (!hip_ok(hipMemcpyAsync(idx, b.idx, bi, hipMemcpyDeviceToHost, c->stream), "copy idx") ||
        !hip_ok(hipMemcpyAsync(val_c64, b.val, bytes - bi, hipMemcpyDeviceToHost, c->stream), "copy val") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
      return -1;
  } else {
    // records without page-locked staging: through a pageable host copy
    (void)hipGetLastError();
    pageable.resize(bytes);
    st = pageable.data();
    if (!hip_ok(hipMemcpyAsync(st, b.idx, bytes, hipMemcpyDeviceToHost, c->stream), "copy results") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
      return -1;
  }
  // the indices to count, `step` int64 words apart
  const int64_t *si = st ? reinterpret_cast<const int64_t *>(st) : idx;
  size_t step = 1;
  if (rec && !staged) {
    // one pass from the pageable copy into {int64 index; complex64 value}
    // records (16 bytes each)
    const cf *sv = reinterpret_cast<const cf *>(st + bi);
    char *r = static_cast<char *>(rec);
    for (size_t i = 0; i < cells; i++) {
      memcpy(r + 16 * i, si + i, 8);
      memcpy(r + 16 * i + 8, sv + i, 8);
    }
    si = reinterpret_cast<const int64_t *>(rec);
    step = 2;
  } else if (rec) {
    // the staging already holds the records
    if (cells) memcpy(rec, st, cells * 16);
    si = reinterpret_cast<const int64_t *>(rec);
    step = 2;
  } else if (staged && cells) {
    memcpy(idx, st, bi);
    memcpy(val_c64, st + bi, bytes - bi);
    si = idx;
  }
  return count_set(si, step, cells, nb, counts);
}

const cf *Ctx::twiddles(size_t len) {
  auto it = tw_by_n.find(len);
  if (it != tw_by_n.end()) return it->second;
  char name[32];
  snprintf(name, sizeof name, "twiddle%zu", len);
  cf *tw = (cf *)scratch(name, (flat_twiddle_count(len) + 1) * sizeof(cf));
  if (!tw || flat_fill_twiddles(this, len, tw) < 0) return nullptr;
  tw_by_n[len] = tw;
  return tw;
}

void *Ctx::scratch(const char *name, size_t bytes) {
  auto found = bufs.find(name);
  if (found != bufs.end() && found->second.bytes >= bytes && found->second.p) return found->second.p;
  if (no_alloc) {
    fail("an allocation-free call needs the context's '%s' buffer (%zu bytes), which is not built yet: "
         "run mfh_hier_peaks or mfh_hier_peaks_async with this band and precision once first",
         name, bytes);
    return nullptr;
  }
  Buf &b = bufs[name];
  if (b.p) {
    hipStreamSynchronize(stream);
    hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    // a captured graph may hold the old address
    if (graph_exec) {
      hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
    drop_hier_graphs();
    graph_key.clear();
  }
  if (!hip_ok(hipMalloc(&b.p, bytes ? bytes : 1), "hipMalloc")) return nullptr;
  b.bytes = bytes;
  return b.p;
}

int Ctx::own_bank(bool templates) {
  cf *&cur = templates ? tmpl : data;
  cf *own = templates ? tmpl_own : data_own;
  if (cur == own) return 0;
  const size_t bytes = (size_t)(templates ? nt : nd) * n * sizeof(cf);
  hipSetDevice(dev);
  if (!hip_ok(hipMemcpyAsync(own, cur, bytes, hipMemcpyDeviceToDevice, stream), "copy bound bank")) return -1;
  cur = own;
  if (templates) coarse_dirty = true;
  return 0;
}

hipEvent_t Ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

// Event pair around one launch, when profiling is on (events are pooled).
Ctx::ProfScope::ProfScope(Ctx *c, const char *name) : c_(c), name_(name) {
  if (!c_->prof_on || c_->capturing || c_->no_alloc) return;
  ev_.a = c_->take_event();
  ev_.b = c_->take_event();
  hipEventRecord(ev_.a, c_->stream);
}
Ctx::ProfScope::~ProfScope() {
  if (!c_->prof_on || c_->capturing || c_->no_alloc) return;
  hipEventRecord(ev_.b, c_->stream);
  ev_.name = name_;
  c_->prof_events.push_back(ev_);
}

void Ctx::prof_collect() {
  if (prof_events.empty()) return;
  hipStreamSynchronize(stream);
  for (auto &e : prof_events) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, e.a, e.b);
    auto &acc = prof_totals[e.name];
    acc.first += ms;
    acc.second += 1;
    event_pool.push_back(e.a);
    event_pool.push_back(e.b);
  }
  prof_events.clear();
}

}  // namespace mfh

using mfh::Ctx;
using mfh::fail;
using mfh::hip_ok;

// ---------------------------------------------------------------- the ABI

extern "C" const char *mfh_last_error(void) { return g_err.c_str(); }

extern "C" int mfh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" int mfh_device_info(int index, char *name, size_t name_len, char *arch, size_t arch_len) {
  hipDeviceProp_t p;
  if (!hip_ok(hipGetDeviceProperties(&p, index), "hipGetDeviceProperties")) return -1;
  if (name && name_len) snprintf(name, name_len, "%s", p.name);
  if (arch && arch_len) snprintf(arch, arch_len, "%s", p.gcnArchName);
  return 0;
}

extern "C" int mfh_supported(size_t n) { return mfh::flat_supported(n) ? 1 : 0; }

extern "C" mfh_ctx *mfh_create(int device, size_t n, int nd, int nt) {
  if (!mfh::flat_supported(n)) {
    fail("transform length %zu is not supported", n);
    return nullptr;
  }
  if (nd < 1 || nt < 1) {
    fail("ndata and ntemplates must be >= 1");
    return nullptr;
  }
  int count = mfh_device_count();
  if (device < 0 || device >= count) {
    fail("no HIP device with index %d (%d visible)", device, count);
    return nullptr;
  }
  if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return nullptr;
  Ctx *c = new Ctx;
  c->dev = device;
  c->n = n;
  c->nd = nd;
  c->nt = nt;
  if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) {
    delete c;
    return nullptr;
  }
  c->data = c->data_own = (cf *)c->scratch("data", (size_t)nd * n * sizeof(cf));
  c->tmpl = c->tmpl_own = (cf *)c->scratch("tmpl", (size_t)nt * n * sizeof(cf));
  if (!c->data || !c->tmpl || (n <= 16384 && !c->twiddles(n))) {
    delete c;
    return nullptr;
  }
  // unset rows read as zero rather than as whatever the allocator left
  hipMemsetAsync(c->data, 0, (size_t)nd * n * sizeof(cf), c->stream);
  hipMemsetAsync(c->tmpl, 0, (size_t)nt * n * sizeof(cf), c->stream);
  if (!hip_ok(hipStreamSynchronize(c->stream), "create")) {
    delete c;
    return nullptr;
  }
  return reinterpret_cast<mfh_ctx *>(c);
}

extern "C" void mfh_destroy(mfh_ctx *ctx) { delete reinterpret_cast<Ctx *>(ctx); }

extern "C" void *mfh_stream(mfh_ctx *ctx) {
  return ctx ? (void *)reinterpret_cast<Ctx *>(ctx)->stream : nullptr;
}

extern "C" void *mfh_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (!hip_ok(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocCoherent), "hipHostMalloc")) {
    (void)hipGetLastError();  // a later launch check must not see this failure
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mfh::host_blocks_mu);
  mfh::host_blocks[(uintptr_t)p] = bytes ? bytes : 1;
  return p;
}

extern "C" int mfh_host_free(void *p) {
  if (!p) return 0;
  {
    std::lock_guard<std::mutex> lk(mfh::host_blocks_mu);
    auto it = mfh::host_blocks.find((uintptr_t)p);
    if (it == mfh::host_blocks.end()) return fail("not a buffer of mfh_host_alloc");
    mfh::host_blocks.erase(it);
  }
  return hip_ok(hipHostFree(p), "hipHostFree") ? 0 : -1;
}

extern "C" int mfh_sync(mfh_ctx *ctx) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  hipSetDevice(c->dev);
  return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ? 0 : -1;
}

extern "C" int mfh_stream_wait(mfh_ctx *ctx, void *producer_stream) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  hipSetDevice(c->dev);
  // a producer stream with nothing left to run has finished every write the
  // caller ordered before this hand-over: no wait to enqueue (the cross-queue
  // wait costs tens of microseconds of host time per call)
  const hipError_t q = hipStreamQuery((hipStream_t)producer_stream);
  if (q == hipSuccess) return 0;
  if (q != hipErrorNotReady) return hip_ok(q, "hipStreamQuery") ? 0 : -1;
  (void)hipGetLastError();  // not-ready is an answer, not an error for later checks
  if (!c->order_ev && !hip_ok(hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming), "hipEventCreate"))
    return -1;
  // a wait enqueued here holds the event's state at this call, so the event
  // can be recorded again by the next hand-over
  return hip_ok(hipEventRecord(c->order_ev, (hipStream_t)producer_stream), "hipEventRecord") &&
                 hip_ok(hipStreamWaitEvent(c->stream, c->order_ev, 0), "hipStreamWaitEvent")
             ? 0
             : -1;
}

// Rows [row0, row0 + rows) of the data (templates = false) or template bank
// from `src`.  A bank bound to a caller's buffer is first copied into the
// context's own allocation, so the library never writes the caller's memory.
static int set_rows(Ctx *c, bool templates, int row0, int rows, const void *src, hipMemcpyKind kind, bool sync) {
  const int total = templates ? c->nt : c->nd;
  if (row0 < 0 || rows < 1 || row0 + rows > total) return fail("rows [%d, %d) out of range [0, %d)", row0, row0 + rows, total);
  if (!src) return fail("null spectrum pointer");
  hipSetDevice(c->dev);
  const bool whole = row0 == 0 && rows == total;
  if (whole) {
    // every row is replaced: nothing of a bound bank needs to survive
    if (templates)
      c->tmpl = c->tmpl_own;
    else
      c->data = c->data_own;
  } else if (c->own_bank(templates) < 0) {
    return -1;
  }
  if (templates) c->coarse_dirty = true;
  cf *base = templates ? c->tmpl : c->data;
  if (!hip_ok(hipMemcpyAsync(base + (size_t)row0 * c->n, src, (size_t)rows * c->n * sizeof(cf), kind, c->stream),
              "hipMemcpyAsync"))
    return -1;
  if (sync && !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return -1;
  return 0;
}

extern "C" int mfh_set_data(mfh_ctx *ctx, int row0, int rows, const void *host_c64) {
  if (!ctx) return fail("null context");
  return set_rows(reinterpret_cast<Ctx *>(ctx), false, row0, rows, host_c64, hipMemcpyHostToDevice, true);
}
extern "C" int mfh_set_templates(mfh_ctx *ctx, int row0, int rows, const void *host_c64) {
  if (!ctx) return fail("null context");
  return set_rows(reinterpret_cast<Ctx *>(ctx), true, row0, rows, host_c64, hipMemcpyHostToDevice, true);
}
extern "C" int mfh_set_data_device_async(mfh_ctx *ctx, int row0, int rows, const void *dev_c64) {
  if (!ctx) return fail("null context");
  return set_rows(reinterpret_cast<Ctx *>(ctx), false, row0, rows, dev_c64, hipMemcpyDeviceToDevice, false);
}
extern "C" int mfh_set_templates_device_async(mfh_ctx *ctx, int row0, int rows, const void *dev_c64) {
  if (!ctx) return fail("null context");
  return set_rows(reinterpret_cast<Ctx *>(ctx), true, row0, rows, dev_c64, hipMemcpyDeviceToDevice, false);
}

extern "C" int mfh_get_rows(mfh_ctx *ctx, int templates, int row0, int rows, void *host_c64) {
  if (!ctx || !host_c64) return fail("null argument");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  const int total = templates ? c->nt : c->nd;
  if (row0 < 0 || rows < 1 || row0 + rows > total) return fail("rows [%d, %d) out of range [0, %d)", row0, row0 + rows, total);
  hipSetDevice(c->dev);
  const cf *base = templates ? c->tmpl : c->data;
  if (!hip_ok(hipMemcpyAsync(host_c64, base + (size_t)row0 * c->n, (size_t)rows * c->n * sizeof(cf),
                             hipMemcpyDeviceToHost, c->stream),
              "hipMemcpyAsync"))
    return -1;
  return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ? 0 : -1;
}

// Point a bank at the caller's device buffer (read in place, never written).
static int bind_bank(Ctx *c, bool templates, const void *dev_c64) {
  if (!dev_c64) return fail("null spectrum pointer");
  if ((uintptr_t)dev_c64 % alignof(cf)) return fail("bound spectra must be 8-byte aligned");
  hipPointerAttribute_t attr;
  if (!hip_ok(hipPointerGetAttributes(&attr, dev_c64), "hipPointerGetAttributes")) return -1;
  if (attr.type != hipMemoryTypeDevice || attr.device != c->dev)
    return fail("bound spectra must be device memory of device %d", c->dev);
  if (templates) {
    c->tmpl = (cf *)dev_c64;
    c->coarse_dirty = true;
  } else {
    c->data = (cf *)dev_c64;
  }
  return 0;
}

extern "C" int mfh_bind_data(mfh_ctx *ctx, const void *dev_c64) {
  if (!ctx) return fail("null context");
  return bind_bank(reinterpret_cast<Ctx *>(ctx), false, dev_c64);
}
extern "C" int mfh_bind_templates(mfh_ctx *ctx, const void *dev_c64) {
  if (!ctx) return fail("null context");
  return bind_bank(reinterpret_cast<Ctx *>(ctx), true, dev_c64);
}

extern "C" size_t mfh_nbins(size_t n, size_t binsize, size_t ws, size_t we) {
  if (!binsize) return 0;
  if (we > n) we = n;
  if (ws >= we) return 0;
  return (we - ws + binsize - 1) / binsize;
}

extern "C" int mfh_flat_peaks_async(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold,
                                    size_t ws, size_t we, int64_t *dev_idx, void *dev_val_c64) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  mfh::FlatRun r;
  r.data = c->data + (size_t)d0 * c->n;
  r.tmpl = c->tmpl + (size_t)t0 * c->n;
  r.data_stride = (long long)c->n;
  r.tmpl_stride = (long long)c->n;
  if (d0 < 0 || t0 < 0 || nd < 1 || nt < 1 || d0 + nd > c->nd || t0 + nt > c->nt)
    return fail("data/templates sub-range out of bounds");
  r.nd = nd;
  r.nt = nt;
  r.binsize = binsize;
  r.threshold = threshold;
  r.ws = ws;
  r.we = we;
  r.out_idx = dev_idx;
  r.out_val = (cf *)dev_val_c64;
  return mfh::launch_flat(c, c->n, r);
}

static long flat_peaks_host(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold, size_t ws,
                            size_t we, int64_t *idx, void *val_c64, void *rec, int32_t *counts) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!rec && (!idx || !val_c64)) return fail("null output pointer");
  const size_t nb = mfh_nbins(c->n, binsize, ws, we);
  if (binsize < 1) return fail("binsize must be >= 1");
  if (nb == 0) return fail("empty window (%zu, %zu)", ws, we);
  const size_t cells = (size_t)nd * (size_t)nt * nb;
  hipSetDevice(c->dev);
  if (!rec && cells && (char *)val_c64 == (char *)idx + cells * sizeof(int64_t) &&
      mfh::host_block(idx, cells * (sizeof(int64_t) + sizeof(cf)))) {
    // [indices | values] in one page-locked mfh_host_alloc buffer: the
    // filter writes the caller's arrays itself, with no readback step
    if (mfh_flat_peaks_async(ctx, d0, nd, t0, nt, binsize, threshold, ws, we, idx, val_c64) < 0 ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
      return -1;
    c->prof_collect();
    return mfh::count_set(idx, 1, cells, nb, counts);
  }
  mfh::PeakBuf pb;
  if (!mfh::peak_buffer(c, cells, pb)) return -1;
  if (mfh_flat_peaks_async(ctx, d0, nd, t0, nt, binsize, threshold, ws, we, pb.idx, pb.val) < 0) return -1;
  const long total = mfh::read_peaks(c, pb, cells, nb, idx, val_c64, counts, rec);
  c->prof_collect();
  return total;
}

extern "C" long mfh_flat_peaks(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold,
                               size_t ws, size_t we, int64_t *idx, void *val_c64, int32_t *counts) {
  return flat_peaks_host(ctx, d0, nd, t0, nt, binsize, threshold, ws, we, idx, val_c64, nullptr, counts);
}

extern "C" long mfh_flat_peaks_rec(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t binsize, float threshold,
                                   size_t ws, size_t we, void *rec, int32_t *counts) {
  if (!rec) return fail("null output pointer");
  return flat_peaks_host(ctx, d0, nd, t0, nt, binsize, threshold, ws, we, nullptr, nullptr, rec, counts);
}

extern "C" int mfh_set_tuning(mfh_ctx *ctx, const char *key, long value) {
  if (!ctx || !key) return fail("null argument");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!strcmp(key, "dpw")) {
    if (value < 0) return fail("dpw must be >= 0");
    c->tune_dpw = (int)value;
    return 0;
  }
  if (!strcmp(key, "graphs")) {
    c->use_graphs = value != 0;
    return 0;
  }
  if (!strcmp(key, "big_chunk_mb")) {
    if (value < 1 || value > 65536) return fail("big_chunk_mb must be in [1, 65536]");
    c->big_chunk_bytes = (long long)value << 20;
    return 0;
  }
  if (!strcmp(key, "coarse_rpw")) {
    if (value != 1 && value != 2) return fail("coarse_rpw must be 1 or 2");
    c->coarse_rpw = (int)value;
    return 0;
  }
  if (!strcmp(key, "coarse_tpw")) {
    if (value < 1 || value > 64) return fail("coarse_tpw must be in [1, 64]");
    c->coarse_tpw = (int)value;
    return 0;
  }
  if (!strcmp(key, "coarse_lds")) {
    if (value != 0 && value != 1) return fail("coarse_lds must be 0 or 1");
    c->coarse_lds = (int)value;
    return 0;
  }
  if (!strcmp(key, "coarse_fuse")) {
    if (value != 0 && value != 1) return fail("coarse_fuse must be 0 or 1");
    c->coarse_fuse = (int)value;
    return 0;
  }
  if (!strcmp(key, "refine_grid")) {
    if (value < 1) return fail("refine_grid must be >= 1");
    c->refine_grid_cap = (int)value;
    return 0;
  }
  return fail("unknown tuning key '%s'", key);
}

extern "C" int mfh_prof_enable(mfh_ctx *ctx, int on) {
  if (!ctx) return fail("null context");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  c->prof_collect();
  c->prof_on = on != 0;
  c->prof_totals.clear();
  return 0;
}

extern "C" int mfh_prof_read(mfh_ctx *ctx, const char *name, double *total_ms, long *count) {
  if (!ctx || !name) return fail("null argument");
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  c->prof_collect();
  auto it = c->prof_totals.find(name);
  if (total_ms) *total_ms = it == c->prof_totals.end() ? 0.0 : it->second.first;
  if (count) *count = it == c->prof_totals.end() ? 0 : it->second.second;
  return 0;
}
