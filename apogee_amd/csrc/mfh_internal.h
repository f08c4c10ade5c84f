// Host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "mfh_dft.h"

namespace mfh {

int fail(const char *fmt, ...);
bool hip_ok(hipError_t e, const char *what);
// Raise kernel `fn`'s dynamic LDS limit to at least `bytes` on the current
// device `dev`, once per (kernel, device) and size: the attribute belongs to
// a device, so a second GPU in the process gets its own call.
bool lds_limit(const void *fn, int dev, size_t bytes);

struct Ctx {
  int dev = -1;
  size_t n = 0;
  int nd = 0, nt = 0;
  hipStream_t stream = nullptr;
  hipEvent_t order_ev = nullptr;  // mfh_stream_wait's hand-over event (created on first use)
  // The banks every launch reads: [nd][n] and [nt][n].  Each points either at
  // the context's own allocation (data_own / tmpl_own) or, after
  // mfh_bind_data / mfh_bind_templates, at a caller's device buffer that is
  // read in place and never written.
  cf *data = nullptr;
  cf *tmpl = nullptr;
  cf *data_own = nullptr;
  cf *tmpl_own = nullptr;
  bool data_bound() const { return data != data_own; }
  bool tmpl_bound() const { return tmpl != tmpl_own; }
  // Before the library writes rows of a bound bank: copy the caller's bank
  // into the context's own allocation (queued on the stream) and read that
  // from then on, so the caller's buffer is never modified.
  int own_bank(bool templates);
  int tune_dpw = 0;          // 0: automatic
  // hierarchical state: coarse band, bank-wide reference fraction (< 0: per
  // template), scaled coarse templates [nt][band], rebuilt when dirty
  int coarse_band = 0;
  double coarse_ref_f = -1.0;
  bool coarse_dirty = true;
  cf *coarse_tmpl = nullptr;
  // coarse screen precision: 0 = FP32 flat screen, 1 = fp16 MFMA, 2 = bf16
  // MFMA; the MFMA statistic is widened by (1 + coarse_delta) at the gate
  int coarse_prec = 0;
  float coarse_delta = 0.f;
  int coarse_tpw = 4;  // templates per wave of the MFMA coarse screen
  int coarse_rpw = 2;  // data rows per wave of the MFMA coarse screen (1 or 2)
  int coarse_lds = 0;  // 1: the screen reads its DFT constants from LDS, not registers
  int coarse_fuse = 1;  // 1: where it applies, the matrix-core screen refines its survivors in the same launch
  // the coarse templates in MFMA fragment order, power-of-two scaled, and
  // their scale exponents (built with coarse_tmpl at band 512)
  cf *coarse_ctm = nullptr;
  int *coarse_sexp = nullptr;
  bool coarse_roots_ready = false;
  // four-step lengths (32768, 65536): roots built, scratch budget per chunk
  bool big_roots_ready[16] = {};  // by log2 of the column length
  long long big_chunk_bytes = 128ll << 20;  // measured best on gfx950 (256: +7%, 64: +16%)
  int refine_grid_cap = 2048;  // workgroups of a list-driven (refine) launch, at most
  // pinned host word holding the last hierarchical refine's survivor count,
  // written by that launch (a sizing hint only: results never depend on it)
  int *refine_hint = nullptr;
  // Per-pass twiddle tables of the flat kernel for length n (FP64-generated
  // on first use, kept for the context's lifetime).
  std::map<size_t, cf *> tw_by_n;
  const cf *twiddles(size_t n);

  // Pinned host staging for results read back into caller host arrays,
  // grown on demand (hipHostMalloc; contents not preserved on growth).
  void *stage = nullptr;
  size_t stage_bytes = 0;
  void *host_stage(size_t bytes);

  struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Buf> bufs;
  // Device buffer `name`, grown on demand (contents not preserved on growth).
  void *scratch(const char *name, size_t bytes);

  // per-kernel event timing
  struct Ev {
    hipEvent_t a = nullptr, b = nullptr;
    const char *name = "";
  };
  bool prof_on = false;
  bool capturing = false;  // a stream capture is recording launches
  // set for the duration of an allocation-free call: scratch() then fails
  // instead of allocating, and no timing events are taken
  bool no_alloc = false;
  std::vector<Ev> prof_events;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  // HIP graph of the hierarchical stage sequence, keyed by its launch
  // parameters: replayed when a call repeats them (captured on the second
  // identical call, so every buffer and kernel attribute exists by then).
  // hierarchical calls: graph replay measured slower than direct launches
  // on gfx950 (18.7 against 13.2 us per call at 16 x 1024, n = 4096), so it
  // is off unless the tuning key "graphs" asks for it; the refine queue
  // turns it on for its own context
  bool use_graphs = false;
  hipGraphExec_t graph_exec = nullptr;
  // the hierarchical sequence's graphs, one per refine-counter parity, both
  // under graph_key; hier_parity is the counter the next call uses
  hipGraphExec_t hier_graph[2] = {nullptr, nullptr};
  int hier_parity = 0;
  void drop_hier_graphs() {
    for (auto &g : hier_graph)
      if (g) {
        hipGraphExecDestroy(g);
        g = nullptr;
      }
  }
  std::vector<long long> graph_key, last_key;
  std::map<std::string, std::pair<double, long>> prof_totals;
  void prof_collect();
  struct ProfScope {
    ProfScope(Ctx *c, const char *name);
    ~ProfScope();
    Ctx *c_;
    const char *name_;
    Ev ev_;
  };

  ~Ctx();
};

// A call's device results, laid out [cells indices | cells values] in the
// context's "out_peaks" buffer so that one copy brings both back.
struct PeakBuf {
  int64_t *idx = nullptr;
  cf *val = nullptr;
};
bool peak_buffer(Ctx *c, size_t cells, PeakBuf &b);
// Bring `cells` results to the caller's host arrays: a kernel queued on the
// context's stream writes them in the caller's layout -- 16-byte {int64
// index; complex64 value} records into `rec` (idx and val unused), else
// [indices | values] -- straight into the caller's memory when it lies in an
// mfh_host_alloc buffer (the values right after the indices, for idx/val),
// else into the page-locked staging buffer followed by host copies (or
// copy-engine transfers when no staging can be had); then waits for the
// stream.  counts[p] (if set) = bins set of pair p, nb bins per pair.
// Returns the total of set bins, or -1 (message set).
long read_peaks(Ctx *c, const PeakBuf &b, size_t cells, size_t nb, int64_t *idx, void *val_c64, int32_t *counts,
                void *rec = nullptr);

// One flat launch: pairs (data rows x template rows) from the given bases.
struct FlatRun {
  const cf *data = nullptr;
  const cf *tmpl = nullptr;
  long long data_stride = 0, tmpl_stride = 0;
  int nd = 0, nt = 0;
  size_t binsize = 0;
  float threshold = 0.f;
  size_t ws = 0, we = 0;
  int64_t *out_idx = nullptr;
  cf *out_val = nullptr;
  // list mode: pairs (data row, template row) read from the device
  const int2 *list = nullptr;
  const int *list_count = nullptr;
  const long long *list_ooff = nullptr;  // output element offset per pair (else d*nt + t rows)
  int list_cap = 0;
  const char *prof_name = "flat";        // event-timing name of a tiled launch
  const char *list_prof_name = "refine";  // and of a list-mode launch
  int *zero = nullptr;                    // if set, the launch stores 0 there
  // list mode: a host-mapped word the launch stores the list length in, and
  // from which the launch's grid is sized (the previous call's length)
  int *hint = nullptr;
};

bool flat_supported(size_t n);
// Per-pass twiddle table size (complex elements) and its FP64 generation.
size_t flat_twiddle_count(size_t n);
int flat_fill_twiddles(Ctx *c, size_t n, cf *tw);
// Validates, splits windows whose bin table would not fit LDS, and launches
// the length-n flat kernel on c->stream.  Returns 0 or -1 (message set).
int launch_flat(Ctx *c, size_t n, const FlatRun &r);
// The flat kernel's arguments for r in one launch (n = 4096, all bins of the
// window in one bin table), for kernels that run its per-pair path.
struct FlatArgs;
int flat_fill_args(Ctx *c, size_t n, const FlatRun &r, FlatArgs &a);
// The two-kernel four-step flat filter for n = 32768 and 65536.
int launch_big(Ctx *c, size_t n, const FlatRun &r);
// Four-step root tables for n = n1 * 256: w_n1^h (h < n1) then w_n^l (l < 256),
// FP64-generated once per context.
const cf *big_roots(Ctx *c, int n1);
// Start offsets of run_series blocks as the forward kernels read them: up to
// kInline in the kernel arguments themselves (no upload), else a device table.
struct BlockStarts {
  static constexpr int kInline = 64;
  const long long *dev = nullptr;  // device table of every block, or null: `inl` holds them
  long long inl[kInline];
  __device__ long long operator[](int b) const { return dev ? dev[b] : inl[b]; }
};
// Forward spectra of run_series blocks into out [nblocks][n] (device).
int launch_series_forward(Ctx *c, const cf *series, size_t nseries, const BlockStarts &starts, int nblocks,
                          cf *out);
// MFMA coarse statistic (max |z_c|^2 over [cs, ce)) per pair, band 512;
// `zero` (may be null) is set to 0 by the launch.
int coarse_mfma_prepare(Ctx *c);
// The gate applied inside the coarse screen when `list` is set: a pair
// fires iff thr2 <= 0 or |z_c|^2 >= thr2, with thr2 = (thr / (1 + delta))^2
// rounded down (a superset of the FP32 screen's decisions whenever the
// statistic's relative error is within delta); fired pairs are appended to
// `list` (count must be zero at launch), the others get -1 / 0 in their nb
// output bins.
struct CoarseGate {
  float thr2 = 0.f;
  int nb = 0;
  int64_t *idx = nullptr;
  cf *val = nullptr;
  int2 *list = nullptr;
  int *count = nullptr;
  int *zero = nullptr;  // set to 0 by the launch (the next call's counter)
};
int launch_coarse_mfma(Ctx *c, const cf *data, long long data_stride, int t0, int nd, int nt, int cs, int ce,
                       bool bf16, float *cmax2, const CoarseGate *gate = nullptr);
// Whether a hierarchical run of length n with nb bins takes the one-launch
// screen-and-refine (matrix-core screen with its constants in LDS, n = 4096,
// one bin, tuning key coarse_fuse), and that launch: the screen of
// launch_coarse_mfma with gate g, whose survivors are refined per rr in the
// same workgroups (g.list is not used).
bool coarse_refine_fits(const Ctx *c, size_t n, size_t nb);
int launch_coarse_refine(Ctx *c, const cf *data, long long data_stride, int t0, int nd, int nt, int cs, int ce,
                         bool bf16, const CoarseGate &g, const FlatRun &rr);
// Hierarchical run over data rows [d0, d0+nd) of the context, or over
// `data_rows` (stride n) when not null.
int *hier_counter(Ctx *c, int parity);
// Per-call buffers of a hierarchical run supplied by the caller instead of
// the context: coarse outputs, the survivor list and its counter (cleared by
// the run), and a spare counter word the screen may clear.
struct HierScratch {
  int64_t *cidx = nullptr;
  cf *cval = nullptr;
  int2 *list = nullptr;
  int *count = nullptr;
  int *spare = nullptr;
};
int hier_run(Ctx *c, int d0, int nd, int t0, int nt, size_t binsize, float threshold, size_t ws, size_t we,
             float coarse_thr, int64_t *idx, cf *val, int *count_out, const cf *data_rows,
             const HierScratch *hs = nullptr);

}  // namespace mfh
