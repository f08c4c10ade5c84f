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

struct Ctx {
  int dev = -1;
  size_t n = 0;
  int nd = 0, nt = 0;
  hipStream_t stream = nullptr;
  cf *data = nullptr;  // [nd][n]
  cf *tmpl = nullptr;  // [nt][n]
  cf *tw = nullptr;    // per-pass twiddle tables of the flat kernel
  int tune_dpw = 0;    // 0: automatic

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
  std::vector<Ev> prof_events;
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
  long long out_row_stride = 0;  // unused (dense [nd][nt][nb])
};

bool flat_supported(size_t n);
// Per-pass twiddle table size (complex elements) and its FP64 generation.
size_t flat_twiddle_count(size_t n);
int flat_fill_twiddles(Ctx *c);
// Validates, splits windows whose bin table would not fit LDS, and launches
// on c->stream.  Returns 0 or -1 (message set).
int launch_flat(Ctx *c, const FlatRun &r);

}  // namespace mfh
