// Single-thread FP32 CPU restatement of the batched matched filter.
//
// TEST INFRASTRUCTURE ONLY: this is the CPU checker and the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it; the shipped HIP path never does.
//
// It restates the reference's CPU engine semantics (paths relative to the
// reference checkout):
//   * ingest conjugates templates once (src/matchfilt.c:195-215);
//   * per pair: product D*conj(T), unnormalised inverse FFT, |z|^2 scan
//     (src/balanced-inl.h:1102-1113, :876-969);
//   * selection key m2 = fma(re, re, im*im) in FP32, seed = thr>0 ? thr*thr : 0,
//     strict '>' against the seed and the running best (src/balanced-inl.h:886,
//     :944, :956), so an all-zero bin reports -1 / 0;
//   * pairs are walked in 8x8 (data, template) tiles (src/matchfilt.c:375-407).
// The FFT is this file's own: a radix-4 Stockham autosort on split re/im
// arrays with FP64-generated twiddles (plus one radix-2 pass for odd log2 n).
// Ties resolve to the lowest lag because the scan is ascending in k.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Plan {
  size_t n = 0;
  int log2n = 0;
  // one twiddle table per pass: tw[p][(r-1)*Ns + k] = exp(+2 pi i r k / (4 Ns))
  std::vector<std::vector<float>> twr, twi;
  std::vector<int> radix, ns;
  std::vector<float> ar, ai, br, bi;   // ping-pong work buffers
};

bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

Plan *make_plan(size_t n) {
  if (!is_pow2(n) || n < 2) return nullptr;
  Plan *p = new Plan;
  p->n = n;
  while ((size_t(1) << p->log2n) < n) p->log2n++;
  size_t Ns = 1;
  int rem = p->log2n;
  if (rem & 1) {  // one radix-2 pass first: no twiddles at Ns = 1
    p->radix.push_back(2);
    p->ns.push_back(int(Ns));
    p->twr.emplace_back();
    p->twi.emplace_back();
    Ns *= 2;
    rem -= 1;
  }
  for (; rem > 0; rem -= 2) {
    p->radix.push_back(4);
    p->ns.push_back(int(Ns));
    std::vector<float> tr(3 * Ns), ti(3 * Ns);
    for (int r = 1; r < 4; r++)
      for (size_t k = 0; k < Ns; k++) {
        double a = 2.0 * M_PI * double(r) * double(k) / double(4 * Ns);
        tr[(r - 1) * Ns + k] = float(std::cos(a));
        ti[(r - 1) * Ns + k] = float(std::sin(a));
      }
    p->twr.push_back(std::move(tr));
    p->twi.push_back(std::move(ti));
    Ns *= 4;
  }
  p->ar.resize(n); p->ai.resize(n); p->br.resize(n); p->bi.resize(n);
  return p;
}

// One Stockham pass: butterflies j in [0, n/R), k = j mod Ns; reads x[j + r n/R],
// twiddles by exp(+2 pi i r k/(R Ns)), R-point inverse DFT, writes
// y[(j/Ns) Ns R + k + r Ns].  The k loop is innermost and contiguous.
void pass4(size_t n, size_t Ns, const float *twr, const float *twi,
           const float *xr, const float *xi, float *yr, float *yi) {
  const size_t q = n / 4;
  for (size_t g = 0; g < q; g += Ns) {
    const size_t ob = g * 4;
    for (size_t k = 0; k < Ns; k++) {
      const size_t j = g + k;
      float a0r = xr[j], a0i = xi[j];
      float a1r = xr[j + q], a1i = xi[j + q];
      float a2r = xr[j + 2 * q], a2i = xi[j + 2 * q];
      float a3r = xr[j + 3 * q], a3i = xi[j + 3 * q];
      if (Ns > 1) {
        float c, s, t;
        c = twr[k]; s = twi[k];
        t = a1r * c - a1i * s; a1i = a1r * s + a1i * c; a1r = t;
        c = twr[Ns + k]; s = twi[Ns + k];
        t = a2r * c - a2i * s; a2i = a2r * s + a2i * c; a2r = t;
        c = twr[2 * Ns + k]; s = twi[2 * Ns + k];
        t = a3r * c - a3i * s; a3i = a3r * s + a3i * c; a3r = t;
      }
      // inverse radix-4: w4 = +i
      const float s0r = a0r + a2r, s0i = a0i + a2i;
      const float d0r = a0r - a2r, d0i = a0i - a2i;
      const float s1r = a1r + a3r, s1i = a1i + a3i;
      const float d1r = a1r - a3r, d1i = a1i - a3i;
      yr[ob + k] = s0r + s1r;          yi[ob + k] = s0i + s1i;
      yr[ob + k + 2 * Ns] = s0r - s1r; yi[ob + k + 2 * Ns] = s0i - s1i;
      // +i * d1 = (-d1i, d1r)
      yr[ob + k + Ns] = d0r - d1i;     yi[ob + k + Ns] = d0i + d1r;
      yr[ob + k + 3 * Ns] = d0r + d1i; yi[ob + k + 3 * Ns] = d0i - d1r;
    }
  }
}

void pass2(size_t n, const float *xr, const float *xi, float *yr, float *yi) {
  const size_t q = n / 2;  // Ns = 1: no twiddle
  for (size_t j = 0; j < q; j++) {
    const float a0r = xr[j], a0i = xi[j], a1r = xr[j + q], a1i = xi[j + q];
    yr[2 * j] = a0r + a1r; yi[2 * j] = a0i + a1i;
    yr[2 * j + 1] = a0r - a1r; yi[2 * j + 1] = a0i - a1i;
  }
}

// In: product in (ar, ai).  Out: pointer pair holding z in natural order.
void inverse_fft(Plan *p, float **outr, float **outi) {
  float *xr = p->ar.data(), *xi = p->ai.data();
  float *yr = p->br.data(), *yi = p->bi.data();
  for (size_t s = 0; s < p->radix.size(); s++) {
    if (p->radix[s] == 2)
      pass2(p->n, xr, xi, yr, yi);
    else
      pass4(p->n, size_t(p->ns[s]), p->twr[s].data(), p->twi[s].data(), xr, xi,
            yr, yi);
    std::swap(xr, yr);
    std::swap(xi, yi);
  }
  *outr = xr;
  *outi = xi;
}

}  // namespace

extern "C" {

// Flat filter over every pair of D (nd x n) and T (nt x n), interleaved
// complex64.  Output idx[(d*nt + t)*nb + b] (int64), val interleaved complex64,
// counts[d*nt + t] (may be NULL).  Returns total crossings, or -1 on bad input.
long mfcpu_flat(size_t n, int nd, int nt, const float *D, const float *T,
                size_t binsize, float threshold, size_t ws, size_t we,
                int64_t *idx, float *val, int *counts) {
  if (!D || !T || !idx || !val || nd < 1 || nt < 1 || binsize < 1) return -1;
  if (we > n) we = n;
  if (ws >= we) return -1;
  Plan *p = make_plan(n);
  if (!p) return -1;
  const size_t nb = (we - ws + binsize - 1) / binsize;
  // ingest: split and conjugate the templates once (src/matchfilt.c:195-215)
  std::vector<float> dr(size_t(nd) * n), di(size_t(nd) * n);
  std::vector<float> tr(size_t(nt) * n), ti(size_t(nt) * n);
  for (size_t i = 0; i < size_t(nd) * n; i++) { dr[i] = D[2 * i]; di[i] = D[2 * i + 1]; }
  for (size_t i = 0; i < size_t(nt) * n; i++) { tr[i] = T[2 * i]; ti[i] = -T[2 * i + 1]; }
  const float t2 = threshold > 0.f ? threshold * threshold : 0.f;
  std::vector<float> best(nb);
  long total = 0;
  const int tile = 8;
  for (int d0 = 0; d0 < nd; d0 += tile)
    for (int t0 = 0; t0 < nt; t0 += tile)
      for (int d = d0; d < nd && d < d0 + tile; d++)
        for (int t = t0; t < nt && t < t0 + tile; t++) {
          const float *xr = &dr[size_t(d) * n], *xi = &di[size_t(d) * n];
          const float *hr = &tr[size_t(t) * n], *hi = &ti[size_t(t) * n];
          float *pr = p->ar.data(), *pi = p->ai.data();
          for (size_t k = 0; k < n; k++) {  // D * conj(T), conj already applied
            pr[k] = xr[k] * hr[k] - xi[k] * hi[k];
            pi[k] = xr[k] * hi[k] + xi[k] * hr[k];
          }
          float *zr, *zi;
          inverse_fft(p, &zr, &zi);
          const size_t row = size_t(d) * nt + t;
          int64_t *oi = idx + row * nb;
          float *ov = val + 2 * row * nb;
          for (size_t b = 0; b < nb; b++) { best[b] = t2; oi[b] = -1; ov[2 * b] = 0.f; ov[2 * b + 1] = 0.f; }
          for (size_t k = ws; k < we; k++) {
            const float m2 = std::fma(zr[k], zr[k], zi[k] * zi[k]);
            const size_t b = (k - ws) / binsize;
            if (m2 > best[b]) {
              best[b] = m2; oi[b] = int64_t(k); ov[2 * b] = zr[k]; ov[2 * b + 1] = zi[k];
            }
          }
          int c = 0;
          for (size_t b = 0; b < nb; b++) c += oi[b] >= 0;
          if (counts) counts[row] = c;
          total += c;
        }
  delete p;
  return total;
}

// Forward spectrum of one run_series block: FFT(block / n), zero-padded
// (src/matchfilt.c:458-470), via conj(IFFT(conj(x))).  out is interleaved.
int mfcpu_block_spectrum(size_t n, const float *series, size_t nseries,
                         size_t start, float *out) {
  Plan *p = make_plan(n);
  if (!p) return -1;
  size_t have = start < nseries ? nseries - start : 0;
  if (have > n) have = n;
  const float inv = 1.0f / float(n);
  for (size_t k = 0; k < n; k++) {
    float r = 0.f, i = 0.f;
    if (k < have) { r = series[2 * (start + k)] * inv; i = series[2 * (start + k) + 1] * inv; }
    p->ar[k] = r; p->ai[k] = -i;
  }
  float *zr, *zi;
  inverse_fft(p, &zr, &zi);
  for (size_t k = 0; k < n; k++) { out[2 * k] = zr[k]; out[2 * k + 1] = -zi[k]; }
  delete p;
  return 0;
}

}  // extern "C"
