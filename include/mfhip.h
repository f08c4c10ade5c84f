/* mfhip -- MI355X (gfx950) batched matched filter, C ABI.
 *
 * The library behind apogee_amd's device="gpu" path.  Every entry point is
 * extern "C", takes plain pointers and sizes, returns an int status (< 0 is an
 * error whose text mfh_last_error() returns for the calling thread), and is
 * synchronous on return unless its name ends in _async, in which case the work
 * is queued on the context's HIP stream (mfh_stream) and mfh_sync() waits.
 *
 * Spectra are complex64 (interleaved float re, im), natural frequency order,
 * the unnormalised forward transform of each segment.  For a (data d,
 * template t) pair the filter evaluates
 *     z[k] = sum_f D_d[f] conj(T_t[f]) exp(+2 pi i f k / n)
 * and reports, per bin of the lag window, the lag of the largest |z| and z
 * there; a bin whose best |z|^2 does not exceed threshold^2 reports index -1
 * and value 0+0j.  Among equal |z|^2 the lowest lag is reported.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference checkout).  The reference's GPU contract is the Python
 * Context protocol of python/matchedfilter/_vkcompute.py; its CPU contract is
 * python/matchedfilter/matchedfilter.h.
 */
#ifndef MFHIP_H
#define MFHIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfh_ctx mfh_ctx;

/* Last error message of the calling thread ("" if none).
 * Replaces: VulkanError text (python/matchedfilter/_vkcompute.py:252-258). */
const char *mfh_last_error(void);

/* Number of HIP devices visible to the process (0 when none).
 * Replaces: GPU enumeration (python/matchedfilter/_vulkan.py:enumerate_devices,
 * python/matchedfilter/device.py:114-142). */
int mfh_device_count(void);

/* Name and gfx architecture of device `index` into caller buffers.
 * Replaces: Device.name / arch_keys (python/matchedfilter/device.py:20-56). */
int mfh_device_info(int index, char *name, size_t name_len, char *arch, size_t arch_len);

/* 1 if the transform length n is supported by this build, else 0.
 * Replaces: _GPU_SIZES (python/matchedfilter/__init__.py:58-59) and
 * ap_supported (python/matchedfilter/matchedfilter.h:204). */
int mfh_supported(size_t n);

/* Create a filter context on `device` for length n with room for nd data and
 * nt template spectra.  NULL on failure.
 * Replaces: Context(index) (python/matchedfilter/_vkcompute.py:962 class) and
 * ap_mf_create (python/matchedfilter/matchedfilter.h:53). */
mfh_ctx *mfh_create(int device, size_t n, int nd, int nt);

/* Release a context and its device memory.
 * Replaces: Context.destroy (python/matchedfilter/_vkcompute.py) and
 * ap_mf_destroy (matchedfilter.h:54). */
void mfh_destroy(mfh_ctx *ctx);

/* The context's HIP stream (hipStream_t), for event timing by the caller. */
void *mfh_stream(mfh_ctx *ctx);

/* Wait for all work queued on the context's stream. */
int mfh_sync(mfh_ctx *ctx);

/* Page-locked host memory the device writes without caching it (coherent),
 * for results: a records array (the *_rec entries' `rec`), or index and
 * value arrays laid out back to back (values right after the indices), that
 * lie inside such a buffer are written directly by the device, with no host
 * copy.  mfh_host_alloc returns NULL on failure (message set);
 * mfh_host_free returns 0 or < 0 (also for a pointer it did not hand out).
 * Nearest reference counterpart: the host-visible shared arrays of
 * empty_shared (python/matchedfilter/_shared.py:58-76). */
void *mfh_host_alloc(size_t bytes);
int mfh_host_free(void *p);

/* Order the context's stream after the work queued so far on
 * `producer_stream` (a hipStream_t of the context's device; null = the null
 * stream): an event recorded there that the context's stream waits on, with
 * no host wait, or nothing when that stream has no work left to run.  The
 * Python binding hands over PyTorch tensors this way,
 * which is what a DLPack __dlpack__(stream=...) exchange does, without the
 * capsule.  No counterpart in the reference, which refuses device arrays
 * (python/matchedfilter/__init__.py:103-131). */
int mfh_stream_wait(mfh_ctx *ctx, void *producer_stream);

/* Copy `rows` spectra (rows x n complex64, host memory) into data rows
 * [row0, row0 + rows).  The library copies; the caller's buffer is free on
 * return.
 * Replaces: ap_mf_set_data (matchedfilter.h:60) and the upload of
 * Context.peaks(upload_data=True) (python/matchedfilter/_vkcompute.py:1042-1047). */
int mfh_set_data(mfh_ctx *ctx, int row0, int rows, const void *host_c64);

/* Same for template rows.  No conjugation is stored: the kernels conjugate.
 * Replaces: ap_mf_set_template (matchedfilter.h:61). */
int mfh_set_templates(mfh_ctx *ctx, int row0, int rows, const void *host_c64);

/* Device-resident variants: `dev_c64` is a device pointer on the context's
 * device (e.g. a torch tensor); copied device to device, queued on the stream.
 * Used for single rows; a whole bank is bound instead (below).
 * Replaces: nothing in the reference (it refuses device arrays,
 * python/matchedfilter/__init__.py:103-137); the device-resident half of the
 * ingest row of the hot path. */
int mfh_set_data_device_async(mfh_ctx *ctx, int row0, int rows, const void *dev_c64);
int mfh_set_templates_device_async(mfh_ctx *ctx, int row0, int rows, const void *dev_c64);

/* Zero-copy ingest: every later run reads the whole data (or template) bank
 * in place from `dev_c64`, nd x n (nt x n) complex64, 8-byte aligned device
 * memory of the context's device (checked with hipPointerGetAttributes).
 * Nothing is copied and nothing is queued: the caller orders the context's
 * stream (mfh_stream) after its own writes to the buffer (a DLPack
 * hand-over does exactly that) and keeps the buffer alive and unchanged
 * until the bank is replaced or the context destroyed.  The library never
 * writes it: a later single-row set, or a series forward into the data rows,
 * first copies the bank into the context's own storage.  Replacing the
 * whole bank (mfh_set_* with row0 0 and every row, or another bind) ends the
 * binding.
 * Replaces: the reference's retained caller spectrum of the hierarchical CPU
 * plan (src/hmf.c:139-143, held by python/matchedfilter/__init__.py:338-353)
 * and, for device memory, its refusal of accelerator arrays
 * (python/matchedfilter/__init__.py:103-131); host-visible shared storage is
 * its GPU alternative (python/matchedfilter/_shared.py:58-76). */
int mfh_bind_data(mfh_ctx *ctx, const void *dev_c64);
int mfh_bind_templates(mfh_ctx *ctx, const void *dev_c64);

/* Copy resident rows back to host memory: data rows (templates = 0) or
 * template rows (1), [row0, row0 + rows), as the kernels read them (a bound
 * bank reads from the caller's buffer).  Synchronous.
 * Replaces: reading the host-visible spectra a forward pass filled
 * (Context.forward into empty_shared storage,
 * python/matchedfilter/_vkcompute.py:855-903). */
int mfh_get_rows(mfh_ctx *ctx, int templates, int row0, int rows, void *host_c64);

/* Number of bins of window [ws, we) at `binsize` (0 if empty).
 * Replaces: ap_mf_nbins (matchedfilter.h:56; src/matchfilt.c:151-156). */
size_t mfh_nbins(size_t n, size_t binsize, size_t ws, size_t we);

/* Flat filter over data rows [d0, d0+nd) x template rows [t0, t0+nt), lags
 * [ws, we) in bins of `binsize`.  Writes host arrays idx (int64) and val
 * (complex64) of shape [nd][nt][nbins]; if counts is not NULL it receives the
 * per-pair number of bins with index >= 0.  Returns the total number of such
 * bins, or < 0 on error.
 * Replaces: Context.peaks (python/matchedfilter/_vkcompute.py:962-1063) and
 * ap_mf_run (matchedfilter.h:66-68). */
long mfh_flat_peaks(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                    size_t binsize, float threshold, size_t ws, size_t we,
                    int64_t *idx, void *val_c64, int32_t *counts);

/* As mfh_flat_peaks, with the results as records: rec receives nd*nt*nbins
 * records {int64 index; float re, im} (16 bytes each, the PEAK_DTYPE layout
 * of the Python surface), written by the device in that layout: straight
 * into rec when it lies in an mfh_host_alloc buffer, else into page-locked
 * staging and then copied.
 * Replaces: the record assembly of _format_result
 * (python/matchedfilter/__init__.py:155-173) after Context.peaks. */
long mfh_flat_peaks_rec(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                        size_t binsize, float threshold, size_t ws, size_t we,
                        void *rec, int32_t *counts);

/* Same with device output pointers; queued on the context stream, no sync.
 * Returns 0 or < 0.  Replaces: the fusedTierB dispatch recorded by
 * Context._make_batch (python/matchedfilter/_vkcompute.py:923-960). */
int mfh_flat_peaks_async(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                         size_t binsize, float threshold, size_t ws, size_t we,
                         int64_t *dev_idx, void *dev_val_c64);

/* ---- hierarchical filter ------------------------------------------------ */

/* Select the coarse band (a power of two, 64 <= band < n) and how coarse
 * templates are normalised: ref_fraction in [0, 1] is the bank-wide in-band
 * fraction of a reference profile (sum_{k<band} ref / sum ref, FP64 on the
 * host); a negative value normalises each template by its own FP64 band
 * fraction, computed on the device.  Coarse templates T[:band]/sqrt(f) are
 * rebuilt on the device whenever templates or this setting change.
 * Replaces: ap_hmf_create_ex band / ap_hmf_set_reference
 * (python/matchedfilter/matchedfilter.h:129-154; src/hmf.c:118-156) and the
 * host-side coarse-template scaling of HierarchicalFilter._gpu_dispatch
 * (python/matchedfilter/__init__.py:1284-1303). */
int mfh_hier_set_band(mfh_ctx *ctx, size_t band, double ref_fraction);

/* Caller-scaled coarse templates: ct0 is nt x band complex64 (host memory,
 * or device memory of the context's device when on_device is 1), already
 * T[:band] / sqrt(f) by the caller's normalisation.  Sets the band and uses
 * ct0 as the coarse bank as given (no device scaling) until the band or the
 * templates change; a later template change rebuilds the coarse bank from
 * the spectra with per-template band fractions.  Synchronous.
 * Replaces: the ct0 argument of Context.hier_peaks
 * (python/matchedfilter/_vkcompute.py:534-535, uploaded at :608-613), which
 * HierarchicalFilter._gpu_dispatch scales on the host
 * (python/matchedfilter/__init__.py:1284-1303). */
int mfh_hier_set_coarse_templates(mfh_ctx *ctx, size_t band, const void *ct0_c64, int on_device);

/* Hierarchical run: coarse band screen (the band-point flat filter on
 * D[:band] and the coarse templates, one bin over the coarse window), gate
 * (fire iff coarse |z|^2 > coarse_thr^2 and |z| >= coarse_thr), survivor
 * compaction, and the exact flat filter on survivors; dismissed pairs report
 * -1 / 0 in every bin.  Reported peaks are those of mfh_flat_peaks.  Outputs
 * and counts as mfh_flat_peaks; *n_refined receives the number of pairs that
 * went to the full filter.  Returns total crossings, or < 0.
 * Replaces: Context.hier_peaks (python/matchedfilter/_vkcompute.py:534-623)
 * and ap_hmf_run (matchedfilter.h:159-161; src/hmf.c:222-328). */
long mfh_hier_peaks(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                    size_t binsize, float threshold, size_t ws, size_t we,
                    float coarse_thr, int64_t *idx, void *val_c64, int32_t *counts,
                    int64_t *n_refined);

/* As mfh_hier_peaks, with the results as 16-byte records (see
 * mfh_flat_peaks_rec).
 * Replaces: the record assembly of _format_result
 * (python/matchedfilter/__init__.py:155-173) after Context.hier_peaks. */
long mfh_hier_peaks_rec(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                        size_t binsize, float threshold, size_t ws, size_t we,
                        float coarse_thr, void *rec, int32_t *counts,
                        int64_t *n_refined);

/* mfh_hier_peaks with device outputs, queued on the stream; if host_refined
 * is not NULL it receives the survivor count once the stream reaches that
 * point.
 * Replaces: the one-command-buffer hierarchical recording of
 * Context._make_hier (python/matchedfilter/_vkcompute.py:650-850). */
int mfh_hier_peaks_async(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                         size_t binsize, float threshold, size_t ws, size_t we,
                         float coarse_thr, int64_t *dev_idx, void *dev_val_c64,
                         int32_t *host_refined);

/* mfh_hier_peaks_async that allocates nothing: no hipMalloc, hipFree or
 * allocating library call, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  The
 * per-call buffers (coarse outputs, survivor list and counter) live in one
 * fixed device array of the library, 6 MiB + 256 B per GPU, held until the
 * process ends; at most 262144 (1 << 18) pairs per call, more is rejected
 * with an error.  Every buffer of the context itself (coarse bank, twiddle
 * and matrix-core tables for this band and precision) must already exist:
 * one mfh_hier_peaks or mfh_hier_peaks_async call on the context builds
 * them; otherwise the call fails with an error and launches nothing.  All
 * calls on one GPU share the array: calls that overlap on one GPU (other
 * streams, contexts or threads) race, and must be ordered by the caller.
 * The counter is cleared by a kernel at the start of each call (no memset).
 * Replaces: nothing in the reference, which has no allocation-free entry. */
int mfh_hier_peaks_noalloc(mfh_ctx *ctx, int d0, int nd, int t0, int nt,
                           size_t binsize, float threshold, size_t ws, size_t we,
                           float coarse_thr, int64_t *dev_idx, void *dev_val_c64);

/* Coarse screen precision: 0 = FP32 (the band-point flat filter, exact
 * reference semantics), 1 = fp16 or 2 = bf16 matrix cores (band 512 only;
 * FP32 accumulation), whose statistic is widened by (1 + delta) at the gate
 * so the survivors are a superset of the FP32 screen's when its relative
 * error stays within delta.  Other bands always use FP32.
 * Replaces: the FP16 coarse pass of the reference's GPU path (fusedTierB with
 * COARSE16, src/gpu/tierb.slang:776-900) -- no exact counterpart. */
int mfh_hier_set_coarse_precision(mfh_ctx *ctx, int precision, float delta);

/* The coarse statistic max |z_c| per pair over the coarse window of [ws, we)
 * at the given precision, into host_max[nd*nt]; for calibrating delta and
 * for tests.  No refine runs.
 * Replaces: the coarse dump of MF_HMF_DUMP (src/hmf.c:67-68, :279-282). */
int mfh_hier_coarse_max(mfh_ctx *ctx, int d0, int nd, int t0, int nt, size_t ws, size_t we,
                        int precision, float *host_max);

/* Run one MFMA instruction on given matrices (row-major FP32 inputs converted
 * to the instruction's type): variant 0 = 16x16x32 f16, 1 = 16x16x32 bf16
 * (A 16x32, B 32x16), 2 = 16x16x16 f16, 3 = 16x16x16 bf16 (A, B 16x16);
 * C is 16x16.  Pins the fragment layouts the coarse screen relies on.
 * Replaces: nothing in the reference (a self-test of this build). */
int mfh_selftest_mfma(int device, int variant, const float *A, const float *B, float *C);

/* Measured VALU ceiling: `blocks` workgroups of 256 lanes each run `chains`
 * (4, 8, 16 or 32) independent chains of iters x 16 dependent scalar FP32
 * operations, op 0 = FMA (2 flop) or 1 = add (1 flop); no packed FP32.
 * Writes the second of two identical launches' duration (ms, HIP events) and
 * its flop count.  Synchronous, on a private stream.
 * Replaces: the FMA / add chain probe of tools/gpu_roofline.py:186-233 (the
 * add rate is "the one that actually bounds an FFT", :225-233). */
int mfh_probe_valu(int device, int op, int chains, int blocks, int iters, double *ms_out,
                   double *flop_out);

/* Measured copy bandwidth: one grid-stride float4 copy of `bytes` (a
 * multiple of 16), device to device; writes the second launch's duration.
 * Replaces: the COPY probe of tools/gpu_roofline.py:236-243. */
int mfh_probe_copy(int device, size_t bytes, double *ms_out);

/* ---- time series --------------------------------------------------------- */

/* Filter a time series over a caller-supplied block layout: block b covers
 * series[start[b] .. start[b]+n) (zero-padded past the end; spectrum
 * FFT(block)/n with the 1/n applied first) and reports lags
 * [win_start[b], win_end[b]) in bins of `binsize`.  Every window must give
 * the same bin count.  `series_c64` is nseries complex64 samples in host
 * memory, or a device pointer when series_on_device is non-zero.  Outputs are
 * [block][template][bin] in the caller's block order; counts (optional) per
 * (block, template).  Runs of consecutive equal windows share one launch.
 * Returns total crossings, or < 0.
 * Replaces: ap_mf_run_series (python/matchedfilter/matchedfilter.h:84-96;
 * src/matchfilt.c:437-490) and Context.forward + peaks of the reference's
 * GPU run_series (python/matchedfilter/_vkcompute.py:855-906,
 * python/matchedfilter/__init__.py:601-645). */
long mfh_run_series(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                    const size_t *start, const size_t *win_start, const size_t *win_end, int nblocks,
                    int t0, int nt, size_t binsize, float threshold,
                    int64_t *idx, void *val_c64, int32_t *counts);

/* As mfh_run_series, with the results as 16-byte records (see
 * mfh_flat_peaks_rec), [block][template][bin].
 * Replaces: the record assembly of _format_result
 * (python/matchedfilter/__init__.py:155-173) after the GPU run_series. */
long mfh_run_series_rec(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                        const size_t *start, const size_t *win_start, const size_t *win_end, int nblocks,
                        int t0, int nt, size_t binsize, float threshold,
                        void *rec, int32_t *counts);

/* The block spectra of run_series on their own: spectrum b = FFT of
 * series[start[b] : start[b] + n] (zero-padded past the end) with the 1/n
 * applied first, written into data rows [row0, row0 + nblocks) of the
 * context, where mfh_flat_peaks / mfh_hier_peaks then read them.
 * Replaces: Context.forward (python/matchedfilter/_vkcompute.py:855-903,
 * src/gpu/series_forward.slang:9-33), which fills the GPU data buffer. */
int mfh_series_forward(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                       const size_t *start, int nblocks, int row0);

/* The hierarchical twin: the block spectra go through the coarse screen and
 * refine of mfh_hier_peaks (band set by mfh_hier_set_band); *n_refined
 * receives the number of refined (block, template) pairs.
 * Replaces: ap_hmf_run_series (python/matchedfilter/matchedfilter.h:186-191;
 * src/hmf.c:158-220). */
long mfh_hier_run_series(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                         const size_t *start, const size_t *win_start, const size_t *win_end,
                         int nblocks, int t0, int nt, size_t binsize, float threshold,
                         float coarse_thr, int64_t *idx, void *val_c64, int32_t *counts,
                         int64_t *n_refined);

/* As mfh_hier_run_series, with the results as 16-byte records (see
 * mfh_flat_peaks_rec).
 * Replaces: the record assembly of _format_result
 * (python/matchedfilter/__init__.py:155-173) after ap_hmf_run_series. */
long mfh_hier_run_series_rec(mfh_ctx *ctx, const void *series_c64, size_t nseries, int series_on_device,
                             const size_t *start, const size_t *win_start, const size_t *win_end,
                             int nblocks, int t0, int nt, size_t binsize, float threshold,
                             float coarse_thr, void *rec, int32_t *counts,
                             int64_t *n_refined);

/* Tuning knobs: key "dpw" (data rows per workgroup; 0 = automatic),
 * "refine_grid" (workgroups of a list-driven refine), "coarse_tpw"
 * (templates per wave of the matrix-core coarse screen, 1..64), "coarse_rpw"
 * (its data rows per wave, 1 or 2), "coarse_lds" (0/1: its DFT constants
 * read from LDS), "coarse_fuse" (0/1, on by default: at n = 4096 with one
 * bin and the constants in LDS, the screen's workgroups refine their own
 * survivors in the same launch), "big_chunk_mb" (scratch
 * per four-step launch pair at n = 32768 / 65536), "graphs" (0/1: replay the
 * hierarchical sequence from HIP graphs; off by default, measured slower).
 * Replaces: MF_MFTILE / MF_PBMAX environment knobs (src/matchfilt.c:93-107). */
int mfh_set_tuning(mfh_ctx *ctx, const char *key, long value);

/* Per-kernel timing with HIP events on the context stream: enable (1) or
 * disable (0) recording, and read the accumulated milliseconds and launch
 * count of kernel `name` ("flat", ...).  Enabling clears the totals.
 * Replaces: the MF_HMF_PROF stage timers (src/hmf.c:65, :94-101). */
int mfh_prof_enable(mfh_ctx *ctx, int on);
int mfh_prof_read(mfh_ctx *ctx, const char *name, double *total_ms, long *count);

/* ---- Derivatives of peak values -----------------------------------------
 * For reported peaks idx [nd][nt][nb] (lags; -1 = nothing reported) of data
 * rows D and template rows T (device pointers, row strides in elements >= n,
 * the context's n), and the incoming gradient gval [nd][nt][nb] of a real
 * loss with respect to each value (dL/dRe z + i dL/dIm z), queue
 *     grad_data[d][f] = sum_{t,b} gval T_t[f] exp(-2 pi i f k / n)
 *     grad_tmpl[t][f] = sum_{d,b} conj(gval) D_d[f] exp(+2 pi i f k / n)
 * into dense [nd][n] / [nt][n] device arrays (either may be null).  Needs only
 * the stored lags: the correlation series is never rebuilt.
 * Replaces: nothing in the reference (the peak-only filter there has no
 * derivative); the forward it differentiates is mfh_flat_peaks. */
int mfh_peak_grad_async(mfh_ctx *ctx, const void *dev_data_c64, long long data_stride, int nd,
                        const void *dev_tmpl_c64, long long tmpl_stride, int nt, size_t nb,
                        const int64_t *dev_idx, const void *dev_gval_c64, void *dev_grad_data_c64,
                        void *dev_grad_tmpl_c64);

/* z at stored lags: out[d][t][b] = sum_f D_d[f] conj(T_t[f]) exp(+2 pi i f k / n)
 * (FP64 sums of FP32 products), 0 where idx is -1.  Recomputes a peak value
 * without the series.  Replaces: nothing in the reference. */
int mfh_peak_values_async(mfh_ctx *ctx, const void *dev_data_c64, long long data_stride, int nd,
                          const void *dev_tmpl_c64, long long tmpl_stride, int nt, size_t nb,
                          const int64_t *dev_idx, void *dev_out_c64);

/* ---- Uneven batched refine over a device work queue ---------------------
 * Requests of mixed transform lengths (64 .. 65536), each one data spectrum
 * and c consecutive template spectra in a shared pool of complex64, are
 * drained by one persistent launch.  Each (request, candidate) pair gives the
 * same index and value as mfh_flat_peaks on that pair.
 * Replaces: a loop of ap_mf_run / Context.peaks calls, one per request
 * (python/matchedfilter/matchedfilter.h:66; python/matchedfilter/_vkcompute.py:962),
 * and the single-length refine list of Context.hier_peaks
 * (python/matchedfilter/_vkcompute.py:534-850). */
typedef struct mfh_queue mfh_queue;
mfh_queue *mfh_queue_create(int device);
void mfh_queue_destroy(mfh_queue *queue);
/* The pool: `elems` complex64 values.  on_device = 0 copies host memory into
 * the queue; 1 uses the caller's device pointer, which must stay valid.
 * Setting a pool drops the plan (plan again before running). */
int mfh_queue_set_pool(mfh_queue *queue, const void *spectra_c64, size_t elems, int on_device);
/* Plan nreq requests.  Request r: length n[r]; data spectrum at pool element
 * data_off[r]; ncand[r] templates at tmpl_off[r] + j*n[r]; bins of binsize[r]
 * over the lag window [ws[r], we[r]); threshold[r] as in mfh_flat_peaks.
 * out_off[r] (may be null) receives the element offset of the request's
 * [ncand][nbins] outputs.  Returns the total output elements or -1. */
long long mfh_queue_plan(mfh_queue *queue, int nreq, const int64_t *n, const int64_t *data_off,
                         const int64_t *tmpl_off, const int32_t *ncand, const int64_t *binsize,
                         const float *threshold, const int64_t *ws, const int64_t *we, int64_t *out_off);
/* Run the plan into device outputs, queued on the queue's stream.  The
 * second identical call is captured into a HIP graph and replayed after. */
int mfh_queue_run_async(mfh_queue *queue, int64_t *dev_idx, void *dev_val_c64);
/* Run into host outputs; counts (may be null) receives the reported bins per
 * pair in plan order.  Returns the number of reported bins or -1. */
long long mfh_queue_run(mfh_queue *queue, int64_t *idx, void *val_c64, int32_t *counts);
/* Knobs: "grid_per_cu" (persistent workgroups per CU, default 1), "grid_cap"
 * (upper bound on persistent workgroups, 0 = none), "graphs". */
int mfh_queue_set_tuning(mfh_queue *queue, const char *key, long value);
int mfh_queue_sync(mfh_queue *queue);
void *mfh_queue_stream(mfh_queue *queue);
/* Event timing of the queue's launches (graph replay disabled meanwhile):
 * total milliseconds over all its kernels and the number of runs timed. */
int mfh_queue_prof_enable(mfh_queue *queue, int on);
int mfh_queue_prof_read(mfh_queue *queue, double *total_ms, long *count);

#ifdef __cplusplus
}
#endif
#endif
