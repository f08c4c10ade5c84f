"""FP64 restatement of the batched matched filter -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The shipped path (``apogee_amd``) never calls into it and fails
loudly when its HIP library is missing.

What it restates, function by function, with the reference lines it follows
(paths relative to the reference checkout):

* ``correlate``      z = n * ifft(D * conj(T)), unnormalised inverse
                     (python/matchedfilter/matchedfilter.h:35-51,
                     tests/test_api.py:68-81).
* ``direct_correlate`` the same from the O(n^2) definition, no FFT
                     (tests/test_independent.py:24-39).
* ``nbins``          ceil((end-start)/binsize), clamped window
                     (src/matchfilt.c:151-156, python/matchedfilter/__init__.py:425-433).
* ``select``         per-bin argmax with the selection rule of
                     src/balanced-inl.h:876-969: strict ``>`` against the
                     seed, seed = threshold^2 (0 when threshold <= 0), a bin
                     whose best is exactly 0 reports -1 / 0+0j.  Ties go to the
                     lowest lag (the reference leaves the tie unspecified,
                     tests/test_independent.py:281-285; lowest-lag is this
                     build's documented choice).
* ``flat_peaks``     the whole flat filter (src/matchfilt.c:349-420).
* ``coarse_templates`` / ``band_fraction`` / ``coarse_window`` /
  ``hier_peaks``     the explicit-band hierarchical filter
                     (src/hmf.c:118-156, :222-328).
* ``series_spectra`` / ``series_peaks``  run_series block spectra
                     FFT(series[s:s+n] zero-padded) / n and grouped filtering
                     (src/matchfilt.c:437-490, python/matchedfilter/_series.py).
* ``peak_value_grad`` closed-form gradient of a peak value with respect to
                     the template spectrum at a fixed lag (no reference
                     counterpart; pinned by finite differences in the tests).

Parity of every function above is pinned in ``tests/test_oracle.py`` against
the O(n^2) definition, closed forms, brute force and the golden fixtures under
``tests/golden/``.  ``gate_for``-style Monte-Carlo calibration is not restated
here.
"""
import numpy as np

__all__ = [
    "correlate", "direct_correlate", "nbins", "window", "select", "flat_peaks",
    "band_fraction", "coarse_templates", "coarse_window", "hier_peaks",
    "series_spectra", "series_peaks", "peak_value_grad", "flops_per_pair",
]


def flops_per_pair(n):
    """Algorithmic flop model per (data, template) pair: 6n + 5n*log2(n) + 3n.

    The reference's own roofline model (tools/gpu_roofline.py:249-260):
    complex product (6n), radix-2-equivalent inverse FFT (5 n log2 n) and the
    |z|^2 scan (3n).
    """
    n = int(n)
    return 6 * n + 5 * n * int(np.log2(n)) + 3 * n


def correlate(D, T):
    """z[..., k] = sum_f D[f] conj(T[f]) exp(+2 pi i f k / n), in complex128.

    Broadcasts over leading axes.  Follows matchedfilter.h:35-51: the inverse
    transform is unnormalised, so a perfect match returns n * energy.
    """
    D = np.asarray(D, dtype=np.complex128)
    T = np.asarray(T, dtype=np.complex128)
    n = D.shape[-1]
    return np.fft.ifft(D * np.conj(T), axis=-1) * n


def direct_correlate(D, T):
    """The correlation from its definition as a direct O(n^2) sum.

    Shares no code with an FFT (tests/test_independent.py:24-39), so agreement
    between this and ``correlate`` pins the sign and normalisation convention.
    """
    D = np.asarray(D, dtype=np.complex128)
    T = np.asarray(T, dtype=np.complex128)
    n = D.shape[-1]
    k = np.arange(n)
    w = np.exp(2j * np.pi * np.outer(k, k) / n)
    return (D * np.conj(T)) @ w


def window(n, win):
    """Clamp ``(start, end)`` to [0, n]; an empty window raises ValueError.

    python/matchedfilter/__init__.py:425-433.
    """
    if win is None:
        return 0, int(n)
    start = max(0, min(int(win[0]), n))
    end = max(0, min(int(win[1]), n))
    if start >= end:
        raise ValueError("empty window (%d, %d)" % (start, end))
    return start, end


def nbins(n, binsize, start=0, end=None):
    """Number of bins over the clamped window (src/matchfilt.c:151-156)."""
    end = n if end is None else min(int(end), n)
    binsize = int(binsize)
    if binsize < 1:
        raise ValueError("binsize must be >= 1")
    if start >= end:
        return 0
    return (end - start + binsize - 1) // binsize


def select(z, binsize, threshold, start, end, mag2=None):
    """Per-bin peak selection over lags [start, end) of ``z`` (last axis).

    Rule (src/balanced-inl.h:878-965): a sample wins only if its squared
    magnitude is strictly greater than both the seed and the bin's running
    best; the seed is threshold^2 for threshold > 0, else 0, so a bin whose
    maximum is exactly zero reports nothing.  Dismissed bins carry index -1
    and value 0+0j (src/balanced-inl.h:912-913).  Ties resolve to the lowest
    lag.  ``mag2`` overrides the squared magnitude used for selection (the
    FP32 key of the device path can be passed here to test the rule itself).

    Returns ``(index int64, value complex128, best |z|^2 float64)`` of shape
    ``z.shape[:-1] + (nbins,)``.
    """
    z = np.asarray(z)
    n = z.shape[-1]
    nb = nbins(n, binsize, start, end)
    lead = z.shape[:-1]
    if mag2 is None:
        mag2 = z.real.astype(np.float64) ** 2 + z.imag.astype(np.float64) ** 2
    thr = float(threshold)
    seed = thr * thr if thr > 0 else 0.0
    seg_m = mag2[..., start:end]
    seg_z = z[..., start:end]
    width = end - start
    pad = nb * binsize - width
    if pad:
        seg_m = np.concatenate(
            [seg_m, np.full(lead + (pad,), -1.0, dtype=seg_m.dtype)], axis=-1)
        seg_z = np.concatenate(
            [seg_z, np.zeros(lead + (pad,), dtype=seg_z.dtype)], axis=-1)
    m = seg_m.reshape(lead + (nb, binsize))
    zz = seg_z.reshape(lead + (nb, binsize))
    # NaN never wins a strict comparison: treat it as below everything.
    m_sel = np.where(np.isnan(m), -np.inf, m)
    arg = np.argmax(m_sel, axis=-1)              # first maximum = lowest lag
    best = np.take_along_axis(m_sel, arg[..., None], axis=-1)[..., 0]
    val = np.take_along_axis(zz, arg[..., None], axis=-1)[..., 0]
    idx = start + np.arange(nb) * binsize + arg
    keep = best > seed
    idx = np.where(keep, idx, -1).astype(np.int64)
    val = np.where(keep, val, 0).astype(np.complex128)
    best = np.where(keep, best, 0.0)
    return idx, val, best


def flat_peaks(D, T, binsize=None, threshold=0.0, win=None):
    """The flat filter over every (data, template) pair.

    Shapes: D (nd, n), T (nt, n) -> index (nd, nt, nb), value (nd, nt, nb),
    best |z|^2 (nd, nt, nb).  src/matchfilt.c:349-420 semantics in FP64.
    """
    D = np.atleast_2d(np.asarray(D))
    T = np.atleast_2d(np.asarray(T))
    n = D.shape[-1]
    binsize = n if binsize is None else int(binsize)
    start, end = window(n, win)
    z = correlate(D[:, None, :], T[None, :, :])
    return select(z, binsize, threshold, start, end)


def band_fraction(T, band, reference=None):
    """In-band power fraction f, in FP64.

    With a reference profile: f = sum_{k<band} ref / sum ref, one value for the
    whole bank (src/hmf.c:125-137; python/matchedfilter/__init__.py:1276-1278).
    Without: per template, sum_{k<band} |T|^2 / sum |T|^2 (0 for an all-zero
    template), kept in FP64 as the reference's GPU dispatch does
    (python/matchedfilter/__init__.py:1290-1300); the reference's CPU engine
    rounds this per-template fraction to float (src/hmf.c:147-152), a
    difference of one FP32 rounding in the coarse scale.
    """
    if reference is not None:
        ref = np.asarray(reference, dtype=np.float64)
        return float(ref[:band].sum() / ref.sum())
    T = np.atleast_2d(np.asarray(T, dtype=np.complex64))
    p = np.abs(T.astype(np.complex128)) ** 2
    tot = p.sum(axis=1)
    low = p[:, :band].sum(axis=1)
    return np.where(tot > 0, low / np.where(tot > 0, tot, 1.0), 0.0)


def coarse_templates(T, band, reference=None):
    """Coarse templates T[:band] / sqrt(f), scaled in FP64 and stored FP32.

    src/hmf.c:118-124 (``refresh_template``): scale = f > 0 ? 1/sqrt(f) : 0,
    applied to the FP32 spectrum in double, then rounded to FP32.  With a
    reference the CPU rounds f itself to FP32 first (src/hmf.c:133).
    """
    T = np.atleast_2d(np.asarray(T, dtype=np.complex64))
    f = band_fraction(T, band, reference)
    if reference is not None:
        f = np.full(T.shape[0], float(np.float32(f)))
    else:
        f = np.asarray(f, dtype=np.float64)
    scale = np.where(f > 0, 1.0 / np.sqrt(np.where(f > 0, f, 1.0)), 0.0)
    ct = T[:, :band].astype(np.complex128) * scale[:, None]
    return ct.astype(np.complex64)


def coarse_window(n, band, start, end):
    """Coarse lag window for a fine window (src/hmf.c:233-239).

    R = n / band; [max(0, floor(start/R) - 1), min(band, ceil(end/R))).
    """
    R = n // band
    cs = start // R
    ce = min(band, (end + R - 1) // R)
    if cs > 0:
        cs -= 1
    return cs, ce


def hier_peaks(D, T, band, coarse_threshold, binsize=None, threshold=0.0,
               win=None, reference=None, coarse_max=None):
    """Explicit-band hierarchical filter: coarse gate, then the flat filter.

    src/hmf.c:222-328: the coarse pass is the band-point flat filter on
    D[:band] and the coarse templates, one bin over the coarse window; a pair
    fires iff its coarse |z|^2 beats threshold^2 and |z| >= threshold; fired
    pairs report exactly the flat peaks, the rest report -1 / 0 in every bin.
    ``coarse_max`` (nd, nt) overrides the coarse maximum |z| (to replay a
    device's coarse screen).  Returns (index, value, best, fired).
    """
    D = np.atleast_2d(np.asarray(D, dtype=np.complex64))
    T = np.atleast_2d(np.asarray(T, dtype=np.complex64))
    n = D.shape[-1]
    start, end = window(n, win)
    binsize = n if binsize is None else int(binsize)
    if coarse_max is None:
        ct = coarse_templates(T, band, reference)
        cs, ce = coarse_window(n, band, start, end)
        zc = correlate(D[:, None, :band], ct[None, :, :])
        m2 = np.abs(zc[..., cs:ce]) ** 2
        coarse_max = np.sqrt(m2.max(axis=-1))
    thr = float(coarse_threshold)
    fired = (coarse_max ** 2 > thr * thr) & (coarse_max >= thr) if thr > 0 \
        else coarse_max > 0
    idx, val, best = flat_peaks(D, T, binsize, threshold, (start, end))
    idx = np.where(fired[..., None], idx, -1)
    val = np.where(fired[..., None], val, 0)
    best = np.where(fired[..., None], best, 0.0)
    return idx, val, best, fired


def series_spectra(series, starts, n):
    """Block spectra FFT(series[s:s+n], zero-padded) / n, in FP64.

    src/matchfilt.c:458-470: scaling by 1/n happens before the transform,
    which is exact for power-of-two n.  A block past the end is zero-padded;
    a start beyond the series gives an all-zero block.
    """
    series = np.asarray(series, dtype=np.complex128)
    out = np.zeros((len(starts), n), dtype=np.complex128)
    for b, s in enumerate(starts):
        s = int(s)
        seg = series[s:s + n]
        blk = np.zeros(n, dtype=np.complex128)
        blk[:seg.size] = seg
        out[b] = np.fft.fft(blk / n)
    return out


def series_peaks(series, starts, win_start, win_end, T, binsize=None,
                 threshold=0.0):
    """run_series: per-block flat filter with per-block windows.

    Every block's window must give the same bin count
    (python/matchedfilter/_series.py:29-36); output is (nblocks, nt, nbins)
    in caller order.
    """
    T = np.atleast_2d(np.asarray(T))
    n = T.shape[-1]
    binsize = n if binsize is None else int(binsize)
    spec = series_spectra(series, starts, n)
    counts = set()
    rows = []
    for b in range(len(starts)):
        lo, hi = min(int(win_start[b]), n), min(int(win_end[b]), n)
        if lo >= hi:
            raise ValueError("every block must have a nonempty search window")
        counts.add(nbins(n, binsize, lo, hi))
        rows.append(flat_peaks(spec[b:b + 1], T, binsize, threshold, (lo, hi)))
    if len(counts) != 1:
        raise ValueError("every block's window must give the same bin count")
    idx = np.concatenate([r[0] for r in rows], axis=0)
    val = np.concatenate([r[1] for r in rows], axis=0)
    best = np.concatenate([r[2] for r in rows], axis=0)
    return idx, val, best


def peak_value_grad(D, T, k):
    """Gradients of z[k] = sum_f D[f] conj(T[f]) w^{fk} at a fixed lag k.

    With w = exp(2 pi i / n): dz/dD[f] = conj(T[f]) w^{fk} (holomorphic in D)
    and dz/dconj(T[f]) = D[f] w^{fk}.  For a real loss L(z), the Wirtinger
    gradient with respect to conj(T) is
        dL/dconj(T[f]) = conj(dL/dz_bar) ... ;
    this helper returns the two per-frequency coefficients
    ``(a_f, b_f) = (conj(T[f]) w^{fk}, D[f] w^{fk})`` in complex128.
    """
    D = np.asarray(D, dtype=np.complex128)
    T = np.asarray(T, dtype=np.complex128)
    n = D.shape[-1]
    f = np.arange(n)
    w = np.exp(2j * np.pi * f * int(k) / n)
    return np.conj(T) * w, D * w


def peak_grads(D, T, idx, g):
    """Gradients of a real loss through reported peak values, FP64.

    ``idx`` (nd, nt, nb) holds the stored lags (-1: nothing reported, a
    constant 0), ``g`` (nd, nt, nb) the incoming gradient dL/dRe z + i dL/dIm z
    of each value (PyTorch's convention for complex tensors).  With
    w = exp(2 pi i / n) and z = sum_f D[f] conj(T[f]) w^(f k):
        dL/dD_d[f] = sum_{t,b} g conj(dz/dD)        = sum g T_t[f] w^(-f k)
        dL/dT_t[f] = sum_{d,b} conj(g) dz/dconj(T)  = sum conj(g) D_d[f] w^(f k)
    (z is holomorphic in D and in conj(T)).  Returns (grad_D, grad_T).
    """
    D = np.asarray(D, dtype=np.complex128)
    T = np.asarray(T, dtype=np.complex128)
    idx = np.asarray(idx)
    g = np.asarray(g, dtype=np.complex128)
    nd, n = D.shape
    nt = T.shape[0]
    f = np.arange(n)
    gD = np.zeros((nd, n), np.complex128)
    gT = np.zeros((nt, n), np.complex128)
    for d in range(nd):
        for t in range(nt):
            for b in range(idx.shape[2]):
                k = int(idx[d, t, b])
                if k < 0:
                    continue
                w = np.exp(2j * np.pi * ((f * k) % n) / n)
                gD[d] += g[d, t, b] * T[t] * np.conj(w)
                gT[t] += np.conj(g[d, t, b]) * D[d] * w
    return gD, gT
