"""Parity rules shared by the tests: device peaks against the FP64 oracle.

Tolerance (stated once, used everywhere): peak values agree to
``TOL(n) = 2e-6 * log2(n)`` relative.  Indices must be identical except at a
documented near-tie:

* two lags whose FP64 magnitudes are within ``4 * TOL(n)`` of the bin's FP64
  maximum may swap (FP32 rounding decides between them);
* a bin whose FP64 maximum lies within ``4 * TOL(n)`` of the threshold may be
  reported on one side and empty on the other;
* a bin whose FP64 maximum is below ``4 * TOL(n)`` of the pair's largest |z|
  is numerically empty (FP32 rounding noise of an exact zero): its index is
  not compared and a reported value must itself be that small.
"""
import math

import numpy as np

from oracle import mf_oracle as O


def tol(n):
    return 2e-6 * math.log2(n)


def _bin_max(absz, binsize, start, end):
    seg = absz[..., start:end]
    nb = -(-(end - start) // binsize)
    pad = nb * binsize - seg.shape[-1]
    if pad:
        seg = np.concatenate([seg, np.zeros(seg.shape[:-1] + (pad,))], axis=-1)
    return seg.reshape(seg.shape[:-1] + (nb, binsize)).max(axis=-1)


def assert_peaks_match(D, T, idx, val, binsize, threshold, window=None, context="",
                       pair_chunk=256):
    """Compare device (idx, val) of shape (nd, nt, nb) with the FP64 oracle.

    Returns the number of index mismatches excused as near-ties.
    """
    D = np.atleast_2d(D)
    T = np.atleast_2d(T)
    nd, n = D.shape
    nt = T.shape[0]
    start, end = O.window(n, window)
    t = tol(n)
    thr = float(threshold)
    excused = 0
    bad = []
    # chunk over templates so the FP64 correlation stays small
    step = max(1, pair_chunk // max(1, nd))
    for t0 in range(0, nt, step):
        t1 = min(nt, t0 + step)
        z = O.correlate(D[:, None, :], T[None, t0:t1, :])
        ridx, rval, rbest = O.select(z, binsize, threshold, start, end)
        gi = idx[:, t0:t1]
        gv = val[:, t0:t1]
        assert gi.shape == ridx.shape, (gi.shape, ridx.shape, context)
        absz = np.abs(z)
        scale = absz[..., start:end].max(axis=-1)[..., None]
        bmax = _bin_max(absz, binsize, start, end)
        empty_num = bmax <= 4 * t * scale
        # numerically empty bins: a reported value must be tiny
        e_bad = empty_num & (gi >= 0) & (np.abs(gv) > 8 * t * scale)
        for d, tt, b in zip(*np.nonzero(e_bad)):
            bad.append((d, t0 + tt, b, "empty-bin value"))
        live = ~empty_num
        same = (gi == ridx) & live
        # values where indices agree and the bin reports
        rep = same & (ridx >= 0)
        err = np.abs(gv.astype(np.complex128) - rval)
        v_bad = rep & (err > t * np.abs(rval) + 1e-3 * t * scale)
        for d, tt, b in zip(*np.nonzero(v_bad)):
            bad.append((d, t0 + tt, b, "value", float(err[d, tt, b] / abs(rval[d, tt, b]))))
        # dismissed bins must carry an exact zero value
        z_bad = same & (ridx < 0) & (gv != 0)
        for d, tt, b in zip(*np.nonzero(z_bad)):
            bad.append((d, t0 + tt, b, "dismissed bin has a value"))
        # index mismatches: excuse documented near-ties only
        for d, tt, b in zip(*np.nonzero(live & (gi != ridx))):
            g, r = int(gi[d, tt, b]), int(ridx[d, tt, b])
            bm = float(bmax[d, tt, b])
            lo = start + b * binsize
            hi = min(end, lo + binsize)
            if g < 0 or r < 0:
                if thr > 0 and abs(bm - thr) <= 4 * t * max(thr, bm):
                    excused += 1
                    continue
                bad.append((d, t0 + tt, b, "reported/dismissed", g, r, bm, thr))
            elif not (lo <= g < hi):
                bad.append((d, t0 + tt, b, "index outside bin", g, r))
            elif absz[d, tt, g] >= bm * (1 - 4 * t):
                excused += 1
                zv = complex(z[d, tt, g])
                if abs(complex(gv[d, tt, b]) - zv) > t * abs(zv) + 1e-3 * t * float(scale[d, tt, 0]):
                    bad.append((d, t0 + tt, b, "near-tie value does not belong to its lag"))
            else:
                bad.append((d, t0 + tt, b, "index", g, r, float(absz[d, tt, g]), bm))
    assert not bad, "%s: %d mismatches, first %s" % (context, len(bad), bad[:5])
    return excused
