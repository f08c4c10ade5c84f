"""A host double of apogee_amd._hip.Context, for testing the Python layer on
machines without a GPU.  It answers with the FP64 oracle (values cast to
complex64); it is test infrastructure and never reachable from the package.
"""
import numpy as np

from oracle import mf_oracle as O


class FakeContext:
    calls = []

    def __init__(self, index, n, ndata, ntemplates):
        self.index, self.n = index, n
        self.ndata, self.ntemplates = ndata, ntemplates
        self.D = np.zeros((ndata, n), np.complex64)
        self.T = np.zeros((ntemplates, n), np.complex64)
        self.band = None
        self.ref_f = -1.0
        self.last_refinements = 0
        FakeContext.calls.append(("create", n, ndata, ntemplates))

    def set_tuning(self, key, value):
        FakeContext.calls.append(("tune", key, value))

    def set_rows(self, kind, row0, a):
        a = np.asarray(a)
        dst = self.D if kind == "data" else self.T
        dst[row0:row0 + a.shape[0]] = a
        FakeContext.calls.append(("set", kind, row0, a.shape[0]))

    @staticmethod
    def _as_records(idx, val, records):
        """The library's records option: one structured array, then None."""
        if records is None:
            return idx, val
        rec = np.empty(idx.shape, records)
        rec[rec.dtype.names[0]] = idx
        rec[rec.dtype.names[1]] = val
        return rec, None

    def peaks(self, d0, nd, t0, nt, binsize, threshold, start, end, counts=False, records=None):
        FakeContext.calls.append(("peaks", d0, nd, t0, nt, binsize, threshold, start, end))
        idx, val, _ = O.flat_peaks(self.D[d0:d0 + nd], self.T[t0:t0 + nt], binsize, threshold,
                                   (start, end))
        cnt = (idx >= 0).sum(axis=2).astype(np.int32) if counts else None
        return self._as_records(idx, val.astype(np.complex64), records) + (cnt,)

    def hier_set_band(self, band, ref_fraction=-1.0):
        self.band, self.ref_f = band, ref_fraction

    def hier_set_coarse_precision(self, precision, delta):
        self.coarse_precision = (precision, delta)

    def hier_peaks(self, d0, nd, t0, nt, binsize, threshold, start, end, coarse_threshold,
                   counts=False, records=None):
        FakeContext.calls.append(("hier", d0, nd, t0, nt, coarse_threshold))
        T = self.T[t0:t0 + nt]
        if self.ref_f >= 0:
            f = np.full(nt, self.ref_f)
        else:
            f = O.band_fraction(T, self.band)
        sc = np.where(f > 0, 1 / np.sqrt(np.where(f > 0, f, 1)), 0)
        ct = (T[:, :self.band].astype(np.complex128) * sc[:, None]).astype(np.complex64)
        cs, ce = O.coarse_window(self.n, self.band, start, end)
        zc = O.correlate(self.D[d0:d0 + nd, None, :self.band], ct[None])
        cmax = np.abs(zc[..., cs:ce]).max(axis=-1)
        idx, val, _, fired = O.hier_peaks(self.D[d0:d0 + nd], T, self.band, coarse_threshold,
                                          binsize, threshold, (start, end), coarse_max=cmax)
        cnt = (idx >= 0).sum(axis=2).astype(np.int32) if counts else None
        self.last_refinements = int(fired.sum())
        return self._as_records(idx, val.astype(np.complex64), records) + (cnt, int(fired.sum()))

    def run_series(self, series, starts, win_start, win_end, t0, nt, binsize, threshold,
                   coarse_threshold=None, records=None):
        FakeContext.calls.append(("series", len(starts), t0, nt, binsize, threshold, coarse_threshold))
        T = self.T[t0:t0 + nt]
        idx, val, _ = O.series_peaks(np.asarray(series), starts, win_start, win_end, T, binsize,
                                     threshold)
        return self._as_records(idx, val.astype(np.complex64), records) + (0,)


class FakeQueue:
    """Host double of apogee_amd._hip.Queue: plans in Python and answers each
    pair with the FP64 oracle's flat filter."""

    def __init__(self, index):
        self.index = index
        self.pool = None
        self.plan_args = None

    def set_tuning(self, key, value):
        pass

    def set_pool(self, pool):
        self.pool = np.array(pool, dtype=np.complex64)

    def plan(self, n, data_off, tmpl_off, ncand, binsize, threshold, ws, we):
        n = np.asarray(n, np.int64)
        nb = (np.asarray(we) - np.asarray(ws) + np.asarray(binsize) - 1) // np.asarray(binsize)
        sizes = np.asarray(ncand, np.int64) * nb
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.plan_args = (n, data_off, tmpl_off, ncand, binsize, threshold, ws, we, nb, off)
        return off, int(sizes.sum())

    def run(self, nout, npairs):
        n, doff, toff, cand, bs, thr, ws, we, nb, off = self.plan_args
        idx = np.empty(nout, np.int64)
        val = np.empty(nout, np.complex64)
        cnt = []
        for r in range(n.size):
            L, c = int(n[r]), int(cand[r])
            if c == 0:
                continue
            D = self.pool[doff[r]:doff[r] + L][None]
            T = self.pool[toff[r]:toff[r] + c * L].reshape(c, L)
            i, v, _ = O.flat_peaks(D, T, int(bs[r]), float(thr[r]), (int(ws[r]), int(we[r])))
            k = c * int(nb[r])
            idx[off[r]:off[r] + k] = i.reshape(-1)
            val[off[r]:off[r] + k] = v.reshape(-1)
            cnt.extend((i[0] >= 0).sum(axis=1).tolist())
        return idx, val, np.asarray(cnt, np.int32)
