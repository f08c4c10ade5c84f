"""HierarchicalFilter: a coarse band screen on the device, then the exact
flat filter on the pairs that survive it.

Mirrors python/matchedfilter/__init__.py:1055-1461 of the reference
(constructor validation, ``set_reference``, ``set_coarse_threshold``,
``set_first_stage``, ``config``, ``stats``, ``refine_rate``).  Execution is
one device call per run (libmfhip ``mfh_hier_peaks``): the coarse templates
T[:band]/sqrt(f) are built on the device with FP64 band fractions, the coarse
pass is the band-point flat filter over the coarse window, the gate compacts
survivors into a device list and the refine kernel drains it; dismissed pairs
report -1 / 0.  Every reported peak is the flat filter's peak; only omissions
are possible.

Configuration: an explicit ``band`` (and optionally ``set_coarse_threshold``)
is used as given.  Otherwise (``set_reference`` required) the band is the
cheapest on this device by measured cost, as the reference ranks its
measured cost rows (``choose_config``, python/matchedfilter/__init__.py:
996-1052): for every band whose modelled gate resolves the budget
(``gatemodel.gate_for`` at (snr, fd)), the expected cost of a call of this
filter's ndata x ntemplates pairs is the measured screen cost plus the
modelled noise refine rate at that gate (``gatemodel.noise_exceedance``)
times the measured cost of refining every pair, both interpolated in the
pair count from the gfx950 table (``autotune``, op ``band``); the smallest
cost wins.  Lengths the table lacks fall back to the band n/8 (at most
16384).
"""
import numpy as np

from . import PEAK_DTYPE, MatchedFilter, _host_view

#: Coarse-screen precisions: name -> (code of mfh_hier_set_coarse_precision,
#: default gate widening).  The widening bounds the relative error of the
#: matrix-core statistic against the FP32 screen with headroom: the largest
#: error measured on noise, loud and inspiral-bank pairs across windows
#: (tests/test_gpu_mfma.py) was 4.9e-4 (fp16) and 4.1e-3 (bf16); the
#: defaults are about 4x that.
_COARSE_PRECISIONS = {"fp32": (0, 0.0), "fp16": (1, 2e-3), "bf16": (2, 1.6e-2)}

#: The largest band the one-pass coarse screen runs.
_BAND_MAX = 16384


def _in_pairs(values, at, pairs):
    """Cost at ``pairs`` from costs measured at two batch sizes, linear in
    the pair count (a fixed launch part plus a per-pair part)."""
    (p0, p1), (c0, c1) = at, values
    if p1 == p0:
        return float(c1)
    slope = (c1 - c0) / (p1 - p0)
    return max(float(c0 + slope * (pairs - p0)), 0.0)


def rank_bands(power, n, snr, fd, pairs, precision="fp32", entry=None):
    """Candidate bands of length n, cheapest first, from the measured cost
    entry (default: the gfx950 table's): a list of dicts with band, gate,
    modelled noise refine rate and expected milliseconds per call.  Bands
    whose gate cannot resolve fd are left out."""
    from . import autotune
    from .gatemodel import gate_for, noise_exceedance
    entry = autotune.band_entry(n) if entry is None else entry
    if entry is None:
        return []
    out = []
    for key, row in entry["bands"].items():
        band = int(key)
        if band >= n or band > _BAND_MAX:
            continue
        gate = gate_for(power, n, band, snr, fd)
        if gate is None:
            continue
        rate = noise_exceedance(power, n, band, gate)
        screen = row.get("screen_ms_fp16") if (precision != "fp32" and band == 512) else None
        screen = row["screen_ms"] if screen is None else screen
        cost = (_in_pairs(screen, entry["pairs"], pairs)
                + rate * _in_pairs(row["refine_ms_all"], entry["pairs"], pairs))
        out.append({"band": band, "gate": gate, "noise_refine_rate": rate, "ms": cost})
    return sorted(out, key=lambda c: (c["ms"], c["band"]))


class HierarchicalFilter(MatchedFilter):
    """Matched filter that screens pairs on the low band first.

    ``snr`` is the |rho| of the weakest signal that must be kept and ``fd`` the
    tolerated false-dismissal probability for it.  Compare values across
    devices within float32 roundoff; the coarse gate can omit peaks, never
    invent or move one.
    """

    def __init__(self, n, ndata=1, ntemplates=1, snr=5.5, fd=1e-2,
                 band=None, taps=None, device=None):
        n = int(n)
        if band is not None:
            band = int(band)
            if band < 64 or band >= n or band & (band - 1):
                raise ValueError("band must be a power of two, >= 64 and < n")
        if taps is not None:
            taps = int(taps)
            if taps < 2 or taps > 64 or taps % 2:
                raise ValueError("taps must be even and between 2 and 64")
        super().__init__(n, ndata, ntemplates, device)
        self.snr = float(snr)
        self.fd = float(fd)
        self._pinned = (band, int(taps or 8)) if band is not None else None
        self._ref = None
        self._cal_thr = None
        self._fs_snr = None
        self._cal = None
        self._band_set = None
        self._pairs = 0
        self._trig = 0
        self._prec = ("fp32", 0.0)
        self._prec_set = None
        self._auto = None

    def _apply_tuning(self):
        from . import autotune
        autotune.apply(self._gpu, "flat", self.n, self.ndata, self.ntemplates)
        if self.n >= 32768:
            autotune.apply(self._gpu, "big", self.n, self.ndata, self.ntemplates)
        autotune.apply(self._gpu, "hier", self.n, self.ndata, self.ntemplates)
        autotune.apply(self._gpu, "coarse", self.n, self.ndata, self.ntemplates)   # the band-512 MFMA screen

    # ----------------------------------------------------------- settings
    def set_reference(self, power):
        """Reference output-power profile (length n, any scale), or None."""
        if power is not None:
            p = np.ascontiguousarray(_host_view(power), dtype=np.float32)
            if p.shape != (self.n,):
                raise ValueError("reference must be a one-dimensional array of length %d" % self.n)
            if not np.isfinite(p).all() or np.any(p < 0) or not np.any(p > 0):
                raise ValueError("reference must be finite, nonnegative, with positive total power")
            self._ref = p
        else:
            self._ref = None
        self._cal = None

    def set_coarse_threshold(self, value):
        """Set the coarse gate directly (requires an explicit band); None
        returns to the reference model."""
        if value is None:
            self._cal_thr = None
            self._cal = None
            return
        value = float(value)
        if not np.isfinite(value) or value < 0 or value > float(np.finfo(np.float32).max):
            raise ValueError("coarse threshold must be finite nonnegative float32, or None")
        if self._pinned is None:
            raise ValueError("an explicit coarse threshold requires an explicit band")
        self._cal_thr = value
        self._cal = None

    def set_first_stage(self, snr):
        """Calibrate the coarse gate at ``snr`` instead of the constructor's;
        None or a non-positive value restores it."""
        self._fs_snr = float(snr) if snr is not None and float(snr) > 0 else None
        self._cal = None

    def set_coarse_precision(self, precision="fp16", delta=None):
        """Run the coarse screen on the matrix cores ("fp16" or "bf16", band
        512) or in FP32 ("fp32", the default; any band).  The matrix-core
        statistic is widened by ``1 + delta`` at the gate (default: the
        measured bound of this build), so it keeps at least the pairs the FP32
        screen keeps whenever its relative error is within ``delta``; reported
        peaks are unaffected either way."""
        if precision not in _COARSE_PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(_COARSE_PRECISIONS))
        d = _COARSE_PRECISIONS[precision][1] if delta is None else float(delta)
        if not np.isfinite(d) or d < 0 or d >= 1:
            raise ValueError("delta must be in [0, 1)")
        self._prec = (precision, d)

    @property
    def coarse_precision(self):
        """``(precision, delta)`` of the coarse screen."""
        return self._prec

    def coarse_statistic(self, precision=None, data=None, templates=None, window=None):
        """The coarse maximum |z_c| per pair at a precision (default: the
        configured one), shape (nd, nt); no refine runs.  Compare "fp16" or
        "bf16" against "fp32" to measure the widening a bank needs."""
        band, f, _ = self._calibration_band()
        prec = self._prec[0] if precision is None else precision
        if prec not in _COARSE_PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(_COARSE_PRECISIONS))
        start, end = self._window(window)
        d0, nd = (0, self.ndata) if data is None else (int(data[0]), int(data[1]))
        t0, nt = (0, self.ntemplates) if templates is None else (int(templates[0]), int(templates[1]))
        self._apply_band(band, f)
        return self._gpu.hier_coarse_max(d0, nd, t0, nt, start, end, _COARSE_PRECISIONS[prec][0])

    def _calibration_band(self):
        band = self._band()
        f = -1.0
        if self._ref is not None:
            ref = self._ref.astype(np.float64)
            f = float(ref[:band].sum() / ref.sum())
        return band, f, None

    def _apply_band(self, band, f):
        if self._band_set != (band, f):
            self._gpu.hier_set_band(band, f)
            self._band_set = (band, f)
        prec, delta = self._prec
        code = _COARSE_PRECISIONS[prec][0] if band == 512 else 0
        if self._prec_set != (code, delta):
            self._gpu.hier_set_coarse_precision(code, delta)
            self._prec_set = (code, delta)

    # ------------------------------------------------------- calibration
    def _band(self):
        if self._pinned is not None:
            return self._pinned[0]
        if self._ref is None:
            raise ValueError("set_reference is required for automatic configuration selection")
        key = (self.snr, self.fd, self._fs_snr, self._prec[0], self._ref.tobytes())
        if self._auto is not None and self._auto[0] == key:
            return self._auto[1]
        ranked = rank_bands(self._ref, self.n, self._fs_snr or self.snr, self.fd,
                            self.ndata * self.ntemplates, self._prec[0])
        if ranked:
            band = ranked[0]["band"]
        else:
            band = min(self.n // 8, _BAND_MAX)
        self._auto = (key, band)
        return band

    def _calibration(self):
        key = (self.snr, self.fd, self._fs_snr, self._pinned, self._cal_thr, self._prec[0],
               None if self._ref is None else self._ref.tobytes())
        if self._cal is not None and self._cal[0] == key:
            return self._cal[1]
        band = self._band()
        if self._cal_thr is not None:
            thr = float(self._cal_thr)
        else:
            if self._ref is None:
                raise ValueError(
                    "no calibrated coarse threshold for n=%d band=%d: provide a reference "
                    "or set both band and coarse threshold" % (self.n, band))
            from .gatemodel import gate_for
            thr = gate_for(self._ref, self.n, band, self._fs_snr or self.snr, self.fd)
            if thr is None:
                raise ValueError("the false-dismissal budget fd=%g cannot be resolved" % self.fd)
            if not np.isfinite(thr) or thr < 0 or thr > float(np.finfo(np.float32).max):
                raise ValueError("calibrated coarse threshold must be finite, nonnegative float32")
        f = -1.0
        if self._ref is not None:
            ref = self._ref.astype(np.float64)
            f = float(ref[:band].sum() / ref.sum())
        out = (band, f, thr)
        self._cal = (key, out)
        return out

    def _dispatch(self, d0, nd, t0, nt, binsize, threshold, start, end, counts, records=False):
        band, f, thr = self._calibration()
        self._apply_band(band, f)
        idx, val, cnt, refined = self._gpu.hier_peaks(d0, nd, t0, nt, binsize, float(threshold),
                                                      start, end, thr, counts=bool(counts),
                                                      records=PEAK_DTYPE if records else None)
        self._pairs += nd * nt
        self._trig += int(refined)
        return idx, val, cnt

    def _dispatch_series(self, ser, layout, binsize, threshold, t0, nt, records=False):
        band, f, thr = self._calibration()
        self._apply_band(band, f)
        first, second, refined = self._gpu.run_series(ser, layout.starts, layout.low, layout.high, t0, nt,
                                                      binsize, threshold, coarse_threshold=thr,
                                                      records=PEAK_DTYPE if records else None)
        self._pairs += first.shape[0] * first.shape[1]
        self._trig += int(refined)
        return first, second

    def _on_ingest(self, kind, index, a):
        if kind == "template":
            self._band_set = None

    # ----------------------------------------------------------- reports
    @property
    def config(self):
        """Selected or pinned ``(band, taps)``."""
        if self._pinned is not None:
            return self._pinned
        return (self._band(), 8)

    @property
    def cost_table(self):
        """Which cost table chose the band: None when pinned, "gfx950" for
        the measured table (or its n/8 fallback for lengths it lacks)."""
        return None if self._pinned is not None else "gfx950"

    @property
    def coarse_threshold(self):
        """The coarse gate in use (resolving it if needed)."""
        return self._calibration()[2]

    @property
    def stats(self):
        """``(pairs, triggers)`` accumulated since construction."""
        return (self._pairs, self._trig)

    @property
    def refine_rate(self):
        """Fraction of pairs that needed the full correlation."""
        return self._trig / self._pairs if self._pairs else 0.0
