"""apogee_amd -- MI355X-native batched matched filter with peak-only output.

A drop-in for the ``device="gpu"`` surface of the reference package
``matchedfilter`` (python/matchedfilter/__init__.py): ``MatchedFilter`` and
``HierarchicalFilter`` with ``set_data`` / ``set_templates`` /
``set_reference``, ``run(binsize, threshold, window, data, templates, counts,
raw)``, ``run_series`` and the ``(ndata, ntemplates, nbins)`` result of
``PEAK_DTYPE`` records, whose below-threshold bins hold ``index == -1`` and
``value == 0``.

For each (data d, template t) pair the filter evaluates
``z = n * ifft(D_d * conj(T_t))`` and reports the loudest lag per bin.  The
work runs in hand-written HIP kernels for gfx950 (libmfhip.so, C ABI in
include/mfhip.h); the correlation series never leaves the chip.  There is no
CPU engine here: without the HIP library or a visible GPU, constructing a
filter raises.
"""
import numpy as np

from ._errors import UnsupportedSize
from . import _hip

__version__ = "0.1.0"

#: dtype of the arrays returned by :meth:`MatchedFilter.run`
#: (python/matchedfilter/__init__.py:45).
PEAK_DTYPE = np.dtype([("index", "<i8"), ("value", "<c8")])

#: Transform lengths of the reference's GPU path
#: (python/matchedfilter/__init__.py:58-59); which of them this build runs is
#: answered by ``_hip.supported`` (others raise UnsupportedSize).
_GPU_SIZES = frozenset((64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384,
                        32768, 65536))
#: Lengths this build adds beyond the reference's GPU set (its CPU engine's
#: larger power-of-two lengths), carried by the same four-step split.
_LARGE_SIZES = frozenset((131072, 262144, 524288, 1048576))

__all__ = ["MatchedFilter", "HierarchicalFilter", "RefineQueue", "PEAK_DTYPE", "devices",
           "Device", "UnsupportedSize", "__version__"]


def devices():
    """Every HIP device this build can dispatch to."""
    from .device import devices as _devices
    return _devices()


# ------------------------------------------------------------------ inputs
#: DLPack device types whose memory the host can read directly (CPU, CUDA
#: pinned host, ROCm pinned host).
_HOST_KINDS = frozenset((1, 3, 11))


def _dlpack_kind(a):
    """DLPack device type of ``a`` (1 = host when it does not say)."""
    probe = getattr(a, "__dlpack_device__", None)
    if probe is None:
        return 1
    try:
        return int(probe()[0])
    except Exception:
        return 1


def _on_rocm(a):
    """True for a ROCm-resident array this build reads through DLPack."""
    return _dlpack_kind(a) == _hip.DLPACK_ROCM and hasattr(a, "__dlpack__")


def _host_view(a):
    """A host array as NumPy (DLPack host producers are wrapped without a
    copy where NumPy can); arrays on any accelerator other than this build's
    ROCm device raise TypeError, because a silent transfer to the host would
    hide the cost the caller came to avoid."""
    if isinstance(a, np.ndarray):
        return a
    kind = _dlpack_kind(a)
    if kind not in _HOST_KINDS:
        raise TypeError("array is on a non-host device (DLPack type %d) that this build "
                        "cannot read; transfer it explicitly" % kind)
    if hasattr(a, "__dlpack__"):
        try:
            return np.from_dlpack(a)
        except Exception:       # read-only or exotic producers: NumPy coerces below
            pass
    return np.asarray(a)


def _c64_rows(a, shape, what):
    """Host complex64 C-contiguous array of exactly ``shape``."""
    out = np.ascontiguousarray(_host_view(a), dtype=np.complex64)
    if len(shape) == 1:
        if out.ndim != 1 or out.size != shape[0]:
            raise ValueError("%s must be a 1-D complex array of %d samples, got shape %s"
                             % (what, shape[0], out.shape))
    elif out.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, out.shape))
    return out


def _rocm_rows(ctx, a, shape, what):
    """Borrow a ROCm producer's complex64 buffer of ``shape`` for ``ctx``.

    A PyTorch tensor is made contiguous first; other producers must hand
    over compact (row-major) complex64 memory on the context's device."""
    if hasattr(a, "is_contiguous") and hasattr(a, "contiguous") and not a.is_contiguous():
        a = a.contiguous()
    view = ctx.borrow(a)
    if not view.complex64:
        raise ValueError("%s must be complex64" % what)
    got = view.shape if len(shape) > 1 else (int(np.prod(view.shape)),)
    if tuple(got) != tuple(shape):
        raise ValueError("expected shape %s, got %s" % (shape, view.shape))
    if not view.compact:
        raise ValueError("%s must be C-contiguous device memory" % what)
    if view.device != ctx.index:
        raise ValueError("%s is on device %d, the filter on %d" % (what, view.device, ctx.index))
    return view


# ----------------------------------------------------------------- results
def _records(idx, val):
    """(index, value) as one PEAK_DTYPE array."""
    out = np.empty(idx.shape, PEAK_DTYPE)
    out["index"] = idx
    out["value"] = val
    return out


def _result(idx, val, raw, counts):
    """The public return value: records (or the raw pair), plus the per-pair
    crossing counts when they were asked for."""
    body = (idx, val) if raw else _records(idx, val)
    return body if counts is None else (body, counts)


# ----------------------------------------------------- request resolution
class _RowBook:
    """Which rows of a bank have been set since the filter was made."""

    def __init__(self, rows):
        self._have = np.zeros(rows, dtype=bool)
        self._full = False

    def record(self, row):
        if row is None:
            self._full = True
        elif not self._full:
            self._have[row] = True
            self._full = bool(self._have.all())

    def forget(self):
        self._have[:] = False
        self._full = False

    def covers(self, first, count):
        return self._full or bool(self._have[first:first + count].all())


def _clamp_window(n, window):
    """Lag window (start, end) clamped to [0, n]; empty raises ValueError."""
    if window is None:
        return 0, n
    lo = min(max(int(window[0]), 0), n)
    hi = min(max(int(window[1]), 0), n)
    if hi <= lo:
        raise ValueError("empty window (%d, %d)" % (lo, hi))
    return lo, hi


def _bin_width(binsize, n):
    width = n if binsize is None else int(binsize)
    if width < 1:
        raise ValueError("binsize must be >= 1")
    return width


def _row_span(span, total):
    """(first, count) of a sub-range argument; None is the whole bank."""
    return (0, total) if span is None else (int(span[0]), int(span[1]))


def _row_index(index, rows):
    if index is None:
        return None
    row = int(index)
    if not 0 <= row < rows:
        raise IndexError("index %d out of range" % row)
    return row


class MatchedFilter:
    """Correlate ``ndata`` data spectra against ``ntemplates`` template spectra.

    Every pair (d, t) gives ``n * IFFT(data_d * conj(template_t))``; only the
    loudest sample per bin of the lag window is reported.  Spectra are the
    unnormalised forward transforms of the segments, natural order, and are
    kept on the device from the moment they are set.
    """

    #: Builds the device context; tests substitute a host double here.
    _context_factory = _hip.Context

    def __init__(self, n, ndata=1, ntemplates=1, device=None):
        from .device import parse as _parse_device
        self.n = int(n)
        self.ndata = int(ndata)
        self.ntemplates = int(ntemplates)
        if self.ndata < 1 or self.ntemplates < 1:
            raise ValueError("ndata and ntemplates must be >= 1")
        if self.n not in _GPU_SIZES | _LARGE_SIZES:
            raise ValueError(
                "device='gpu' supports n in %s; got %d" % (sorted(_GPU_SIZES | _LARGE_SIZES), self.n))
        self.device = _parse_device(device)
        self._init_state()
        self._gpu = self._context_factory(self.device.index, self.n, self.ndata, self.ntemplates)
        self._apply_tuning()

    def _apply_tuning(self):
        """Launch shapes from the gfx950 autotuner table (defaults if none)."""
        from . import autotune
        autotune.apply(self._gpu, "flat", self.n, self.ndata, self.ntemplates)
        if self.n >= 32768:
            autotune.apply(self._gpu, "big", self.n, self.ndata, self.ntemplates)

    def _init_state(self):
        self._books = {"data": _RowBook(self.ndata), "template": _RowBook(self.ntemplates)}
        self._has_data = False
        self._series_batch_bytes = 64 * 1024 * 1024
        self._cache_bytes = 512 * 1024 * 1024

    # ------------------------------------------------------------- ingest
    def _ingest(self, kind, spectra, row):
        """Put one row (``row`` an int) or the whole bank (``row`` None) on
        the device.  Host arrays are copied.  ROCm arrays are taken through
        DLPack, ordered after the producer's pending work on the device: a
        whole bank is read in place (bound, not copied), a single row is
        copied device to device."""
        rows = self.ndata if kind == "data" else self.ntemplates
        shape = (rows, self.n) if row is None else (self.n,)
        what = "spectra" if row is None else "spectrum"
        if _on_rocm(spectra):
            view = _rocm_rows(self._gpu, spectra, shape, what)
            if row is None:
                self._gpu.bind_rows(kind, view)
            else:
                self._gpu.copy_rows(kind, row, view)
            a = spectra
        else:
            a = _c64_rows(spectra, shape, what)
            self._gpu.set_rows(kind, 0 if row is None else row, a.reshape(-1, self.n))
        self._on_ingest(kind, row, a)

    def _on_ingest(self, kind, index, a):
        """Hook for subclasses that derive state from the spectra."""

    def set_data(self, spectra, index=None):
        """Set one data spectrum (with ``index``) or all, from (ndata, n).

        Host arrays are copied.  A ROCm array (any DLPack producer, e.g. a
        PyTorch tensor) of the whole bank is read in place by every later
        run: writing to it afterwards requires calling ``set_data`` again,
        which orders the next run after the write."""
        row = _row_index(index, self.ndata)
        self._ingest("data", spectra, row)
        self._has_data = True
        self._books["data"].record(row)

    def set_templates(self, spectra, index=None):
        """Set one template spectrum (with ``index``) or all, from
        (ntemplates, n); ROCm arrays as for :meth:`set_data`."""
        row = _row_index(index, self.ntemplates)
        self._ingest("template", spectra, row)
        self._books["template"].record(row)

    # ---------------------------------------------------------------- run
    def _window(self, window):
        return _clamp_window(self.n, window)

    def nbins(self, binsize, window=None):
        width = _bin_width(binsize, self.n)
        lo, hi = self._window(window)
        return (hi - lo + width - 1) // width

    def _require_templates(self, first, count):
        if not self._books["template"].covers(first, count):
            raise ValueError("no templates for requested rows: call set_templates() first")

    def _check_run(self, binsize, window, data, templates):
        """Resolve a run's arguments to (binsize, start, end, d0, nd, t0, nt),
        with the reference's errors and their order."""
        width = _bin_width(binsize, self.n)
        lo, hi = self._window(window)
        d0, nd = _row_span(data, self.ndata)
        t0, nt = _row_span(templates, self.ntemplates)
        inside = (min(nd, nt) >= 1 and min(d0, t0) >= 0
                  and d0 + nd <= self.ndata and t0 + nt <= self.ntemplates)
        if not inside:
            raise ValueError("data/templates sub-range out of bounds")
        if not (self._has_data and self._books["data"].covers(d0, nd)):
            raise ValueError("no data: call set_data() before run()")
        self._require_templates(t0, nt)
        return width, lo, hi, d0, nd, t0, nt

    def run(self, binsize=None, threshold=0.0, window=None,
            data=None, templates=None, counts=False, raw=False):
        """Correlate and report the loudest sample per bin.

        Returns a ``PEAK_DTYPE`` array of shape (nd, nt, nbins); with
        ``counts=True`` also the per-pair number of bins that crossed; with
        ``raw=True`` the plain ``(index, value)`` arrays.  ``data`` /
        ``templates`` = ``(start, count)`` restrict the run to a sub-range
        (identical to the matching slice of a full run); ``window=(start,
        end)`` restricts the lags searched.  Each result owns its storage
        (up to 64 MiB, page-locked host memory the device writes directly),
        so a later call never overwrites it, unlike the reference's reused
        output buffer (python/matchedfilter/__init__.py:463-470).
        """
        req = self._check_run(binsize, window, data, templates)
        width, lo, hi, d0, nd, t0, nt = req
        first, second, cnt = self._dispatch(d0, nd, t0, nt, width, threshold, lo, hi, counts, records=not raw)
        body = (first, second) if raw else first
        return body if not counts else (body, cnt)

    def _dispatch(self, d0, nd, t0, nt, binsize, threshold, start, end, counts, records=False):
        """(index, value, counts) of a run, or with ``records`` (PEAK_DTYPE
        records the library fills, None, counts)."""
        return self._gpu.peaks(d0, nd, t0, nt, binsize, float(threshold), start, end,
                               counts=bool(counts), records=PEAK_DTYPE if records else None)

    # ------------------------------------------------------------ series
    def _series_layout(self, series, starts, win_start, win_end, binsize, templates):
        """Check a run_series request (the reference's contract,
        python/matchedfilter/__init__.py:517-543) and build its layout."""
        from ._series import SeriesLayout, offsets_table, _FLAT
        if _on_rocm(series):
            import torch
            ser = torch.from_dlpack(series) if not isinstance(series, torch.Tensor) else series
            if ser.dim() != 1:
                raise ValueError(_FLAT)
            ser = ser.to(torch.complex64).contiguous()
        else:
            ser = np.ascontiguousarray(_host_view(series), dtype=np.complex64)
            if ser.ndim != 1:
                raise ValueError(_FLAT)
        table, low, high = offsets_table([np.asarray(_host_view(a)) for a in (starts, win_start, win_end)],
                                         self.n)
        width = _bin_width(binsize, self.n)
        t0, nt = _row_span(templates, self.ntemplates)
        if nt < 1 or t0 < 0 or t0 + nt > self.ntemplates:
            raise ValueError("templates sub-range out of bounds")
        self._require_templates(t0, nt)
        return ser, SeriesLayout(self.n, table, low, high, width), width, t0, nt

    def run_series(self, series, starts, win_start, win_end,
                   binsize=None, threshold=0.0, templates=None, raw=False):
        """Filter a time series over a caller-supplied block layout.

        Block b covers ``series[starts[b]:starts[b]+n]`` (zero-padded; its
        spectrum is FFT(block)/n) and reports lags [win_start[b],
        win_end[b]).  Every window must give the same number of bins.  The
        forward transforms and the filter run on the device.  Returns
        (nblocks, ntemplates, nbins) in caller order.  A later ``run()``
        requires ``set_data()`` again (python/matchedfilter/__init__.py:545-594).
        """
        ser, layout, width, t0, nt = self._series_layout(series, starts, win_start, win_end,
                                                         binsize, templates)
        self._has_data = False
        self._books["data"].forget()
        first, second = self._dispatch_series(ser, layout, width, float(threshold), t0, nt,
                                              records=not raw)
        return (first, second) if raw else first

    def _dispatch_series(self, ser, layout, binsize, threshold, t0, nt, records=False):
        """(index, value) of a series run, or with ``records`` (PEAK_DTYPE
        records the library fills, None)."""
        first, second, _ = self._gpu.run_series(ser, layout.starts, layout.low, layout.high, t0, nt,
                                                binsize, threshold,
                                                records=PEAK_DTYPE if records else None)
        return first, second

    # ---------------------------------------------------------- memory
    def empty_shared(self, shape, dtype=np.complex64):
        """Host array for the setters, in page-locked memory the device reads
        by DMA (the role of the reference's host-visible device storage,
        python/matchedfilter/_shared.py:58-76).  It is an ordinary NumPy
        array; PyTorch's pinned allocator backs it, and without PyTorch it is
        plain NumPy storage."""
        dt = np.dtype(dtype)
        try:
            import torch
        except ImportError:
            return np.empty(shape, dtype=dt)
        nbytes = int(np.prod(shape, dtype=np.int64)) * dt.itemsize
        buf = torch.empty(max(nbytes, 1), dtype=torch.uint8, pin_memory=True)
        return buf.numpy()[:nbytes].view(dt).reshape(shape)

    def set_memory_limits(self, *, cache_bytes=None, series_bytes=None):
        """Set the dispatch-cache and series working-storage budgets in bytes."""
        for value in (cache_bytes, series_bytes):
            if value is not None and int(value) < 1:
                raise ValueError("memory budgets must be positive")
        if cache_bytes is not None:
            self.clear_cache()
            self._cache_bytes = int(cache_bytes)
        if series_bytes is not None:
            self._series_batch_bytes = int(series_bytes)

    def clear_cache(self):
        """Release per-run device scratch, keeping the resident spectra."""


from .hierarchical import HierarchicalFilter  # noqa: E402
from .refine import RefineQueue  # noqa: E402


def __getattr__(name):
    if name == "Device":
        from .device import Device
        return Device
    raise AttributeError(name)
