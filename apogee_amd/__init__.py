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

#: DLPack device types read as host memory: CPU, CUDA host, ROCm host.
_DLPACK_HOST = {1, 3, 11}
_DLPACK_ROCM = 10


def devices():
    """Every HIP device this build can dispatch to."""
    from .device import devices as _devices
    return _devices()


def _from_any(a):
    """numpy arrays pass through; DLPack host arrays are wrapped; ROCm
    tensors pass through untouched for a device-to-device copy; any other
    accelerator array is refused (python/matchedfilter/__init__.py:96-131)."""
    if isinstance(a, np.ndarray):
        return a
    if hasattr(a, "__dlpack_device__"):
        try:
            kind = int(a.__dlpack_device__()[0])
        except Exception:
            kind = 1
        if kind == _DLPACK_ROCM and hasattr(a, "data_ptr"):
            return a
        if kind not in _DLPACK_HOST:
            raise TypeError("array is on a non-host device (DLPack type %d) that "
                            "this build cannot read; transfer it explicitly" % kind)
        try:
            return np.from_dlpack(a)
        except Exception:
            pass
    return a


def _is_device_tensor(a):
    return hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


def _device_c64(a, shape, what):
    """Validate a ROCm tensor as complex64 of ``shape``; returns it contiguous."""
    import torch
    if a.dtype != torch.complex64:
        raise ValueError("%s must be complex64, got %s" % (what, a.dtype))
    if tuple(a.shape) != tuple(shape):
        raise ValueError("expected shape %s, got %s" % (tuple(shape), tuple(a.shape)))
    return a.contiguous()


def _as_c64(a, n, what):
    a = np.ascontiguousarray(_from_any(a), dtype=np.complex64)
    if a.ndim != 1 or a.size != n:
        raise ValueError("%s must be a 1-D complex array of %d samples, got shape %s"
                         % (what, n, a.shape))
    return a


def _format_result(idx, val, *, raw=False, counts=None, order=None):
    """Assemble PEAK_DTYPE records or return raw arrays, optionally restoring
    caller order (python/matchedfilter/__init__.py:155-173)."""
    if counts is True:
        counts = (idx >= 0).sum(axis=2).astype(np.int32)
    if raw:
        if order is None:
            result = (idx, val)
        else:
            ri = np.empty(idx.shape, np.int64)
            rv = np.empty(val.shape, np.complex64)
            ri[order], rv[order] = idx, val
            result = (ri, rv)
    else:
        result = np.empty(idx.shape, PEAK_DTYPE)
        if order is None:
            result["index"], result["value"] = idx, val
        else:
            result["index"][order], result["value"][order] = idx, val
    if counts is not None and counts is not False:
        return result, counts
    return result


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

    def _init_state(self):
        self._dataset = False
        self._data_ready = set()
        self._template_ready = set()
        self._series_batch_bytes = 64 * 1024 * 1024
        self._cache_bytes = 512 * 1024 * 1024

    # ------------------------------------------------------------- ingest
    def _input_index(self, index, size):
        if index is None:
            return None
        index = int(index)
        if index < 0 or index >= size:
            raise IndexError("index %d out of range" % index)
        return index

    def _mark_ready(self, kind, index):
        attr = "_" + kind + "_ready"
        ready = getattr(self, attr)
        if index is None:
            setattr(self, attr, None)
        elif ready is not None:
            ready.add(index)
            size = self.ndata if kind == "data" else self.ntemplates
            if len(ready) == size:
                setattr(self, attr, None)

    def _missing(self, kind, start, count):
        ready = getattr(self, "_" + kind + "_ready")
        return ready is not None and any(i not in ready for i in range(start, start + count))

    def _require_templates(self, start, count):
        if self._missing("template", start, count):
            raise ValueError("no templates for requested rows: call set_templates() first")

    def _ingest(self, kind, spectra, index):
        rows = self.ndata if kind == "data" else self.ntemplates
        a = _from_any(spectra)
        if index is None:
            if _is_device_tensor(a):
                a = _device_c64(a, (rows, self.n), "spectra")
            else:
                a = np.ascontiguousarray(a, dtype=np.complex64)
                if a.shape != (rows, self.n):
                    raise ValueError("expected shape (%d, %d), got %s" % (rows, self.n, a.shape))
            self._gpu.set_rows(kind, 0, a)
        else:
            if _is_device_tensor(a):
                a = _device_c64(a.reshape(-1), (self.n,), "spectrum").reshape(1, self.n)
            else:
                a = _as_c64(a, self.n, "spectrum")[None, :]
            self._gpu.set_rows(kind, index, a)
        self._on_ingest(kind, index, a)

    def _on_ingest(self, kind, index, a):
        """Hook for subclasses that derive state from the spectra."""

    def set_data(self, spectra, index=None):
        """Set one data spectrum (with ``index``) or all, from (ndata, n)."""
        index = self._input_index(index, self.ndata)
        self._ingest("data", spectra, index)
        self._dataset = True
        self._mark_ready("data", index)

    def set_templates(self, spectra, index=None):
        """Set one template spectrum (with ``index``) or all, from (ntemplates, n)."""
        index = self._input_index(index, self.ntemplates)
        self._ingest("template", spectra, index)
        self._mark_ready("template", index)

    # ---------------------------------------------------------------- run
    def _window(self, window):
        if window is None:
            return 0, self.n
        start, end = int(window[0]), int(window[1])
        start = max(0, min(start, self.n))
        end = max(0, min(end, self.n))
        if start >= end:
            raise ValueError("empty window (%d, %d)" % (start, end))
        return start, end

    def nbins(self, binsize, window=None):
        binsize = int(binsize)
        if binsize < 1:
            raise ValueError("binsize must be >= 1")
        start, end = self._window(window)
        return -(-(end - start) // binsize)

    def _check_run(self, binsize, window, data, templates):
        binsize = self.n if binsize is None else int(binsize)
        if binsize < 1:
            raise ValueError("binsize must be >= 1")
        start, end = self._window(window)
        d0, nd = (0, self.ndata) if data is None else (int(data[0]), int(data[1]))
        t0, nt = (0, self.ntemplates) if templates is None else (int(templates[0]), int(templates[1]))
        if nd < 1 or nt < 1 or d0 < 0 or t0 < 0 \
                or d0 + nd > self.ndata or t0 + nt > self.ntemplates:
            raise ValueError("data/templates sub-range out of bounds")
        if not self._dataset or self._missing("data", d0, nd):
            raise ValueError("no data: call set_data() before run()")
        self._require_templates(t0, nt)
        return binsize, start, end, d0, nd, t0, nt

    def run(self, binsize=None, threshold=0.0, window=None,
            data=None, templates=None, counts=False, raw=False):
        """Correlate and report the loudest sample per bin.

        Returns a ``PEAK_DTYPE`` array of shape (nd, nt, nbins); with
        ``counts=True`` also the per-pair number of bins that crossed; with
        ``raw=True`` the plain ``(index, value)`` arrays.  ``data`` /
        ``templates`` = ``(start, count)`` restrict the run to a sub-range
        (identical to the matching slice of a full run); ``window=(start,
        end)`` restricts the lags searched.
        """
        binsize, start, end, d0, nd, t0, nt = self._check_run(binsize, window, data, templates)
        idx, val, cnt = self._dispatch(d0, nd, t0, nt, binsize, threshold, start, end, counts)
        return _format_result(idx, val, raw=raw, counts=cnt if counts else None)

    def _dispatch(self, d0, nd, t0, nt, binsize, threshold, start, end, counts):
        return self._gpu.peaks(d0, nd, t0, nt, binsize, float(threshold), start, end,
                               counts=bool(counts))

    # ------------------------------------------------------------ series
    def _series_layout(self, series, starts, win_start, win_end, binsize, templates):
        """Validate the run_series contract (python/matchedfilter/__init__.py:517-543)."""
        ser = _from_any(series)
        if _is_device_tensor(ser):
            import torch
            if ser.dim() != 1:
                raise ValueError("series, starts, win_start and win_end must be one-dimensional")
            ser = ser.to(torch.complex64).contiguous()
        else:
            ser = np.ascontiguousarray(ser, dtype=np.complex64)
            if ser.ndim != 1:
                raise ValueError("series, starts, win_start and win_end must be one-dimensional")
        arrs = [np.asarray(_from_any(a)) for a in (starts, win_start, win_end)]
        if any(a.ndim != 1 for a in arrs):
            raise ValueError("series, starts, win_start and win_end must be one-dimensional")
        if any(np.issubdtype(a.dtype, np.signedinteger) and (a < 0).any() for a in arrs):
            raise ValueError("starts and windows must be nonnegative")
        st, ws, we = (np.ascontiguousarray(a, dtype=np.uintp) for a in arrs)
        if not (st.size == ws.size == we.size):
            raise ValueError("starts, win_start and win_end must be the same length")
        if st.size < 1:
            raise ValueError("run_series needs at least one block")
        if np.any(st > np.iinfo(np.intp).max - self.n):
            raise ValueError("starts must be nonnegative and fit in the sample index range")
        binsize = self.n if binsize is None else int(binsize)
        if binsize < 1:
            raise ValueError("binsize must be >= 1")
        t0, nt = (0, self.ntemplates) if templates is None else (int(templates[0]), int(templates[1]))
        if nt < 1 or t0 < 0 or t0 + nt > self.ntemplates:
            raise ValueError("templates sub-range out of bounds")
        self._require_templates(t0, nt)
        from ._series import SeriesLayout
        layout = SeriesLayout(self.n, st, ws, we, binsize)
        return ser, layout, binsize, t0, nt

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
        ser, layout, binsize, t0, nt = self._series_layout(series, starts, win_start, win_end,
                                                           binsize, templates)
        self._dataset = False
        self._data_ready = set()
        idx, val = self._dispatch_series(ser, layout, binsize, float(threshold), t0, nt)
        return _format_result(idx, val, raw=raw)

    def _dispatch_series(self, ser, layout, binsize, threshold, t0, nt):
        idx, val, _ = self._gpu.run_series(ser, layout.starts, layout.low, layout.high, t0, nt,
                                           binsize, threshold)
        return idx, val

    # ---------------------------------------------------------- memory
    def empty_shared(self, shape, dtype=np.complex64):
        """Host array suitable for the setters (ordinary NumPy storage)."""
        return np.empty(shape, dtype=dtype)

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
