"""ctypes binding of libmfhip.so and the device Context used by the filters.

The Context plays the role of the reference's backend Context protocol
(python/matchedfilter/_vkcompute.py: ``peaks`` :962-1063, ``hier_peaks``
:534-623, ``forward`` :855-906): the public classes call it at exactly the
places the reference calls ``self._gpu.peaks`` / ``.hier_peaks`` /
``.forward``.  Unlike the reference it keeps the spectra on the device from
the moment they are set, so nothing is re-uploaded per run.

There is no fallback: if the HIP library is missing or no HIP device is
visible, constructing a Context raises.
"""
import ctypes
import os
import threading

import numpy as np

from ._errors import UnsupportedSize

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfhip.so")

_lib = None
_lock = threading.Lock()

_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p


def _declare(lib):
    lib.mfh_last_error.restype = ctypes.c_char_p
    lib.mfh_device_count.restype = ctypes.c_int
    lib.mfh_device_info.argtypes = [ctypes.c_int, ctypes.c_char_p, _c_size, ctypes.c_char_p, _c_size]
    lib.mfh_supported.argtypes = [_c_size]
    lib.mfh_create.restype = _vp
    lib.mfh_create.argtypes = [ctypes.c_int, _c_size, ctypes.c_int, ctypes.c_int]
    lib.mfh_destroy.argtypes = [_vp]
    lib.mfh_destroy.restype = None
    lib.mfh_stream.argtypes = [_vp]
    lib.mfh_stream.restype = _vp
    lib.mfh_sync.argtypes = [_vp]
    for name in ("mfh_set_data", "mfh_set_templates", "mfh_set_data_device_async",
                 "mfh_set_templates_device_async"):
        getattr(lib, name).argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
    lib.mfh_nbins.restype = _c_size
    lib.mfh_nbins.argtypes = [_c_size, _c_size, _c_size, _c_size]
    lib.mfh_flat_peaks.restype = ctypes.c_long
    lib.mfh_flat_peaks.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _c_size, ctypes.c_float, _c_size, _c_size, _vp, _vp, _vp]
    lib.mfh_flat_peaks_async.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _c_size, ctypes.c_float, _c_size, _c_size, _vp, _vp]
    lib.mfh_hier_set_band.argtypes = [_vp, _c_size, ctypes.c_double]
    lib.mfh_hier_peaks.restype = ctypes.c_long
    lib.mfh_hier_peaks.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _c_size, ctypes.c_float, _c_size, _c_size, ctypes.c_float,
                                   _vp, _vp, _vp, _vp]
    lib.mfh_hier_peaks_async.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _c_size, ctypes.c_float, _c_size, _c_size, ctypes.c_float,
                                         _vp, _vp, _vp]
    for name in ("mfh_run_series", "mfh_hier_run_series"):
        getattr(lib, name).restype = ctypes.c_long
    lib.mfh_run_series.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _c_size, ctypes.c_float, _vp, _vp, _vp]
    lib.mfh_hier_run_series.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _c_size, ctypes.c_float,
                                        ctypes.c_float, _vp, _vp, _vp, _vp]
    lib.mfh_series_forward.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]
    lib.mfh_hier_set_coarse_precision.argtypes = [_vp, ctypes.c_int, ctypes.c_float]
    lib.mfh_hier_coarse_max.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _c_size, _c_size, ctypes.c_int, _vp]
    lib.mfh_selftest_mfma.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    lib.mfh_set_tuning.argtypes = [_vp, ctypes.c_char_p, ctypes.c_long]
    lib.mfh_prof_enable.argtypes = [_vp, ctypes.c_int]
    lib.mfh_prof_read.argtypes = [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_long)]
    lib.mfh_peak_grad_async.argtypes = [_vp, _vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_longlong,
                                         ctypes.c_int, _c_size, _vp, _vp, _vp, _vp]
    lib.mfh_peak_values_async.argtypes = [_vp, _vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_longlong,
                                          ctypes.c_int, _c_size, _vp, _vp]
    lib.mfh_queue_create.restype = _vp
    lib.mfh_queue_create.argtypes = [ctypes.c_int]
    lib.mfh_queue_destroy.restype = None
    lib.mfh_queue_destroy.argtypes = [_vp]
    lib.mfh_queue_set_pool.argtypes = [_vp, _vp, _c_size, ctypes.c_int]
    lib.mfh_queue_plan.restype = ctypes.c_longlong
    lib.mfh_queue_plan.argtypes = [_vp, ctypes.c_int] + [_vp] * 9
    lib.mfh_queue_run_async.argtypes = [_vp, _vp, _vp]
    lib.mfh_queue_run.restype = ctypes.c_longlong
    lib.mfh_queue_run.argtypes = [_vp, _vp, _vp, _vp]
    lib.mfh_queue_set_tuning.argtypes = [_vp, ctypes.c_char_p, ctypes.c_long]
    lib.mfh_queue_sync.argtypes = [_vp]
    lib.mfh_queue_stream.restype = _vp
    lib.mfh_queue_stream.argtypes = [_vp]
    lib.mfh_queue_prof_enable.argtypes = [_vp, ctypes.c_int]
    lib.mfh_queue_prof_read.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    return lib


def _share_torch_hip_runtime():
    """Make libmfhip.so and PyTorch use ONE HIP runtime.

    PyTorch-ROCm bundles its own libamdhip64 and names it in DT_NEEDED as
    "libamdhip64.so"; libmfhip.so names the system runtime by its soname
    "libamdhip64.so.7".  Loaded in the other order the process ends up with
    two runtimes and PyTorch's finds no device.  Preloading PyTorch's copy
    (same soname) first makes both resolve to it.  Nothing happens when
    PyTorch is not installed.
    """
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def lib():
    """The loaded HIP library; raises RuntimeError if it was never built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "the HIP library %s is missing: build it with "
                    "`python -m apogee_amd._build` (hipcc, gfx950)" % LIB_PATH)
            _share_torch_hip_runtime()
            _lib = _declare(ctypes.CDLL(LIB_PATH))
        return _lib


def last_error():
    return lib().mfh_last_error().decode(errors="replace")


def check(status, what="mfh"):
    if status < 0:
        raise RuntimeError("%s failed: %s" % (what, last_error()))
    return status


def device_count():
    return int(lib().mfh_device_count())


def device_info(index):
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    check(lib().mfh_device_info(int(index), name, 256, arch, 64), "mfh_device_info")
    return name.value.decode(), arch.value.decode()


def supported(n):
    return bool(lib().mfh_supported(int(n)))


def selftest_mfma(device, variant, A, B):
    """Run one MFMA variant on FP32 matrices (see include/mfhip.h)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    C = np.empty((16, 16), dtype=np.float32)
    check(lib().mfh_selftest_mfma(int(device), int(variant), _ptr(A), _ptr(B), _ptr(C)),
          "mfh_selftest_mfma")
    return C


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _device_pointer(a):
    """Device address of a ROCm tensor (torch), or None for host arrays."""
    if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
        return int(a.data_ptr())
    return None


class Context:
    """Device state of one filter: spectra resident in HBM plus a HIP stream.

    ``n``, ``ndata`` and ``ntemplates`` fix the device allocation; rows are
    written by :meth:`set_rows` and read by every run until overwritten.
    """

    def __init__(self, index, n, ndata, ntemplates):
        L = lib()
        if not supported(n):
            raise UnsupportedSize("transform length %d is not supported on this device" % n)
        self.index = int(index)
        self.n, self.ndata, self.ntemplates = int(n), int(ndata), int(ntemplates)
        h = L.mfh_create(self.index, self.n, self.ndata, self.ntemplates)
        if not h:
            raise RuntimeError("mfh_create failed: %s" % last_error())
        self._h = ctypes.c_void_p(h)
        self.last_refinements = 0

    # ------------------------------------------------------------ lifetime
    def destroy(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.mfh_destroy(h)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("context destroyed")
        return self._h

    def stream(self):
        return int(lib().mfh_stream(self.handle) or 0)

    def sync(self):
        check(lib().mfh_sync(self.handle), "mfh_sync")

    # -------------------------------------------------------------- ingest
    def set_rows(self, kind, row0, spectra):
        """Copy spectra into data (kind="data") or template rows from row0.

        ``spectra`` is a C-contiguous complex64 host array (rows, n), or a
        complex64 ROCm tensor of the same shape (copied device to device).
        """
        dev = _device_pointer(spectra)
        rows = int(spectra.shape[0])
        L = lib()
        if dev is not None:
            # the producer's work runs on PyTorch's stream: finish it first
            import torch
            torch.cuda.current_stream(spectra.device).synchronize()
            fn = L.mfh_set_data_device_async if kind == "data" else L.mfh_set_templates_device_async
            check(fn(self.handle, int(row0), rows, ctypes.c_void_p(dev)), "set %s" % kind)
            self.sync()
            return
        fn = L.mfh_set_data if kind == "data" else L.mfh_set_templates
        check(fn(self.handle, int(row0), rows, _ptr(spectra)), "set %s" % kind)

    # ---------------------------------------------------------------- runs
    def peaks(self, d0, nd, t0, nt, binsize, threshold, start, end, counts=False):
        """Flat filter over rows [d0, d0+nd) x [t0, t0+nt); host results.

        Returns ``(index int64, value complex64)`` of shape (nd, nt, nbins), and
        the per-pair crossing counts (int32, (nd, nt)) when ``counts``.
        """
        nb = int(lib().mfh_nbins(self.n, int(binsize), int(start), int(end)))
        idx = np.empty((nd, nt, nb), dtype=np.int64)
        val = np.empty((nd, nt, nb), dtype=np.complex64)
        cnt = np.empty((nd, nt), dtype=np.int32) if counts else None
        check(lib().mfh_flat_peaks(self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize),
                                   float(threshold), int(start), int(end), _ptr(idx), _ptr(val),
                                   _ptr(cnt) if cnt is not None else None), "mfh_flat_peaks")
        return idx, val, cnt

    def peaks_async(self, d0, nd, t0, nt, binsize, threshold, start, end, idx_ptr, val_ptr):
        """Queue a flat run writing device buffers (int64 / complex64 pointers)."""
        check(lib().mfh_flat_peaks_async(self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize),
                                         float(threshold), int(start), int(end),
                                         ctypes.c_void_p(int(idx_ptr)), ctypes.c_void_p(int(val_ptr))),
              "mfh_flat_peaks_async")

    # ---------------------------------------------------------- hierarchy
    def hier_set_band(self, band, ref_fraction=-1.0):
        """Coarse band and normalisation (bank-wide fraction, or < 0 per template)."""
        check(lib().mfh_hier_set_band(self.handle, int(band), float(ref_fraction)), "mfh_hier_set_band")

    def hier_peaks(self, d0, nd, t0, nt, binsize, threshold, start, end, coarse_threshold, counts=False):
        """Hierarchical run; returns (index, value, counts or None, refined pairs)."""
        nb = int(lib().mfh_nbins(self.n, int(binsize), int(start), int(end)))
        idx = np.empty((nd, nt, nb), dtype=np.int64)
        val = np.empty((nd, nt, nb), dtype=np.complex64)
        cnt = np.empty((nd, nt), dtype=np.int32) if counts else None
        refined = ctypes.c_int64(0)
        check(lib().mfh_hier_peaks(self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize),
                                   float(threshold), int(start), int(end), float(coarse_threshold),
                                   _ptr(idx), _ptr(val), _ptr(cnt) if cnt is not None else None,
                                   ctypes.byref(refined)), "mfh_hier_peaks")
        self.last_refinements = int(refined.value)
        return idx, val, cnt, int(refined.value)

    def hier_set_coarse_precision(self, precision, delta):
        """0 = FP32 screen, 1 = fp16 MFMA, 2 = bf16 MFMA with gate widening delta."""
        check(lib().mfh_hier_set_coarse_precision(self.handle, int(precision), float(delta)),
              "mfh_hier_set_coarse_precision")

    def hier_coarse_max(self, d0, nd, t0, nt, start, end, precision):
        """Coarse statistic max |z_c| per pair (nd, nt) at a precision."""
        out = np.empty((nd, nt), dtype=np.float32)
        check(lib().mfh_hier_coarse_max(self.handle, int(d0), int(nd), int(t0), int(nt), int(start),
                                        int(end), int(precision), _ptr(out)), "mfh_hier_coarse_max")
        return out

    def hier_peaks_async(self, d0, nd, t0, nt, binsize, threshold, start, end, coarse_threshold,
                         idx_ptr, val_ptr):
        """Queue a hierarchical run writing device buffers."""
        check(lib().mfh_hier_peaks_async(self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize),
                                         float(threshold), int(start), int(end), float(coarse_threshold),
                                         ctypes.c_void_p(int(idx_ptr)), ctypes.c_void_p(int(val_ptr)), None),
              "mfh_hier_peaks_async")

    # ------------------------------------------------------------ series
    def run_series(self, series, starts, win_start, win_end, t0, nt, binsize, threshold,
                   coarse_threshold=None):
        """Block spectra of a series on the device, then the flat (or, with a
        coarse threshold, hierarchical) filter; returns (index, value, refined)."""
        dev = _device_pointer(series)
        if dev is not None:
            import torch
            torch.cuda.current_stream(series.device).synchronize()
            sp, ns, on_dev = ctypes.c_void_p(dev), int(series.numel()), 1
        else:
            sp, ns, on_dev = _ptr(series), int(series.size), 0
        st = np.ascontiguousarray(starts, dtype=np.uintp)
        ws = np.ascontiguousarray(win_start, dtype=np.uintp)
        we = np.ascontiguousarray(win_end, dtype=np.uintp)
        nblk = int(st.size)
        nb = int(lib().mfh_nbins(self.n, int(binsize), int(ws[0]), int(we[0])))
        idx = np.empty((nblk, nt, nb), dtype=np.int64)
        val = np.empty((nblk, nt, nb), dtype=np.complex64)
        refined = ctypes.c_int64(0)
        if coarse_threshold is None:
            check(lib().mfh_run_series(self.handle, sp, ns, on_dev, _ptr(st), _ptr(ws), _ptr(we), nblk,
                                       int(t0), int(nt), int(binsize), float(threshold),
                                       _ptr(idx), _ptr(val), None), "mfh_run_series")
        else:
            check(lib().mfh_hier_run_series(self.handle, sp, ns, on_dev, _ptr(st), _ptr(ws), _ptr(we),
                                            nblk, int(t0), int(nt), int(binsize), float(threshold),
                                            float(coarse_threshold), _ptr(idx), _ptr(val), None,
                                            ctypes.byref(refined)), "mfh_hier_run_series")
        self.last_refinements = int(refined.value)
        return idx, val, int(refined.value)

    # ------------------------------------------------------- derivatives
    def peak_grad_async(self, data_ptr, nd, tmpl_ptr, nt, nb, idx_ptr, gval_ptr, gdata_ptr, gtmpl_ptr):
        """Queue dL/dD and/or dL/dT (device pointers; a null output is skipped)."""
        vp = lambda p: ctypes.c_void_p(int(p)) if p else None
        check(lib().mfh_peak_grad_async(self.handle, vp(data_ptr), self.n, int(nd), vp(tmpl_ptr), self.n,
                                        int(nt), int(nb), vp(idx_ptr), vp(gval_ptr), vp(gdata_ptr),
                                        vp(gtmpl_ptr)), "mfh_peak_grad_async")

    def peak_values_async(self, data_ptr, nd, tmpl_ptr, nt, nb, idx_ptr, out_ptr):
        """Queue z at stored lags (device pointers)."""
        vp = lambda p: ctypes.c_void_p(int(p))
        check(lib().mfh_peak_values_async(self.handle, vp(data_ptr), self.n, int(nd), vp(tmpl_ptr), self.n,
                                          int(nt), int(nb), vp(idx_ptr), vp(out_ptr)),
              "mfh_peak_values_async")

    def series_forward(self, series, starts, row0=0):
        """Block spectra of a series into data rows [row0, row0 + len(starts))."""
        dev = _device_pointer(series)
        if dev is not None:
            import torch
            torch.cuda.current_stream(series.device).synchronize()
            sp, ns, on_dev = ctypes.c_void_p(dev), int(series.numel()), 1
        else:
            sp, ns, on_dev = _ptr(series), int(series.size), 0
        st = np.ascontiguousarray(starts, dtype=np.uintp)
        check(lib().mfh_series_forward(self.handle, sp, ns, on_dev, _ptr(st), int(st.size), int(row0)),
              "mfh_series_forward")

    # ------------------------------------------------------------- tuning
    def set_tuning(self, key, value):
        check(lib().mfh_set_tuning(self.handle, key.encode(), int(value)), "mfh_set_tuning")

    def prof_enable(self, on=True):
        check(lib().mfh_prof_enable(self.handle, 1 if on else 0), "mfh_prof_enable")

    def prof_read(self, name):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_long(0)
        check(lib().mfh_prof_read(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)),
              "mfh_prof_read")
        return ms.value, cnt.value


class Queue:
    """Device side of :class:`apogee_amd.refine.RefineQueue`: a spectrum pool,
    a planned work queue and the persistent refine launch that drains it."""

    def __init__(self, index):
        L = lib()
        self.index = int(index)
        h = L.mfh_queue_create(self.index)
        if not h:
            raise RuntimeError("mfh_queue_create failed: %s" % last_error())
        self._h = ctypes.c_void_p(h)
        self._pool_ref = None

    def destroy(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.mfh_queue_destroy(h)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("queue destroyed")
        return self._h

    def set_pool(self, pool):
        """pool: 1-D complex64 host array (copied) or ROCm tensor (referenced)."""
        dev = _device_pointer(pool)
        if dev is not None:
            import torch
            torch.cuda.current_stream(pool.device).synchronize()
            check(lib().mfh_queue_set_pool(self.handle, ctypes.c_void_p(dev), int(pool.numel()), 1),
                  "mfh_queue_set_pool")
            self._pool_ref = pool     # the queue reads it in place
        else:
            check(lib().mfh_queue_set_pool(self.handle, _ptr(pool), int(pool.size), 0),
                  "mfh_queue_set_pool")
            self._pool_ref = None

    def plan(self, n, data_off, tmpl_off, ncand, binsize, threshold, ws, we):
        """Returns (per-request output offsets int64, total output elements)."""
        arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (n, data_off, tmpl_off)]
        nc = np.ascontiguousarray(ncand, dtype=np.int32)
        bs = np.ascontiguousarray(binsize, dtype=np.int64)
        th = np.ascontiguousarray(threshold, dtype=np.float32)
        w0 = np.ascontiguousarray(ws, dtype=np.int64)
        w1 = np.ascontiguousarray(we, dtype=np.int64)
        nreq = int(nc.size)
        off = np.empty(nreq, dtype=np.int64)
        total = lib().mfh_queue_plan(self.handle, nreq, _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                     _ptr(nc), _ptr(bs), _ptr(th), _ptr(w0), _ptr(w1), _ptr(off))
        check(total, "mfh_queue_plan")
        return off, int(total)

    def run(self, nout, npairs):
        """Host results: (index int64 [nout], value complex64 [nout], counts int32 [npairs])."""
        idx = np.empty(max(nout, 1), dtype=np.int64)
        val = np.empty(max(nout, 1), dtype=np.complex64)
        cnt = np.empty(max(npairs, 1), dtype=np.int32)
        check(lib().mfh_queue_run(self.handle, _ptr(idx), _ptr(val), _ptr(cnt)), "mfh_queue_run")
        return idx[:nout], val[:nout], cnt[:npairs]

    def run_async(self, idx_ptr, val_ptr):
        check(lib().mfh_queue_run_async(self.handle, ctypes.c_void_p(int(idx_ptr)),
                                        ctypes.c_void_p(int(val_ptr))), "mfh_queue_run_async")

    def sync(self):
        check(lib().mfh_queue_sync(self.handle), "mfh_queue_sync")

    def stream(self):
        return int(lib().mfh_queue_stream(self.handle) or 0)

    def set_tuning(self, key, value):
        check(lib().mfh_queue_set_tuning(self.handle, key.encode(), int(value)), "mfh_queue_set_tuning")

    def prof_enable(self, on=True):
        check(lib().mfh_queue_prof_enable(self.handle, 1 if on else 0), "mfh_queue_prof_enable")

    def prof_read(self):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_long(0)
        check(lib().mfh_queue_prof_read(self.handle, ctypes.byref(ms), ctypes.byref(cnt)),
              "mfh_queue_prof_read")
        return ms.value, cnt.value
