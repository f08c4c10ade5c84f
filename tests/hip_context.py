"""The reference's backend Context protocol over libmfhip.so, as a
maintainer would add it to the reference package
(python/matchedfilter/_hipcompute.py beside _vkcompute.py).

Only ctypes and NumPy: no PyTorch, nothing from apogee_amd.  Signatures,
defaults, return types (int32 index, complex64 value), errors and the
``last_refinements`` attribute follow python/matchedfilter/_vkcompute.py:

* ``peaks`` (:962-1063), ``hier_peaks`` (:534-623), ``forward`` and
  ``cancel_forward`` (:855-906);
* ``empty_shared``, ``clear_cache``, ``destroy``, ``max_dispatch_x``,
  ``cache_limit_bytes``.

One mfh context per (n, ndata, ntemplates) shape is kept, as the reference
caches one dispatch batch per shape; ``upload_data`` / ``upload_tmpl`` False
reuse the spectra resident from the previous call on that shape.  Test
infrastructure: tests/test_gpu_context_shim.py runs it against the FP64
oracle and the package's own classes.
"""
import ctypes

import numpy as np

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


class HipError(RuntimeError):
    pass


def _ptr(a):
    return _vp(a.ctypes.data)


class Context:
    """GPU context on HIP device ``index`` (the Vulkan Context's contract)."""

    max_dispatch_x = 1 << 31
    cache_limit_bytes = 8 << 30

    def __init__(self, index=0, lib_path=None, coarse_precision=0, coarse_delta=0.0):
        self.lib = L = ctypes.CDLL(lib_path)
        L.mfh_last_error.restype = ctypes.c_char_p
        L.mfh_create.restype = _vp
        L.mfh_create.argtypes = [ctypes.c_int, _sz, ctypes.c_int, ctypes.c_int]
        L.mfh_destroy.argtypes = [_vp]
        L.mfh_destroy.restype = None
        for name in ("mfh_set_data", "mfh_set_templates"):
            getattr(L, name).argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
        L.mfh_nbins.restype = _sz
        L.mfh_nbins.argtypes = [_sz, _sz, _sz, _sz]
        L.mfh_flat_peaks.restype = ctypes.c_long
        L.mfh_flat_peaks.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz,
                                     ctypes.c_float, _sz, _sz, _vp, _vp, _vp]
        L.mfh_hier_set_coarse_templates.argtypes = [_vp, _sz, _vp, ctypes.c_int]
        L.mfh_hier_set_coarse_precision.argtypes = [_vp, ctypes.c_int, ctypes.c_float]
        L.mfh_hier_peaks.restype = ctypes.c_long
        L.mfh_hier_peaks.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz,
                                     ctypes.c_float, _sz, _sz, ctypes.c_float, _vp, _vp, _vp,
                                     ctypes.POINTER(ctypes.c_int64)]
        L.mfh_series_forward.argtypes = [_vp, _vp, _sz, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]
        L.mfh_get_rows.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
        self.index = int(index)
        self.coarse = (int(coarse_precision), float(coarse_delta))
        self.last_refinements = 0
        self._ctx = {}
        self._band = {}

    # ----------------------------------------------------------- plumbing
    def _check(self, status, what):
        if status < 0:
            raise HipError("%s: %s" % (what, self.lib.mfh_last_error().decode()))
        return status

    def _handle(self, n, nd, nt):
        """(handle, fresh): the context of this shape, created on first use."""
        key = (n, nd, nt)
        h = self._ctx.get(key)
        if h is not None:
            return h, False
        h = self.lib.mfh_create(self.index, n, nd, nt)
        if not h:
            raise HipError("mfh_create: %s" % self.lib.mfh_last_error().decode())
        h = _vp(h)
        self._check(self.lib.mfh_hier_set_coarse_precision(h, *self.coarse), "coarse precision")
        self._ctx[key] = h
        return h, True

    @staticmethod
    def _window(n, window, binsize):
        lo, hi = (0, n) if window is None else (int(window[0]), int(window[1]))
        lo, hi = max(0, min(lo, n)), max(0, min(hi, n))
        if lo >= hi:
            raise ValueError("empty window (%d, %d)" % (lo, hi))
        binsize = n if binsize is None else int(binsize)
        return lo, hi, binsize, -(-(hi - lo) // binsize)

    def _upload(self, h, data, tmpl, upload_data, upload_tmpl, fresh):
        if upload_data or fresh:
            d = np.ascontiguousarray(data, dtype=np.complex64)
            self._check(self.lib.mfh_set_data(h, 0, d.shape[0], _ptr(d)), "upload data")
        if upload_tmpl or fresh:
            t = np.ascontiguousarray(tmpl, dtype=np.complex64)
            self._check(self.lib.mfh_set_templates(h, 0, t.shape[0], _ptr(t)), "upload templates")

    # ------------------------------------------------------------- protocol
    def peaks(self, n, data, tmpl, binsize=None, threshold=0.0, window=None,
              upload_data=True, upload_tmpl=True):
        """Flat filter of every (data, template) pair: (index int32, value
        complex64), each (nd, nt, nbins)."""
        nd, nt = data.shape[0], tmpl.shape[0]
        lo, hi, binsize, nb = self._window(n, window, binsize)
        h, fresh = self._handle(n, nd, nt)
        self._upload(h, data, tmpl, upload_data, upload_tmpl, fresh)
        idx = np.empty((nd, nt, nb), np.int64)
        val = np.empty((nd, nt, nb), np.complex64)
        self._check(self.lib.mfh_flat_peaks(h, 0, nd, 0, nt, binsize, float(threshold), lo, hi,
                                            _ptr(idx), _ptr(val), None), "peaks")
        return idx.astype(np.int32), val

    def hier_peaks(self, n, band, data, tmpl, ct0, raw_thr,
                   binsize=None, threshold=0.0, window=None,
                   upload_data=True, upload_tmpl=True):
        """Coarse screen with the caller's coarse templates ct0 (nt, band) and
        gate raw_thr, then the exact filter on the survivors; sets
        ``last_refinements`` to the number of refined pairs."""
        nd, nt = data.shape[0], tmpl.shape[0]
        lo, hi, binsize, nb = self._window(n, window, binsize)
        h, fresh = self._handle(n, nd, nt)
        self._upload(h, data, tmpl, upload_data, upload_tmpl, fresh)
        if upload_tmpl or fresh or self._band.get((n, nd, nt)) != band:
            c = np.ascontiguousarray(ct0, dtype=np.complex64)
            if c.shape != (nt, band):
                raise ValueError("ct0 must be (%d, %d), got %s" % (nt, band, c.shape))
            self._check(self.lib.mfh_hier_set_coarse_templates(h, band, _ptr(c), 0), "coarse templates")
            self._band[(n, nd, nt)] = band
        idx = np.empty((nd, nt, nb), np.int64)
        val = np.empty((nd, nt, nb), np.complex64)
        refined = ctypes.c_int64(0)
        self._check(self.lib.mfh_hier_peaks(h, 0, nd, 0, nt, binsize, float(threshold), lo, hi,
                                            float(raw_thr), _ptr(idx), _ptr(val), None,
                                            ctypes.byref(refined)), "hier_peaks")
        self.last_refinements = int(refined.value)
        return idx.astype(np.int32), val

    def forward(self, n, series, starts, spectra, *, defer=False):
        """Block spectra FFT(series[s:s+n], zero-padded)/n into ``spectra``
        (nblocks, n) for every start.  ``defer`` is accepted for the
        protocol; the work runs at once."""
        ser = np.ascontiguousarray(series, dtype=np.complex64)
        st = np.ascontiguousarray(starts, dtype=np.uintp)
        nblk = int(st.size)
        if spectra.dtype != np.complex64 or spectra.shape != (nblk, n) or not spectra.flags.c_contiguous:
            raise ValueError("spectra must be C-contiguous complex64 (%d, %d)" % (nblk, n))
        h, _ = self._handle(n, nblk, 1)
        self._check(self.lib.mfh_series_forward(h, _ptr(ser), ser.size, 0, _ptr(st), nblk, 0), "forward")
        self._check(self.lib.mfh_get_rows(h, 0, 0, nblk, _ptr(spectra)), "forward readback")

    def cancel_forward(self):
        pass

    def empty_shared(self, shape, dtype=np.complex64):
        return np.empty(shape, dtype=dtype)

    def clear_cache(self):
        for h in self._ctx.values():
            self.lib.mfh_destroy(h)
        self._ctx.clear()
        self._band.clear()

    def destroy(self):
        self.clear_cache()
