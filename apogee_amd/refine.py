"""Uneven batched refine: many small matched-filter requests of mixed
transform lengths in one persistent device launch.

A request is one data spectrum and ``c`` template spectra of one length ``n``
(64 ... 65536) with its own binsize, threshold and lag window; its result is
what ``MatchedFilter(n, 1, c).run(binsize, threshold, window)`` returns for
those spectra, index for index and value for value.  Requests are staged with
:meth:`RefineQueue.add` (host spectra, copied into one pool) or described
over a caller-owned pool with :meth:`RefineQueue.plan` (a 1-D complex64 host
array or ROCm tensor, read in place), then drained by :meth:`RefineQueue.run`.
The second identical run is captured into a HIP graph and replayed after.

The reference refines survivors of its coarse gate one length at a time
(python/matchedfilter/_vkcompute.py:534-850, ``refineListed`` in
src/gpu/tierb.slang:986-1034) and otherwise runs one ``ap_mf_run`` per request
(src/matchfilt.c:349-420); the mixed-length queue has no counterpart there.
"""
import numpy as np

from . import _hip
from ._errors import UnsupportedSize

_LENGTHS = frozenset(1 << k for k in range(6, 21))


def _peak_dtype():
    from . import PEAK_DTYPE
    return PEAK_DTYPE


class RefineQueue:
    """Mixed-length (data, templates) requests drained by one persistent launch."""

    #: Builds the device queue; tests substitute a host double here.
    _queue_factory = _hip.Queue

    def __init__(self, device=None):
        from .device import parse as _parse_device
        self.device = _parse_device(device)
        self._q = self._queue_factory(self.device.index)
        from . import autotune
        autotune.apply(self._q, "queue")
        self._staged = []          # add(): (data, templates, binsize, threshold, window)
        self._planned = None       # arrays of the current plan
        self._dirty = True

    # ----------------------------------------------------------- staging
    def add(self, data, templates, binsize=None, threshold=0.0, window=None):
        """Stage one request; returns its position in the results of run().

        ``data`` is one spectrum (n,), ``templates`` is (c, n) or (n,); both
        complex64-convertible host arrays.  ``binsize``, ``threshold`` and
        ``window`` mean what they mean for :meth:`MatchedFilter.run`.
        """
        d = np.ascontiguousarray(data, dtype=np.complex64)
        t = np.ascontiguousarray(templates, dtype=np.complex64)
        if d.ndim != 1:
            raise ValueError("data must be one spectrum (1-D), got shape %s" % (d.shape,))
        if t.ndim == 1:
            t = t[None, :]
        if t.ndim != 2 or t.shape[1] != d.size:
            raise ValueError("templates must be (c, %d), got shape %s" % (d.size, t.shape))
        n = int(d.size)
        self._check_length(n)
        bs, ws, we = self._bins(n, binsize, window)
        thr = float(threshold)
        self._staged.append((d, t, bs, thr, ws, we))
        self._planned = None
        self._dirty = True
        return len(self._staged) - 1

    def clear(self):
        """Forget every staged request and plan."""
        self._staged = []
        self._planned = None
        self._dirty = True

    def __len__(self):
        if self._planned is not None:
            return int(self._planned["n"].size)
        return len(self._staged)

    # ------------------------------------------------------ explicit plan
    def plan(self, pool, n, data_offset, template_offset, candidates,
             binsize=None, threshold=0.0, window=None):
        """Describe requests over a caller-owned spectrum pool.

        ``pool`` is a 1-D complex64 array (host: copied once; ROCm tensor:
        read in place, keep it alive and unchanged in shape).  Per request r
        (array arguments, scalars broadcast): length ``n[r]``, data spectrum
        at ``pool[data_offset[r]:][:n[r]]``, ``candidates[r]`` templates at
        ``template_offset[r] + j*n[r]``.  ``binsize`` defaults to n[r];
        ``window`` is None, one (start, end) pair, or an (nreq, 2) array.
        """
        from . import _host_view, _on_rocm
        if _on_rocm(pool):
            import torch
            p = torch.from_dlpack(pool) if not isinstance(pool, torch.Tensor) else pool
            if p.dtype != torch.complex64 or p.dim() != 1:
                raise ValueError("pool must be a 1-D complex64 tensor")
            p = p.contiguous()
            size = int(p.numel())
        else:
            p = np.ascontiguousarray(_host_view(pool), dtype=np.complex64)
            if p.ndim != 1:
                raise ValueError("pool must be 1-D, got shape %s" % (p.shape,))
            size = int(p.size)
        n = np.atleast_1d(np.asarray(n, dtype=np.int64))
        nreq = int(n.size)
        for v in np.unique(n):
            self._check_length(int(v))

        def per_request(x, dtype, what):
            a = np.asarray(x, dtype=dtype)
            if a.ndim == 0:
                a = np.full(nreq, a, dtype=dtype)
            if a.shape != (nreq,):
                raise ValueError("%s must be a scalar or have %d entries" % (what, nreq))
            return a

        doff = per_request(data_offset, np.int64, "data_offset")
        toff = per_request(template_offset, np.int64, "template_offset")
        cand = per_request(candidates, np.int64, "candidates")
        if (cand < 0).any():
            raise ValueError("candidates must be >= 0")
        if (doff < 0).any() or (doff + n > size).any():
            raise ValueError("a data spectrum lies outside the pool")
        if (toff < 0).any() or (toff + cand * n > size).any():
            raise ValueError("template spectra lie outside the pool")
        bs = n.copy() if binsize is None else per_request(binsize, np.int64, "binsize")
        if (bs < 1).any():
            raise ValueError("binsize must be >= 1")
        thr = per_request(threshold, np.float64, "threshold")
        if window is None:
            ws, we = np.zeros(nreq, np.int64), n.copy()
        else:
            w = np.asarray(window, dtype=np.int64)
            if w.shape == (2,):
                w = np.broadcast_to(w, (nreq, 2))
            if w.shape != (nreq, 2):
                raise ValueError("window must be (start, end) or an (nreq, 2) array")
            ws = np.clip(w[:, 0], 0, n)
            we = np.clip(w[:, 1], 0, n)
            if (ws >= we).any():
                r = int(np.argmax(ws >= we))
                raise ValueError("empty window (%d, %d) for request %d" % (ws[r], we[r], r))
        self._q.set_pool(p)
        self._set_plan(n, doff, toff, cand, bs, thr.astype(np.float32), ws, we)
        self._staged = []

    def _set_plan(self, n, doff, toff, cand, bs, thr, ws, we):
        off, total = self._q.plan(n, doff, toff, cand, bs, thr, ws, we)
        nb = (we - ws + bs - 1) // bs
        self._planned = {"n": n, "cand": cand, "nb": nb, "off": off, "total": total}
        self._dirty = False

    def _stage_plan(self):
        if not self._staged:
            raise ValueError("no requests: call add() or plan() first")
        sizes = [d.size + t.size for d, t, *_ in self._staged]
        pool = np.empty(int(sum(sizes)), dtype=np.complex64)
        nreq = len(self._staged)
        n = np.empty(nreq, np.int64)
        doff = np.empty(nreq, np.int64)
        toff = np.empty(nreq, np.int64)
        cand = np.empty(nreq, np.int64)
        bs = np.empty(nreq, np.int64)
        thr = np.empty(nreq, np.float32)
        ws = np.empty(nreq, np.int64)
        we = np.empty(nreq, np.int64)
        pos = 0
        for r, (d, t, b, th, w0, w1) in enumerate(self._staged):
            n[r] = d.size
            doff[r] = pos
            pool[pos:pos + d.size] = d
            pos += d.size
            toff[r] = pos
            pool[pos:pos + t.size] = t.reshape(-1)
            pos += t.size
            cand[r] = t.shape[0]
            bs[r], thr[r], ws[r], we[r] = b, th, w0, w1
        self._q.set_pool(pool)
        self._set_plan(n, doff, toff, cand, bs, thr, ws, we)

    # --------------------------------------------------------------- run
    def run(self, raw=False):
        """Drain the queue; returns one (c, nbins) result per request.

        Each is a PEAK_DTYPE array (``raw=False``) or an ``(index, value)``
        pair of views into one flat buffer, in request order.
        """
        if self._dirty or self._planned is None:
            self._stage_plan()
        pl = self._planned
        idx, val, _ = self._q.run(pl["total"], int(pl["cand"].sum()))
        out = []
        for r in range(pl["n"].size):
            c, nb, o = int(pl["cand"][r]), int(pl["nb"][r]), int(pl["off"][r])
            i = idx[o:o + c * nb].reshape(c, nb)
            v = val[o:o + c * nb].reshape(c, nb)
            if raw:
                out.append((i, v))
            else:
                rec = np.empty((c, nb), _peak_dtype())
                rec["index"], rec["value"] = i, v
                out.append(rec)
        return out

    def run_async(self, index_out, value_out):
        """Queue a run writing device tensors (int64 and complex64, at least
        :attr:`total` elements each) on the queue's stream; see :meth:`sync`."""
        if self._dirty or self._planned is None:
            self._stage_plan()
        need = self._planned["total"]
        for t, what in ((index_out, "index_out"), (value_out, "value_out")):
            if int(t.numel()) < need:
                raise ValueError("%s holds %d elements, the plan writes %d" % (what, t.numel(), need))
        self._q.run_async(index_out.data_ptr(), value_out.data_ptr())

    def sync(self):
        self._q.sync()

    # ------------------------------------------------------ introspection
    @property
    def total(self):
        """Output elements of the plan (sum over requests of c * nbins)."""
        if self._dirty or self._planned is None:
            self._stage_plan()
        return self._planned["total"]

    @property
    def offsets(self):
        """Element offset of each request's [c][nbins] block in run_async outputs."""
        if self._dirty or self._planned is None:
            self._stage_plan()
        return self._planned["off"].copy()

    def set_tuning(self, key, value):
        """"grid_per_cu" (persistent workgroups per CU) or "graphs" (0/1)."""
        self._q.set_tuning(key, value)

    # ----------------------------------------------------------- helpers
    @staticmethod
    def _check_length(n):
        if n not in _LENGTHS:
            if 64 <= n <= 1048576 and n & (n - 1) == 0:
                raise UnsupportedSize("transform length %d is not supported" % n)
            raise ValueError("transform length must be a power of two in [64, 1048576], got %d" % n)

    @staticmethod
    def _bins(n, binsize, window):
        bs = n if binsize is None else int(binsize)
        if bs < 1:
            raise ValueError("binsize must be >= 1")
        if window is None:
            ws, we = 0, n
        else:
            ws = max(0, min(int(window[0]), n))
            we = max(0, min(int(window[1]), n))
            if ws >= we:
                raise ValueError("empty window (%d, %d)" % (ws, we))
        return bs, ws, we
