"""Differentiable peak values on the MI355X.

``PeakFilter(n, ndata, ntemplates)(data, templates, binsize, threshold,
window)`` takes complex64 ROCm tensors of spectra and returns
``(value, index)``: the flat filter's peak values (nd, nt, nbins) complex64,
differentiable with respect to both inputs, and the lags (int64, constant).
The forward is the fused flat kernel (only the peaks leave the chip); the
backward is a hand-written kernel that needs only the stored lags: with
g = dL/dRe z + i dL/dIm z,

    dL/dD_d[f] = sum_{t,b} g       T_t[f] exp(-2 pi i f k / n)
    dL/dT_t[f] = sum_{d,b} conj(g) D_d[f] exp(+2 pi i f k / n)

Bins that report nothing hold the constant 0 and pass no gradient.  The
argmax itself is piecewise constant, so these are the derivatives of the
loss on the piece where the reported lags do not change.

The reference filter (python/matchedfilter/__init__.py:435, the flat path of
src/balanced-inl.h:876-969) has no derivative; the forward values and lags
are exactly MatchedFilter.run's.
"""
import numpy as np
import torch

from . import _hip
from ._errors import UnsupportedSize


class _Peaks(torch.autograd.Function):

    @staticmethod
    def forward(ctx, data, templates, pf, binsize, threshold, start, end):
        value, index = pf._forward(data, templates, binsize, threshold, start, end)
        ctx.save_for_backward(data, templates, index)
        ctx.pf = pf
        ctx.mark_non_differentiable(index)
        return value, index

    @staticmethod
    def backward(ctx, gvalue, gindex):
        data, templates, index = ctx.saved_tensors
        need_d, need_t = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if gvalue is None or not (need_d or need_t):
            return None, None, None, None, None, None, None
        gd, gt = ctx.pf._backward(data, templates, index, gvalue, need_d, need_t)
        return gd, gt, None, None, None, None, None


class PeakFilter:
    """A flat filter of fixed shape whose peak values carry gradients."""

    def __init__(self, n, ndata=1, ntemplates=1, device=None):
        from .device import parse as _parse_device
        self.n, self.ndata, self.ntemplates = int(n), int(ndata), int(ntemplates)
        if self.ndata < 1 or self.ntemplates < 1:
            raise ValueError("ndata and ntemplates must be >= 1")
        self.device = _parse_device(device)
        if not _hip.supported(self.n):
            raise UnsupportedSize("transform length %d is not supported" % self.n)
        self._ctx = _hip.Context(self.device.index, self.n, self.ndata, self.ntemplates)
        self._torch_device = torch.device("cuda", self.device.index)

    def nbins(self, binsize=None, window=None):
        binsize, start, end = self._bins(binsize, window)
        return -(-(end - start) // binsize)

    def _bins(self, binsize, window):
        binsize = self.n if binsize is None else int(binsize)
        if binsize < 1:
            raise ValueError("binsize must be >= 1")
        if window is None:
            return binsize, 0, self.n
        start = max(0, min(int(window[0]), self.n))
        end = max(0, min(int(window[1]), self.n))
        if start >= end:
            raise ValueError("empty window (%d, %d)" % (start, end))
        return binsize, start, end

    def _check(self, x, rows, what):
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise TypeError("%s must be a ROCm tensor" % what)
        if x.dtype != torch.complex64:
            raise ValueError("%s must be complex64, got %s" % (what, x.dtype))
        if tuple(x.shape) != (rows, self.n):
            raise ValueError("%s: expected shape (%d, %d), got %s" % (what, rows, self.n, tuple(x.shape)))

    def __call__(self, data, templates, binsize=None, threshold=0.0, window=None):
        """Returns (value complex64 (nd, nt, nb), index int64 (nd, nt, nb))."""
        self._check(data, self.ndata, "data")
        self._check(templates, self.ntemplates, "templates")
        binsize, start, end = self._bins(binsize, window)
        return _Peaks.apply(data, templates, self, binsize, float(threshold), start, end)

    # ------------------------------------------------------------ kernels
    # Ordering against PyTorch is by host waits: PyTorch's stream is drained
    # before a launch reads its tensors, and the context's stream before the
    # outputs are handed back.  The banks are bound to the input tensors in
    # place (no copies).
    def _forward(self, data, templates, binsize, threshold, start, end):
        d = data.detach().contiguous()
        t = templates.detach().contiguous()
        nb = -(-(end - start) // binsize)
        shape = (self.ndata, self.ntemplates, nb)
        index = torch.empty(shape, dtype=torch.int64, device=self._torch_device)
        value = torch.empty(shape, dtype=torch.complex64, device=self._torch_device)
        torch.cuda.current_stream(self._torch_device).synchronize()
        self._ctx.bind_tensor("data", d)
        self._ctx.bind_tensor("template", t)
        self._ctx.peaks_async(0, self.ndata, 0, self.ntemplates, binsize, threshold, start, end,
                              index.data_ptr(), value.data_ptr())
        self._ctx.sync()
        return value, index

    def _backward(self, data, templates, index, gvalue, need_d, need_t):
        d = data.detach().contiguous()
        t = templates.detach().contiguous()
        g = gvalue.to(torch.complex64).contiguous()
        gd = torch.empty_like(d) if need_d else None
        gt = torch.empty_like(t) if need_t else None
        torch.cuda.current_stream(self._torch_device).synchronize()
        self._ctx.peak_grad_async(d.data_ptr(), self.ndata, t.data_ptr(), self.ntemplates, index.shape[2],
                                  index.data_ptr(), g.data_ptr(),
                                  gd.data_ptr() if gd is not None else 0,
                                  gt.data_ptr() if gt is not None else 0)
        self._ctx.sync()
        return gd, gt

    def recompute(self, data, templates, index):
        """z at the stored lags (a direct sum per peak), without gradients."""
        self._check(data, self.ndata, "data")
        self._check(templates, self.ntemplates, "templates")
        d = data.detach().contiguous()
        t = templates.detach().contiguous()
        idx = index.contiguous()
        out = torch.empty(idx.shape, dtype=torch.complex64, device=self._torch_device)
        torch.cuda.current_stream(self._torch_device).synchronize()
        self._ctx.peak_values_async(d.data_ptr(), self.ndata, t.data_ptr(), self.ntemplates, idx.shape[2],
                                    idx.data_ptr(), out.data_ptr())
        self._ctx.sync()
        return out


def peak_values(data, templates, binsize=None, threshold=0.0, window=None):
    """Functional form: differentiable (value, index) of data x templates."""
    key = (data.device.index or 0, int(data.shape[-1]), int(data.shape[0]), int(templates.shape[0]))
    pf = _CACHE.get(key)
    if pf is None:
        pf = PeakFilter(key[1], key[2], key[3], device="gpu:%d" % key[0])
        _CACHE[key] = pf
    return pf(data, templates, binsize=binsize, threshold=threshold, window=window)


_CACHE = {}
