"""The reference-side binding (tests/hip_context.py: the
reference's Context protocol of python/matchedfilter/_vkcompute.py over
libmfhip.so through ctypes) run against the FP64 oracle and this package's
own classes.  Covers ``peaks`` (int32 index), ``hier_peaks`` with
caller-scaled coarse templates and a raw gate exactly as the reference's
HierarchicalFilter._gpu_dispatch prepares them
(python/matchedfilter/__init__.py:1284-1303), and ``forward``.
"""
import numpy as np
import pytest

import apogee_amd as mf
from apogee_amd import _hip
from oracle import mf_oracle as O
from _parity import assert_peaks_match, tol
from hip_context import Context

pytestmark = pytest.mark.gpu


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.fixture(scope="module")
def ctx(gpu_device):
    _hip.lib()          # one HIP runtime with PyTorch's in this process (see _hip)
    c = Context(gpu_device, _hip.LIB_PATH)
    yield c
    c.destroy()


def reference_ct0(H, band, ref=None):
    """ct0 as python/matchedfilter/__init__.py:1289-1302 computes it."""
    if ref is None:
        power = np.abs(H.astype(np.complex128)) ** 2
        total = power.sum(axis=1)
        fraction = np.divide(power[:, :band].sum(axis=1), total, out=np.zeros_like(total), where=total > 0)
        sc = np.divide(1.0, np.sqrt(fraction), out=np.zeros_like(fraction), where=fraction > 0)[:, None]
    else:
        r = np.asarray(ref, dtype=np.float64)
        f = float(r[:band].sum() / r.sum())
        sc = 1.0 / np.sqrt(f) if f > 0 else 0.0
    return (H[:, :band] * sc).astype(np.complex64)


@pytest.mark.parametrize("n,binsize,window,threshold", [(4096, 512, None, 0.0), (1024, 100, (37, 1000), 40.0),
                                                       (65536, 4096, (1000, 60000), 0.0)])
def test_peaks_protocol(ctx, n, binsize, window, threshold):
    rng = np.random.default_rng(n + binsize)
    nd, nt = 3, 7
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    D[1] += 2 * T[4]
    idx, val = ctx.peaks(n, D, T, binsize=binsize, threshold=threshold, window=window)
    assert idx.dtype == np.int32 and val.dtype == np.complex64
    assert_peaks_match(D, T, idx.astype(np.int64), val, binsize, threshold, window, context="shim peaks")
    # resident spectra are reused when the caller says nothing changed
    i2, v2 = ctx.peaks(n, np.zeros_like(D), np.zeros_like(T), binsize=binsize, threshold=threshold,
                       window=window, upload_data=False, upload_tmpl=False)
    np.testing.assert_array_equal(i2, idx)
    np.testing.assert_array_equal(v2, val)
    with pytest.raises(ValueError):
        ctx.peaks(n, D, T, window=(n, n + 5))


@pytest.mark.parametrize("use_ref", [False, True])
def test_hier_peaks_protocol_with_caller_coarse_templates(ctx, use_ref):
    n, nd, nt, band = 4096, 8, 64, 512
    rng = np.random.default_rng(31 + use_ref)
    k = np.arange(1, n // 2, dtype=np.float64)
    ref = np.zeros(n, np.float32)
    ref[1:n // 2] = k ** (-7.0 / 3.0) / ((0.015 * n / k) ** 4 + 1.0)
    H = (np.sqrt(ref.astype(np.float64))[None, :] * np.exp(2j * np.pi * rng.random((nt, n))))
    H = (H / np.linalg.norm(H, axis=1, keepdims=True)).astype(np.complex64)
    D = cnoise(rng, (nd, n))
    for d in range(0, nd, 2):
        D[d] += (7.0 * H[3 * d + 1] * np.exp(2j * np.pi * np.arange(n) * (200 + 91 * d) / n)).astype(np.complex64)
    ct0 = reference_ct0(H, band, ref if use_ref else None)
    raw_thr = 4.5
    idx, val = ctx.hier_peaks(n, band, D, H, ct0, raw_thr, binsize=1024, threshold=5.0)
    assert idx.dtype == np.int32
    # the FP64 oracle with the same coarse bank and gate
    zc = O.correlate(D[:, None, :band], ct0.astype(np.complex128)[None])
    cmax = np.abs(zc).max(axis=-1)
    ridx, rval, _, fired = O.hier_peaks(D, H, band, raw_thr, 1024, 5.0, coarse_max=cmax)
    edge = np.abs(cmax - raw_thr) <= 4 * tol(band) * raw_thr
    assert 0 < fired.sum() < nd * nt
    assert int((fired & ~edge).sum()) <= ctx.last_refinements <= int((fired | edge).sum())
    sure = ~edge
    # wherever the gate decision is unambiguous, the result is the oracle's
    # hierarchical result: the flat peaks of fired pairs, -1 / 0 elsewhere
    dismissed = sure & ~fired
    assert (idx[dismissed] == -1).all() and (val[dismissed] == 0).all()
    keep = sure & fired
    flat_i, flat_v = ctx.peaks(n, D, H, binsize=1024, threshold=5.0)
    assert_peaks_match(D, H, flat_i.astype(np.int64), flat_v, 1024, 5.0, context="shim flat")
    np.testing.assert_array_equal(idx[keep], flat_i[keep])
    np.testing.assert_array_equal(val[keep], flat_v[keep])
    # the package's HierarchicalFilter with the same band and gate agrees bit for bit
    h = mf.HierarchicalFilter(n, nd, nt, band=band, device="gpu")
    if use_ref:
        h.set_reference(ref)
    h.set_coarse_threshold(raw_thr)
    h.set_data(D)
    h.set_templates(H)
    pk = h.run(binsize=1024, threshold=5.0)
    np.testing.assert_array_equal(pk["index"][sure], idx[sure])
    np.testing.assert_array_equal(pk["value"][sure], val[sure])


def test_forward_protocol(ctx):
    n, nblk = 4096, 6
    rng = np.random.default_rng(44)
    series = cnoise(rng, nblk * n // 2 + 700)
    starts = (np.arange(nblk) * (n // 2)).astype(np.uint32)
    spectra = ctx.empty_shared((nblk, n))
    ctx.forward(n, series, starts, spectra, defer=True)
    ctx.cancel_forward()
    exact = O.series_spectra(series, starts, n)
    err = np.abs(spectra - exact).max() / np.abs(exact).max()
    assert err < 4 * tol(n), err
    # the spectra feed peaks exactly as run_series filters its blocks
    T = cnoise(rng, (5, n))
    idx, val = ctx.peaks(n, spectra, T, binsize=512)
    f = mf.MatchedFilter(n, 1, 5, device="gpu")
    f.set_templates(T)
    rs = f.run_series(series, starts, np.zeros(nblk, np.uint64), np.full(nblk, n, np.uint64), binsize=512)
    np.testing.assert_array_equal(idx, rs["index"])
    np.testing.assert_array_equal(val, rs["value"])
