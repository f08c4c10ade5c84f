"""mfh_hier_peaks_noalloc against mfh_hier_peaks_async, through the C ABI.

The allocation-free entry must give exactly the existing entry's results
(same kernels; its per-call buffers live in the library's fixed device
array), on two calls in a row, the second with a smaller input, for the FP32
and the matrix-core screens; it must not change the device's free memory,
must refuse more pairs than its array holds, and must refuse a context whose
own buffers are not built yet, launching nothing.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


def lib():
    from apogee_amd import _hip
    L = _hip.lib()
    L.mfh_hier_peaks_noalloc.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz,
                                         ctypes.c_float, _sz, _sz, ctypes.c_float, _vp, _vp]
    L.mfh_hier_peaks_async.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz,
                                       ctypes.c_float, _sz, _sz, ctypes.c_float, _vp, _vp, _vp]
    return L


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def make_ctx(L, n, nd, nt, band, prec, seed):
    rng = np.random.default_rng(seed)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    for d in range(nd):                       # loud pairs among noise pairs
        D[d] += (9.0 * T[(11 * d) % nt] * np.exp(2j * np.pi * np.arange(n) * (50 + 97 * d) / n)).astype(np.complex64)
    h = L.mfh_create(0, n, nd, nt)
    assert h, L.mfh_last_error()
    h = _vp(h)
    assert L.mfh_set_data(h, 0, nd, D.ctypes.data) == 0
    assert L.mfh_set_templates(h, 0, nt, T.ctypes.data) == 0
    assert L.mfh_hier_set_band(h, band, -1.0) == 0
    assert L.mfh_hier_set_coarse_precision(h, prec, 2e-3 if prec == 1 else (1.6e-2 if prec == 2 else 0.0)) == 0
    return h


def run_pair(L, h, d0, nd, t0, nt, binsize, thr, ws, we, cthr):
    """The allocating entry, then the allocation-free one, on the same
    request; the second must leave the device's free memory unchanged."""
    import torch
    nb = -(-(we - ws) // binsize)
    out = []
    for fn in (L.mfh_hier_peaks_async, L.mfh_hier_peaks_noalloc):
        idx = torch.full((nd, nt, nb), -7, dtype=torch.int64, device="cuda")
        val = torch.full((nd, nt, nb), 7.0, dtype=torch.complex64, device="cuda")
        torch.cuda.synchronize()
        args = [h, d0, nd, t0, nt, binsize, thr, ws, we, cthr, _vp(idx.data_ptr()), _vp(val.data_ptr())]
        if fn is L.mfh_hier_peaks_async:
            args.append(None)
        free0 = torch.cuda.mem_get_info()[0]
        rc = fn(*args)
        assert rc == 0, L.mfh_last_error()
        assert L.mfh_sync(h) == 0
        if fn is L.mfh_hier_peaks_noalloc:
            assert torch.cuda.mem_get_info()[0] == free0
        out.append((idx.cpu().numpy(), val.cpu().numpy()))
    return out


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["fp32", "fp16", "bf16"])
def test_noalloc_matches_the_allocating_entry(prec):
    L = lib()
    n, nd, nt, band = 4096, 16, 256, 512
    h = make_ctx(L, n, nd, nt, band, prec, seed=900 + prec)
    try:
        # the allocating entry builds the context's own buffers on its first call
        first = run_pair(L, h, 0, nd, 0, nt, 1024, 5.0, 0, n, 4.0)
        calls = [(0, nd, 0, nt, 1024, 5.0, 0, n, 4.0),            # the full input
                 (3, 5, 17, 77, 700, 4.5, 100, 3900, 3.5)]       # then a smaller one
        for args in calls:
            (ai, av), (bi, bv) = run_pair(L, h, *args)
            np.testing.assert_array_equal(bi, ai, err_msg=str(args))
            np.testing.assert_array_equal(bv, av, err_msg=str(args))
            assert (ai >= 0).any() and (ai < 0).any(), "both fired and dismissed pairs are exercised"
        np.testing.assert_array_equal(first[1][0], first[0][0])
    finally:
        L.mfh_destroy(h)


def test_noalloc_limits_and_unbuilt_context():
    import torch
    L = lib()
    n = 1024
    # more pairs than the fixed array holds: rejected before any work
    h = make_ctx(L, n, 600, 512, 128, 0, seed=5)
    try:
        run_pair(L, h, 0, 4, 0, 8, 256, 5.0, 0, n, 4.0)        # builds the context's buffers
        idx = torch.empty((600, 512, 1), dtype=torch.int64, device="cuda")
        val = torch.empty((600, 512, 1), dtype=torch.complex64, device="cuda")
        rc = L.mfh_hier_peaks_noalloc(h, 0, 600, 0, 512, n, 5.0, 0, n, 4.0, _vp(idx.data_ptr()),
                                      _vp(val.data_ptr()))
        assert rc < 0 and b"allocation-free limit" in L.mfh_last_error()
    finally:
        L.mfh_destroy(h)
    # a fresh context whose coarse bank was never built: refused, nothing launched
    h = make_ctx(L, n, 2, 3, 128, 0, seed=6)
    try:
        idx = torch.full((2, 3, 1), -7, dtype=torch.int64, device="cuda")
        val = torch.zeros((2, 3, 1), dtype=torch.complex64, device="cuda")
        torch.cuda.synchronize()
        rc = L.mfh_hier_peaks_noalloc(h, 0, 2, 0, 3, n, 5.0, 0, n, 4.0, _vp(idx.data_ptr()),
                                      _vp(val.data_ptr()))
        assert rc < 0 and b"not built yet" in L.mfh_last_error()
        assert L.mfh_sync(h) == 0
        assert (idx.cpu().numpy() == -7).all()
        # after one allocating call the same request succeeds
        (ai, av), (bi, bv) = run_pair(L, h, 0, 2, 0, 3, n, 5.0, 0, n, 4.0)
        np.testing.assert_array_equal(bi, ai)
        np.testing.assert_array_equal(bv, av)
    finally:
        L.mfh_destroy(h)
