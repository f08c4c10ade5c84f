"""Zero-copy ingest of ROCm arrays through DLPack.

A whole bank handed over as a ROCm DLPack producer is read in place (the
context's bank points at the producer's memory; nothing is copied), ordered
after the producer's pending work on the device by the DLPack stream
hand-over.  Results must equal the host-ingest results bit for bit, the
library must never write the producer's buffer, and a later update of the
buffer takes effect when the bank is set again.  The reference refuses
device arrays (python/matchedfilter/__init__.py:103-131); here they are used
in place.
"""
import numpy as np
import pytest

import apogee_amd as mf

pytestmark = pytest.mark.gpu


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


class Producer:
    """A DLPack producer that is not a PyTorch tensor: it forwards the
    protocol of a tensor it wraps, as any other ROCm array library would."""

    def __init__(self, t):
        self._t = t

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()

    def __dlpack__(self, stream=None, **kw):
        return self._t.__dlpack__(stream=stream, **kw)


def host_run(D, T, **kw):
    f = mf.MatchedFilter(D.shape[1], D.shape[0], T.shape[0], device="gpu")
    f.set_data(D)
    f.set_templates(T)
    return f.run(**kw)


@pytest.mark.parametrize("n", [256, 4096, 65536])
def test_bound_banks_equal_host_ingest(n, gpu_device):
    import torch
    rng = np.random.default_rng(400 + n)
    nd, nt = 3, 5
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    ref = host_run(D, T, binsize=n // 4, threshold=0.0)
    dt = torch.from_numpy(D).to("cuda")
    tt = torch.from_numpy(T).to("cuda")
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(dt)
    f.set_templates(Producer(tt))
    # bound, not copied: the banks are the producers' own memory
    assert f._gpu._bound["data"].ptr == dt.data_ptr()
    assert f._gpu._bound["template"].ptr == tt.data_ptr()
    got = f.run(binsize=n // 4, threshold=0.0)
    np.testing.assert_array_equal(got["index"], ref["index"])
    np.testing.assert_array_equal(got["value"], ref["value"])


def test_stream_ordering_and_rebinding(gpu_device):
    """The bank is produced by PyTorch kernels immediately before set_data (no
    host wait in between), then rewritten in place and set again."""
    import torch
    n, nd, nt = 4096, 4, 64
    rng = np.random.default_rng(401)
    T = cnoise(rng, (nt, n))
    tt = torch.from_numpy(T).to("cuda")
    base = torch.from_numpy(cnoise(rng, (nd, n))).to("cuda")
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_templates(tt)
    for scale in (1.0, 3.0, -0.5):
        dt = torch.empty_like(base)
        torch.mul(base, scale, out=dt)           # pending on PyTorch's stream
        dt[1] += 2.0 * tt[7]                     # more pending work
        f.set_data(dt)
        got = f.run(binsize=512, threshold=0.0)
        ref = host_run(dt.cpu().numpy(), T, binsize=512, threshold=0.0)
        np.testing.assert_array_equal(got["index"], ref["index"])
        np.testing.assert_array_equal(got["value"], ref["value"])
    # the same buffer rewritten in place: set again, then the new contents count
    dt.mul_(2.0)
    f.set_data(dt)
    got = f.run(binsize=512)
    ref = host_run(dt.cpu().numpy(), T, binsize=512)
    np.testing.assert_array_equal(got["value"], ref["value"])


def test_row_update_never_writes_the_producer(gpu_device):
    import torch
    n, nd, nt = 1024, 3, 4
    rng = np.random.default_rng(402)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    row = cnoise(rng, n)
    dt = torch.from_numpy(D).to("cuda")
    tt = torch.from_numpy(T).to("cuda")
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(dt)
    f.set_templates(tt)
    f.set_data(row, index=1)                 # host row into a bound bank
    f.set_templates(torch.from_numpy(T[0]).to("cuda"), index=2)   # device row
    np.testing.assert_array_equal(dt.cpu().numpy(), D)
    np.testing.assert_array_equal(tt.cpu().numpy(), T)
    D2, T2 = D.copy(), T.copy()
    D2[1] = row
    T2[2] = T[0]
    got = f.run(binsize=128)
    ref = host_run(D2, T2, binsize=128)
    np.testing.assert_array_equal(got["index"], ref["index"])
    np.testing.assert_array_equal(got["value"], ref["value"])


def test_series_forward_into_a_bound_bank_keeps_the_producer(gpu_device):
    """Block spectra written into the data rows of a bound bank go to the
    context's own copy: the producer is untouched, the other rows keep the
    producer's contents, and the written rows equal run_series's blocks."""
    import torch
    from apogee_amd import _hip
    n, nt, nd = 1024, 8, 3
    rng = np.random.default_rng(403)
    T = cnoise(rng, (nt, n))
    D = cnoise(rng, (nd, n))
    dt = torch.from_numpy(D).to("cuda")
    ctx = _hip.Context(gpu_device, n, nd, nt)
    ctx.set_rows("template", 0, T)
    ctx.bind_rows("data", ctx.borrow(dt))
    series = rng.standard_normal(2 * n).astype(np.complex64)
    starts = (np.arange(2) * (n // 2)).astype(np.uintp)
    ctx.series_forward(series, starts, row0=1)
    np.testing.assert_array_equal(dt.cpu().numpy(), D)
    idx, val, _ = ctx.peaks(0, nd, 0, nt, 128, 0.0, 0, n)
    ref0 = host_run(D[:1], T, binsize=128, threshold=0.0)
    np.testing.assert_array_equal(val[:1], ref0["value"])
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    blocks = f.run_series(series, starts, np.zeros(2, np.uintp), np.full(2, n, np.uintp), binsize=128)
    np.testing.assert_array_equal(idx[1:], blocks["index"])
    np.testing.assert_array_equal(val[1:], blocks["value"])
    ctx.destroy()


def test_hierarchical_rebuilds_coarse_templates_on_bind(gpu_device):
    import torch
    n, nd, nt, band = 4096, 4, 32, 512
    rng = np.random.default_rng(404)
    D = cnoise(rng, (nd, n))
    T1, T2 = cnoise(rng, (nt, n)), cnoise(rng, (nt, n))
    D[2] += 5 * T2[9]
    h = mf.HierarchicalFilter(n, nd, nt, band=band, device="gpu")
    h.set_coarse_threshold(0.0)
    h.set_data(torch.from_numpy(D).to("cuda"))
    for T in (T1, T2):
        h.set_templates(torch.from_numpy(T).to("cuda"))
        got = h.run(binsize=1024, threshold=0.0)
        ref = host_run(D, T, binsize=1024, threshold=0.0)
        # a zero coarse gate keeps every pair: hierarchical == flat exactly
        np.testing.assert_array_equal(got["index"], ref["index"])
        np.testing.assert_array_equal(got["value"], ref["value"])


def test_noncontiguous_and_wrong_dtype(gpu_device):
    import torch
    n, nd, nt = 512, 2, 3
    rng = np.random.default_rng(405)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(torch.from_numpy(np.ascontiguousarray(D.T)).to("cuda").T)   # a transposed view
    f.set_templates(torch.from_numpy(np.conj(T)).to("cuda").conj())       # lazy conjugate
    ref = host_run(D, T, binsize=64)
    got = f.run(binsize=64)
    np.testing.assert_array_equal(got["value"], ref["value"])
    with pytest.raises(ValueError):
        f.set_data(torch.zeros((nd, n), dtype=torch.complex128, device="cuda"))
    with pytest.raises(ValueError):
        f.set_data(torch.zeros((nd + 1, n), dtype=torch.complex64, device="cuda"))


def test_empty_shared_is_pinned_and_usable(gpu_device):
    n, nd, nt = 1024, 2, 3
    rng = np.random.default_rng(406)
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    D = f.empty_shared((nd, n))
    T = f.empty_shared((nt, n), np.complex64)
    D[...] = cnoise(rng, (nd, n))
    T[...] = cnoise(rng, (nt, n))
    f.set_data(D)
    f.set_templates(T)
    got = f.run(binsize=128)
    ref = host_run(np.array(D), np.array(T), binsize=128)
    np.testing.assert_array_equal(got["value"], ref["value"])
