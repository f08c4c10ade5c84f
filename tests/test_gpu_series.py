"""run_series on the MI355X: device block spectra + filter against the FP64
oracle (block spectrum FFT(series[s:s+n])/n, zero-padded, then the flat
filter with the block's window), the hierarchical one-sided guarantee, and
the reference's contract (tests/test_run_series_flat.py, tests/test_api.py:376-476).
"""
import numpy as np
import pytest

import apogee_amd as mf
from oracle import mf_oracle as O
from _parity import assert_peaks_match

pytestmark = pytest.mark.gpu


def overlap_save_layout(nseries, n, ntaps):
    """Overlap-save blocks stepping n - ntaps + 1; the last block is ragged."""
    valid = n - ntaps + 1
    starts, ws, we = [], [], []
    s = 0
    while s < nseries:
        starts.append(s)
        ws.append(ntaps - 1)
        we.append(min(n, ntaps - 1 + (nseries - s)))
        s += valid
    return (np.array(starts, np.uintp), np.array(ws, np.uintp), np.array(we, np.uintp))


def check_against_oracle(series, starts, ws, we, T, binsize, threshold, got):
    n = T.shape[1]
    spec = O.series_spectra(series, starts, n)
    for b in range(len(starts)):
        assert_peaks_match(spec[b:b + 1], T, got["index"][b:b + 1], got["value"][b:b + 1], binsize,
                           threshold, (int(ws[b]), int(we[b])), context="block %d" % b)


@pytest.mark.parametrize("n", [256, 1024, 4096, 16384, 32768, 65536])
def test_half_overlap_real_series(n, gpu_device):
    """A real float32 series with 50% overlap (starts = b n/2), window [n/4, 3n/4)."""
    rng = np.random.default_rng(n)
    nt, nblk = (5, 12) if n <= 4096 else (3, 5)
    series = rng.standard_normal(nblk * n // 2 + n // 2).astype(np.float32)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts = (np.arange(nblk) * (n // 2)).astype(np.uintp)
    ws = np.full(nblk, n // 4, np.uintp)
    we = np.full(nblk, 3 * n // 4, np.uintp)
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    got = f.run_series(series, starts, ws, we, binsize=n // 8, threshold=0.0)
    assert got.shape == (nblk, nt, 4)
    check_against_oracle(series.astype(np.complex64), starts, ws, we, T, n // 8, 0.0, got)


@pytest.mark.parametrize("n,nblk", [(1024, 64), (1024, 65), (256, 300), (1048576, 40)])
def test_start_tables_inline_and_on_the_device(n, nblk, gpu_device):
    """Up to 64 block starts travel in the forward kernels' arguments, more
    go through a device table; at n = 2^20 the four-step split takes 32
    blocks per chunk, so the inline starts are read past a chunk offset.
    Against the oracle, starts unordered and some past the end."""
    rng = np.random.default_rng(nblk)
    nt = 2
    nser = 3 * n + 1000
    series = (rng.standard_normal(nser) + 1j * rng.standard_normal(nser)).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts = rng.integers(0, nser + n // 2, nblk).astype(np.uintp)
    ws = np.full(nblk, n // 8, np.uintp)
    we = np.full(nblk, n - n // 8, np.uintp)
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    got = f.run_series(series, starts, ws, we, binsize=n // 4, threshold=0.0)
    assert got.shape == (nblk, nt, 3)
    check_against_oracle(series, starts, ws, we, T, n // 4, 0.0, got)


def test_ragged_edges_and_zero_padding(gpu_device):
    n, nt = 2048, 3
    rng = np.random.default_rng(2)
    series = (rng.standard_normal(9000) + 1j * rng.standard_normal(9000)).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts, ws, we = overlap_save_layout(series.size, n, 451)
    nb_first = -(-(int(we[0]) - int(ws[0])) // n)
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    got = f.run_series(series, starts, ws, we, binsize=n, threshold=0.0)
    assert got.shape == (len(starts), nt, nb_first)
    check_against_oracle(series, starts, ws, we, T, n, 0.0, got)


def test_blocks_past_the_end_and_caller_order(gpu_device):
    n, nt = 1024, 2
    rng = np.random.default_rng(3)
    series = (rng.standard_normal(3000) + 1j * rng.standard_normal(3000)).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts = np.array([2500, 0, 5000, 1200, 0], np.uintp)          # one block entirely past the end
    ws = np.array([0, 100, 0, 100, 0], np.uintp)
    we = np.array([512, 612, 512, 612, 512], np.uintp)
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    got = f.run_series(series, starts, ws, we, binsize=128, threshold=0.0)
    check_against_oracle(series, starts, ws, we, T, 128, 0.0, got)
    assert (got["index"][2] == -1).all()                          # an all-zero block reports nothing
    idx, val = f.run_series(series, starts, ws, we, binsize=128, threshold=0.0, raw=True)
    np.testing.assert_array_equal(idx, got["index"])
    with pytest.raises(ValueError, match="no data"):
        f.run()                                                    # data slots consumed


def test_series_matches_block_by_block_runs(gpu_device):
    """run_series == set_data(FFT(block)/n) + run(window), to FP32 rounding."""
    n, nt = 4096, 4
    rng = np.random.default_rng(21)
    series = rng.standard_normal(1 << 15).astype(np.float32)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts, ws, we = overlap_save_layout(series.size, n, 451)
    keep = [i for i in range(len(starts)) if we[i] - ws[i] == we[0] - ws[0]]
    starts, ws, we = starts[keep], ws[keep], we[keep]
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    got = f.run_series(series, starts, ws, we, binsize=n, threshold=0.0).copy()
    for b, s in enumerate(starts):
        blk = np.zeros(n, np.complex64)
        seg = series[int(s):int(s) + n]
        blk[:seg.size] = seg
        f.set_data((np.fft.fft(blk) / n).astype(np.complex64)[None, :])
        want = f.run(binsize=n, threshold=0.0, window=(int(ws[b]), int(we[b])))
        same = got["index"][b] == want["index"][0]
        np.testing.assert_allclose(np.abs(got["value"][b][same]), np.abs(want["value"][0][same]),
                                   rtol=2e-6 * 12)
        # the two FP32 forward transforms (device, NumPy) may only swap lags
        # whose FP64 magnitudes are within the near-tie margin of each other
        absz = np.abs(O.correlate(O.series_spectra(series, [int(s)], n)[:, None, :], T[None]))[0]
        for t in np.nonzero(~same)[0]:
            a, w = int(got["index"][b][t, 0]), int(want["index"][0][t, 0])
            top = absz[t, int(ws[b]):int(we[b])].max()
            assert abs(absz[t, a] - absz[t, w]) <= 4 * 2e-6 * 12 * top, (b, t, a, w)


def test_hierarchical_series_is_identical_or_absent(gpu_device):
    n, nt = 4096, 6
    rng = np.random.default_rng(11)
    k = np.arange(1, n // 2)
    power = np.zeros(n, np.float32)
    power[1:n // 2] = k ** (-7 / 3.0) / ((0.015 * n / k) ** 4 + 1.0)
    power /= power.sum()
    H = np.stack([(np.sqrt(power) * np.exp(2j * np.pi * rng.random(n))) for _ in range(nt)])
    H = (H / np.linalg.norm(H, axis=1, keepdims=True)).astype(np.complex64)
    series = (rng.standard_normal(1 << 16) + 1j * rng.standard_normal(1 << 16)).astype(np.complex64)
    # plant a loud template copy in the series
    h_time = np.fft.ifft(H[2].astype(np.complex128)) * n
    series[20000:20000 + n] += (12 * n * h_time / n).astype(np.complex64)
    starts, ws, we = overlap_save_layout(series.size, n, 451)
    keep = [i for i in range(len(starts)) if we[i] - ws[i] == we[0] - ws[0]]
    starts, ws, we = starts[keep], ws[keep], we[keep]
    flat = mf.MatchedFilter(n, 1, nt, device="gpu")
    hier = mf.HierarchicalFilter(n, 1, nt, snr=5.0, fd=1e-2, device="gpu")
    hier.set_reference(power)
    flat.set_templates(H)
    hier.set_templates(H)
    a = flat.run_series(series, starts, ws, we, binsize=512, threshold=5.0)
    b = hier.run_series(series, starts, ws, we, binsize=512, threshold=5.0)
    fired = b["index"] >= 0
    assert fired.any(), "the planted copy was dismissed"
    assert not (fired & (a["index"] < 0)).any()
    np.testing.assert_array_equal(a["index"][fired], b["index"][fired])
    np.testing.assert_array_equal(a["value"][fired], b["value"][fired])
    pairs, trig = hier.stats
    assert pairs == len(starts) * nt and trig < pairs
    # the records the library fills (mfh_run_series_rec, mfh_hier_run_series_rec)
    # equal the raw arrays of the same call bit for bit
    for f, rec in ((flat, a), (hier, b)):
        idx, val = f.run_series(series, starts, ws, we, binsize=512, threshold=5.0, raw=True)
        assert rec.dtype == mf.PEAK_DTYPE and rec.shape == idx.shape
        np.testing.assert_array_equal(rec["index"], idx)
        np.testing.assert_array_equal(rec["value"], val)


@pytest.mark.parametrize("nblk", [1, 4, 6])
def test_hierarchical_series_dismisses_pure_noise(nblk, gpu_device):
    """The reference's tests/test_hier_series_scale.py:45-75: block spectra
    carry the 1/n of the forward transform on the hierarchical route as on the
    flat one, so on unit complex noise the flat peaks stay near 0.07 and no
    pair clears a gate calibrated for SNR 5.5."""
    n, nt = 4096, 8
    rng = np.random.default_rng(3)
    k = np.arange(1, n // 2)
    power = np.zeros(n, np.float32)
    power[1:n // 2] = k ** (-7 / 3.0) / ((0.015 * n / k) ** 4 + 1.0)
    power /= power.sum()
    H = np.stack([np.sqrt(power) * np.exp(2j * np.pi * rng.random(n)) for _ in range(nt)])
    H = (H / np.linalg.norm(H, axis=1, keepdims=True)).astype(np.complex64)
    step = n // 2
    series = (rng.standard_normal(n + step * nblk) + 1j * rng.standard_normal(n + step * nblk)).astype(np.complex64)
    starts = (np.arange(nblk) * step).astype(np.uintp)
    ws = np.zeros(nblk, np.uintp)
    we = np.full(nblk, n // 2, np.uintp)
    flat = mf.MatchedFilter(n, nblk, nt, device="gpu")
    flat.set_templates(H)
    a = flat.run_series(series, starts, ws, we, binsize=n, threshold=0.0)
    assert (a["index"] >= 0).all() and np.abs(a["value"]).max() < 0.2
    hier = mf.HierarchicalFilter(n, nblk, nt, snr=5.5, fd=1e-2, device="gpu")
    hier.set_reference(power)
    hier.set_templates(H)
    b = hier.run_series(series, starts, ws, we, binsize=n, threshold=0.0)
    assert (b["index"] < 0).all(), "%d peaks reported on pure noise" % (b["index"] >= 0).sum()
    assert hier.stats == (nblk * nt, 0)


def test_device_resident_series(gpu_device):
    import torch
    n, nt = 1024, 2
    rng = np.random.default_rng(5)
    series = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts = np.arange(0, 4000, 512).astype(np.uintp)
    ws = np.zeros(len(starts), np.uintp)
    we = np.full(len(starts), 512, np.uintp)
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    host = f.run_series(series, starts, ws, we, binsize=256).copy()
    dev = f.run_series(torch.from_numpy(series).to("cuda"), starts, ws, we, binsize=256)
    np.testing.assert_array_equal(host["index"], dev["index"])
    np.testing.assert_array_equal(host["value"], dev["value"])


@pytest.mark.parametrize("n", [1024, 65536, 262144])
def test_series_forward_fills_data_rows(n, gpu_device):
    """mfh_series_forward writes the block spectra run_series uses: filtering
    those rows equals run_series, and the rows equal FFT(block)/n."""
    from apogee_amd import _hip
    rng = np.random.default_rng(7 + n)
    nt, nblk = 3, 4
    series = rng.standard_normal(nblk * n // 2 + n // 3).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    starts = (np.arange(nblk) * (n // 2)).astype(np.uintp)
    ctx = _hip.Context(gpu_device, n, nblk + 1, nt)
    ctx.set_rows("template", 0, T)
    ctx.series_forward(series, starts, row0=1)
    idx, val, _ = ctx.peaks(1, nblk, 0, nt, n // 8, 0.0, 0, n)
    f = mf.MatchedFilter(n, 1, nt, device="gpu")
    f.set_templates(T)
    got = f.run_series(series, starts, np.zeros(nblk, np.uintp), np.full(nblk, n, np.uintp),
                       binsize=n // 8, threshold=0.0)
    np.testing.assert_array_equal(idx, got["index"])
    np.testing.assert_array_equal(val, got["value"])
    spec = O.series_spectra(series, starts, n)
    check_against_oracle(series, starts, np.zeros(nblk, np.uintp), np.full(nblk, n, np.uintp), T, n // 8, 0.0, got)
    assert spec.shape == (nblk, n)
    with pytest.raises(RuntimeError):
        ctx.series_forward(series, starts, row0=2)        # rows 2..5 exceed ndata = 5
