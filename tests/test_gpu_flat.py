"""Flat filter on the MI355X against the FP64 oracle, the golden fixtures and
the FP32 CPU restatement.

The cases follow the ideas of the reference's tests/test_independent.py,
tests/test_adversarial.py and tests/test_api.py:68-150 (direct sums, closed
forms, bin and window edges, ragged bins, amplitude range, degenerate inputs,
strict threshold, odd batch shapes, every length, binsize 1, the underflow
cliff, batch isolation, NaN containment, sub-range = slice).  Tolerances and
the near-tie rule are in tests/_parity.py.
"""
import os

import numpy as np
import pytest

import apogee_amd as mf
from oracle import mf_oracle as O
from _parity import assert_peaks_match, tol

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def run(D, T, binsize=None, threshold=0.0, window=None, **kw):
    f = mf.MatchedFilter(D.shape[1], D.shape[0], T.shape[0], device="gpu")
    f.set_data(D)
    f.set_templates(T)
    w = {} if window is None else {"window": window}
    return f.run(binsize=binsize, threshold=threshold, **w, **kw)


def planted(n, lag, amp=1.0, seed=0):
    rng = np.random.default_rng(seed)
    h = cnoise(rng, n)
    h /= np.linalg.norm(h)
    ramp = np.exp(-2j * np.pi * lag * np.arange(n) / n)
    return (amp * h * ramp).astype(np.complex64)[None], h[None]


@pytest.mark.parametrize("name", ["config1_small", "ragged", "planted", "closed_forms"])
def test_golden_fixtures(name, gpu_device):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as g:
        g = {k: g[k] for k in g.files}
    bs, thr, win = int(g["binsize"]), float(g["threshold"]), (int(g["ws"]), int(g["we"]))
    pk = run(g["D"], g["T"], bs, thr, win)
    assert_peaks_match(g["D"], g["T"], pk["index"], pk["value"], bs, thr, win, context=name)
    scale = np.sqrt(g["best"]).max()
    live = g["best"] > (16 * tol(g["D"].shape[1]) * scale) ** 2
    # on clearly separated bins the index is exactly the definition's
    sep = live & (g["idx"] >= 0)
    agree = (pk["index"] == g["idx"])[sep].mean()
    assert agree > 0.99, agree


def test_config1_cpu_parity_workload(gpu_device, mfcpu):
    """n=4096, 16 data x 256 templates, binsize 512: the CPU-parity config."""
    from oracle import cpu
    rng = np.random.default_rng(20261202 + 1)
    n, nd, nt = 4096, 16, 256
    t_time = rng.standard_normal((nt, n))
    T = np.fft.fft(t_time).astype(np.complex64)
    x = np.empty((nd, n))
    for d in range(nd):
        x[d] = 0.5 * np.roll(t_time[d % nt], int(rng.integers(0, n))) + rng.standard_normal(n)
    D = np.fft.fft(x).astype(np.complex64)
    pk = run(D, T, 512, 0.0)
    ex = assert_peaks_match(D, T, pk["index"], pk["value"], 512, 0.0, context="config1")
    ci, cv, _ = cpu.flat(D, T, 512, 0.0)
    assert (ci == pk["index"]).mean() > 0.999
    print("config1: near-ties excused %d of %d bins; CPU restatement index agreement %.5f"
          % (ex, pk.size, (ci == pk["index"]).mean()))


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536])
def test_every_length_against_the_oracle(n, gpu_device):
    rng = np.random.default_rng(n)
    D, T = cnoise(rng, (3, n)), cnoise(rng, (5, n))
    for bs in (n, max(1, n // 8), 37):
        pk = run(D, T, bs, 0.0)
        assert_peaks_match(D, T, pk["index"], pk["value"], bs, 0.0, context="n=%d bs=%d" % (n, bs))


@pytest.mark.parametrize("n", [64, 256, 1024, 4096])
def test_direct_sum_every_lag(n, gpu_device):
    rng = np.random.default_rng(n + 1)
    D, T = cnoise(rng, (2, n)), cnoise(rng, (2, n))
    pk = run(D, T, 1, 0.0)
    for i in range(2):
        for j in range(2):
            want = O.direct_correlate(D[i], T[j]) if n <= 1024 else O.correlate(D[i], T[j])
            got = pk["value"][i, j]
            assert np.abs(got - want).max() <= tol(n) * np.abs(want).max()
            np.testing.assert_array_equal(pk["index"][i, j], np.arange(n))


def test_self_correlation_and_flat_spectrum(gpu_device):
    n = 1024
    rng = np.random.default_rng(1)
    h = cnoise(rng, n)
    pk = run(h[None], h[None], n, 0.0)
    want = float(np.sum(np.abs(h.astype(np.complex128)) ** 2))
    assert pk["index"][0, 0, 0] == 0
    assert abs(pk["value"][0, 0, 0]) == pytest.approx(want, rel=1e-5)
    ones = np.ones((1, 512), np.complex64)
    pk = run(ones, ones, 1, 0.0)
    m = np.abs(pk["value"][0, 0])
    assert m[0] == pytest.approx(512.0, rel=1e-6)
    assert m[1:].max() < 1e-3 * 512


def test_known_lag_sign_convention(gpu_device):
    n = 4096
    rng = np.random.default_rng(1)
    x = cnoise(rng, n)
    for lag in (0, 1, 37, n // 2, n - 1):
        T = (x * np.exp(2j * np.pi * np.arange(n) * lag / n)).astype(np.complex64)
        assert run(x[None], T[None], n, 0.0)["index"][0, 0, 0] == lag


@pytest.mark.parametrize("n,binsize", [(1024, 256), (4096, 512), (2048, 128)])
def test_peak_on_every_bin_boundary(n, binsize, gpu_device):
    for b in range(n // binsize):
        for lag in (b * binsize, (b + 1) * binsize - 1):
            d, h = planted(n, lag, amp=10.0)
            pk = run(d, h, binsize, 1.0)
            assert pk["index"][0, 0, b] == lag
            others = [int(v) for j, v in enumerate(pk["index"][0, 0]) if j != b]
            assert all(v == -1 or abs(v - lag) > binsize for v in others)


def test_window_edges(gpu_device):
    n, ws, we = 4096, 1024, 3072
    for lag in (ws, ws + 1, we - 2, we - 1):
        d, h = planted(n, lag, amp=10.0)
        assert run(d, h, we - ws, 1.0, (ws, we))["index"][0, 0, 0] == lag
    for lag in (ws - 1, we):
        d, h = planted(n, lag, amp=10.0)
        got = int(run(d, h, we - ws, 1.0, (ws, we))["index"][0, 0, 0])
        assert got == -1 or ws <= got < we
        assert got != lag


@pytest.mark.parametrize("n,ws,we,binsize", [
    (4096, 0, 4096, 384), (4096, 16, 4000, 512), (2048, 100, 1900, 300),
    (4096, 37, 4000, 300), (4096, 13, 3999, 7), (4096, 100, 900, 8192), (1024, 500, 501, 3),
])
def test_ragged_bins_and_windows(n, ws, we, binsize, gpu_device):
    rng = np.random.default_rng(3)
    D, T = cnoise(rng, (2, n)), cnoise(rng, (2, n))
    pk = run(D, T, binsize, 0.0, (ws, we))
    assert_peaks_match(D, T, pk["index"], pk["value"], binsize, 0.0, (ws, we),
                       context="%s" % ((n, ws, we, binsize),))


def fuzz_cases(seed, n, count, with_window):
    """Seeded (binsize, ws, we): binsize log-uniform in [1, n], so both many
    small bins and one large bin occur; windows anywhere in [0, n)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        binsize = max(1, min(n, int(round(float(n) ** rng.uniform(0.0, 1.0)))))
        if with_window:
            ws = int(rng.integers(0, n - 1))
            we = int(rng.integers(ws + 1, n + 1))
        else:
            ws, we = 0, n
        out.append((binsize, ws, we))
    return out


@pytest.mark.parametrize("n", [1024, 2048, 4096, 65536])
@pytest.mark.parametrize("with_window,seed", [(False, 20260923), (True, 8675309)])
def test_fuzz_binsizes_and_windows(n, with_window, seed, gpu_device):
    """Random bins and windows, as the reference's
    tests/test_bin_window_matrix.py:197-231 fuzz, on the one-pass kernels and
    (n = 65536) the four-step path."""
    rng = np.random.default_rng(seed + n)
    D, T = cnoise(rng, (1, n)), cnoise(rng, (2, n))
    for binsize, ws, we in fuzz_cases(seed, n, 25, with_window):
        pk = run(D, T, binsize, 1.0, (ws, we))
        assert_peaks_match(D, T, pk["index"], pk["value"], binsize, 1.0, (ws, we),
                           context="n=%d binsize=%d window=(%d,%d)" % (n, binsize, ws, we))


@pytest.mark.parametrize("n", [1024, 16384])
def test_fuzz_with_a_threshold_that_bites(n, gpu_device):
    """Random bins and windows under a threshold at the 99th percentile of
    |z|, so most bins are dismissed (the reference's
    tests/test_bin_window_matrix.py:234-251)."""
    rng = np.random.default_rng(3 + n)
    D, T = cnoise(rng, (1, n)), cnoise(rng, (2, n))
    thr = float(np.percentile(np.abs(O.correlate(D[0], T[0])), 99.0))
    for binsize, ws, we in fuzz_cases(4242, n, 20, True):
        pk = run(D, T, binsize, thr, (ws, we))
        assert_peaks_match(D, T, pk["index"], pk["value"], binsize, thr, (ws, we),
                           context="n=%d binsize=%d window=(%d,%d) thr=%.3f" % (n, binsize, ws, we, thr))


@pytest.mark.parametrize("amp", [1e-12, 1e-6, 1.0, 1e6, 1e12])
def test_amplitude_scale_moves_nothing(amp, gpu_device):
    n, lag = 2048, 733
    d, h = planted(n, lag, amp=amp)
    pk = run(d, h, n, 0.0)
    assert pk["index"][0, 0, 0] == lag
    assert abs(pk["value"][0, 0, 0]) == pytest.approx(amp, rel=1e-3)


def test_wildly_different_template_scales(gpu_device):
    n = 1024
    rng = np.random.default_rng(5)
    scales = np.array([1e-9, 1e-3, 1.0, 1e3, 1e9])
    h = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    h = (h * scales[:, None]).astype(np.complex64)
    d = cnoise(rng, (2, n))
    pk = run(d, h, n, 0.0)
    assert_peaks_match(d, h, pk["index"], pk["value"], n, 0.0, context="scales")


def test_degenerate_inputs(gpu_device):
    n = 1024
    pk = run(np.zeros((1, n), np.complex64), np.ones((1, n), np.complex64), n, 1e-6)
    assert pk["index"][0, 0, 0] == -1 and pk["value"][0, 0, 0] == 0
    rng = np.random.default_rng(7)
    pk = run(cnoise(rng, (1, n)), np.zeros((1, n), np.complex64), n, 0.0)
    assert pk["index"][0, 0, 0] == -1 and pk["value"][0, 0, 0] == 0


def test_all_ties_report_the_lowest_lag(gpu_device):
    """A single non-zero frequency gives |z| equal at every lag.

    The reference leaves the tied index unspecified; this build reports the
    lowest lag whose FP32 |z|^2 attains the bin maximum.
    """
    n = 1024
    d = np.zeros((1, n), np.complex64)
    h = np.zeros((1, n), np.complex64)
    d[0, 7] = 1.0
    h[0, 7] = 1.0
    for bs in (n, 100, 1):
        pk = run(d, h, bs, 0.0)
        z = np.exp(2j * np.pi * 7 * np.arange(n) / n)
        for b in range(pk.shape[-1]):
            i = int(pk["index"][0, 0, b])
            assert b * bs <= i < min(n, (b + 1) * bs)
            assert abs(pk["value"][0, 0, b] - z[i]) < 1e-5     # value belongs to its lag
        m2 = np.abs(pk["value"][0, 0]) ** 2
        assert m2.min() >= 1 - 1e-5


@pytest.mark.parametrize("n", [1024, 4096, 16384, 32768, 65536])
@pytest.mark.parametrize("binsize", [None, 257, 1])
def test_tied_peaks_keep_index_and_value_together(n, binsize, gpu_device):
    """The reference's tests/test_gpu_peak_ties.py:8-29: a delta at n/4 in data
    and templates gives z[m] = i**m, |z| = 1 at every lag.  Whichever lag a bin
    reports, it lies in the window and the bin, and its value is that lag's
    z; a threshold of 2 empties every bin.  Repeated runs of one filter."""
    rows = 32 if n == 16384 and binsize is None else 4
    bank = np.zeros((rows, n), np.complex64)
    bank[:, n // 4] = 1
    f = mf.MatchedFilter(n, rows, rows, device="gpu")
    f.set_data(bank)
    f.set_templates(bank)
    phases = np.array([1, 1j, -1, -1j], np.complex64)
    for _ in range(12 if rows == 32 else 2):
        out = f.run(binsize=binsize, window=(7, n - 3))
        index = out["index"]
        assert ((index >= 7) & (index < n - 3)).all()
        np.testing.assert_allclose(out["value"], phases[index % 4], atol=2e-5)
        if binsize is not None:
            np.testing.assert_array_equal((index - 7) // binsize,
                                          np.broadcast_to(np.arange(index.shape[-1]), index.shape))
    empty = f.run(binsize=binsize, threshold=2.0, window=(7, n - 3))
    np.testing.assert_array_equal(empty["index"], -1)
    np.testing.assert_array_equal(empty["value"], 0)


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 4096, 16384, 65536])
def test_equal_maxima_everywhere_report_each_bins_first_lag(n, gpu_device):
    """The reference's tests/test_cpu_single_peak.py:7-34 contract, on the
    device: a DC-only product makes z[m] = D[0] exactly at every lag (the
    transform of a delta at frequency 0 only ever multiplies it by 1), so every
    bin ties over all its lags and reports its first lag, unless the strict
    threshold equals |D[0]| = 5.  Sub-ranges of templates, unaligned windows,
    a one-sample window, binsizes above the window and the counts output."""
    f = mf.MatchedFilter(n, 3, 37, device="gpu")
    data = np.zeros((3, n), np.complex64)
    data[:, 0] = [3 + 4j, -5j, 0]
    templates = np.zeros((37, n), np.complex64)
    templates[:, 0] = 1
    f.set_templates(templates)
    f.set_data(data)
    for start, count in ((0, 1), (0, 37), (1, 33)):
        for lo, hi in ((0, n), (7, n - 9), (11, 12)):
            for binsize in (hi - lo, 17, n + 3):
                for threshold in (0.0, 4.0, 5.0, 6.0):
                    out, counts = f.run(templates=(start, count), window=(lo, hi),
                                        binsize=binsize, threshold=threshold, counts=True)
                    nbins = (hi - lo + binsize - 1) // binsize
                    want_i = np.full((3, count, nbins), -1)
                    want_v = np.zeros((3, count, nbins), np.complex64)
                    if threshold < 5:
                        want_i[:2] = lo + np.arange(nbins) * binsize
                        want_v[:2] = data[:2, :1, None]
                    ctx = "templates %s window %s binsize %d threshold %g" % (
                        (start, count), (lo, hi), binsize, threshold)
                    np.testing.assert_array_equal(out["index"], want_i, err_msg=ctx)
                    np.testing.assert_array_equal(out["value"], want_v, err_msg=ctx)
                    np.testing.assert_array_equal(counts, (want_i >= 0).sum(axis=-1), err_msg=ctx)


def test_threshold_is_strict_on_the_fp32_key(gpu_device):
    n, lag = 1024, 300
    d, h = planted(n, lag, amp=4.0)
    pk = run(d, h, n, 0.0)
    v = complex(pk["value"][0, 0, 0])
    re, im = np.float32(v.real), np.float32(v.imag)
    # the device key: fma(re, re, im*im) in FP32 (exact products in FP64)
    m2 = np.float32(float(re) * float(re) + float(np.float32(im * im)))
    hi = np.float32(np.sqrt(np.float64(m2)))
    while np.float32(hi * hi) < m2:
        hi = np.nextafter(hi, np.float32(np.inf))
    lo = np.nextafter(np.float32(np.sqrt(np.float64(m2))), np.float32(0))
    while np.float32(lo * lo) >= m2:
        lo = np.nextafter(lo, np.float32(0))
    assert run(d, h, n, float(hi))["index"][0, 0, 0] == -1
    assert run(d, h, n, float(lo))["index"][0, 0, 0] == lag


@pytest.mark.parametrize("nd,nt", [(1, 1), (1, 17), (17, 1), (3, 5), (13, 7), (40, 3)])
def test_odd_batch_shapes(nd, nt, gpu_device):
    n = 1024
    rng = np.random.default_rng(nd * 100 + nt)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    pk = run(D, T, 256, 0.0)
    assert_peaks_match(D, T, pk["index"], pk["value"], 256, 0.0, context="%dx%d" % (nd, nt))


def test_binsize_one_reports_every_lag(gpu_device):
    n = 4096
    rng = np.random.default_rng(11)
    D, T = cnoise(rng, (1, n)), cnoise(rng, (1, n))
    pk = run(D, T, 1, 0.0)
    np.testing.assert_array_equal(pk["index"][0, 0], np.arange(n))
    z = O.correlate(D[0], T[0])
    assert np.abs(pk["value"][0, 0] - z).max() <= tol(n) * np.abs(z).max()


def test_underflow_cliff_reports_nothing(gpu_device):
    n, lag = 1024, 321
    rng = np.random.default_rng(8)
    h = cnoise(rng, n)
    h /= np.linalg.norm(h)
    ramp = np.exp(-2j * np.pi * lag * np.arange(n) / n)

    def at(scale):
        d = (np.float32(scale) * h * ramp).astype(np.complex64)
        pk = run(d[None], h[None], n, 0.0)
        return int(pk["index"][0, 0, 0]), complex(pk["value"][0, 0, 0])

    i, v = at(1e-20)
    assert i == lag and abs(v) == pytest.approx(1e-20, rel=1e-2)
    i, v = at(1e-24)
    assert i == -1 and v == 0


def test_subnormal_inputs_are_clean(gpu_device):
    n, lag = 1024, 321
    rng = np.random.default_rng(8)
    h = cnoise(rng, n)
    h /= np.linalg.norm(h)
    ramp = np.exp(-2j * np.pi * lag * np.arange(n) / n)
    for scale in (1e-30, 1e-35, 1e-38, 1e-40, 1e-44):
        d = (np.float32(scale) * h * ramp).astype(np.complex64)
        pk = run(d[None], h[None], n, 0.0)
        m, i = abs(pk["value"][0, 0, 0]), int(pk["index"][0, 0, 0])
        assert np.isfinite(m) and m >= 0
        if m > 0:
            assert i == lag


def test_a_pair_alone_equals_the_pair_in_a_batch(gpu_device):
    n, nd, nt = 1024, 6, 5
    rng = np.random.default_rng(4)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    ref = run(D, T, 256, 0.0)
    for i in range(nd):
        for j in range(nt):
            one = run(D[i:i + 1], T[j:j + 1], 256, 0.0)
            np.testing.assert_array_equal(one["index"][0, 0], ref["index"][i, j])
            np.testing.assert_array_equal(one["value"][0, 0], ref["value"][i, j])


def test_nan_is_contained_to_its_segment(gpu_device):
    n, nd, nt = 512, 4, 3
    rng = np.random.default_rng(6)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    ref = run(D, T, 128, 0.0)
    bad = D.copy()
    bad[2, 17] = np.complex64(complex(np.nan, 0.0))
    got = run(bad, T, 128, 0.0)
    for i in range(nd):
        if i != 2:
            np.testing.assert_array_equal(got["index"][i], ref["index"][i])
            np.testing.assert_array_equal(got["value"][i], ref["value"][i])


@pytest.mark.parametrize("n", [1024, 4096, 16384])
def test_one_bin_nan_and_infinity(n, gpu_device):
    """One bin over the whole transform (the wave-reduction selection): a
    NaN spectrum turns every lag of its pairs into NaN, which never wins, so
    those pairs report -1 / 0 and every other pair is untouched; an infinite
    sample is allowed to win, and then only a lag whose |z|^2 is +inf."""
    nd, nt = 4, 3
    rng = np.random.default_rng(61)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    ref = run(D, T, None, 0.0)
    bad = D.copy()
    bad[2, 17] = np.complex64(complex(np.nan, 0.0))
    got = run(bad, T, None, 0.0)
    assert (got["index"][2] == -1).all() and (got["value"][2] == 0).all()
    for i in range(nd):
        if i != 2:
            np.testing.assert_array_equal(got["index"][i], ref["index"][i])
            np.testing.assert_array_equal(got["value"][i], ref["value"][i])
    big = D.copy()
    big[1, 0] = np.complex64(complex(np.inf, 0.0))   # f = 0: z[k] += inf for every k
    got = run(big, T, None, 0.0)
    for j in range(nt):
        k = got["index"][1, j, 0]
        if k >= 0:
            z = got["value"][1, j, 0]
            assert np.isinf(np.float32(z.real) ** 2 + np.float32(z.imag) ** 2), (j, k, z)
    for i in range(nd):
        if i != 1:
            np.testing.assert_array_equal(got["index"][i], ref["index"][i])


def test_window_subrange_raw_and_counts(gpu_device):
    n, nd, nt = 4096, 4, 3
    rng = np.random.default_rng(3)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(D)
    f.set_templates(T)
    full = f.run(binsize=1024, threshold=0.0, window=(500, 3500))
    assert ((full["index"] >= 500) & (full["index"] < 3500)).all()
    sub = f.run(binsize=1024, threshold=0.0, window=(500, 3500), data=(1, 2), templates=(0, 2))
    np.testing.assert_array_equal(sub["index"], full["index"][1:3, 0:2])
    np.testing.assert_array_equal(sub["value"], full["value"][1:3, 0:2])
    idx, val = f.run(binsize=1024, threshold=0.0, window=(500, 3500), raw=True)
    np.testing.assert_array_equal(idx, full["index"])
    np.testing.assert_array_equal(val, full["value"])
    pk, cnt = f.run(binsize=512, threshold=150.0, counts=True)
    np.testing.assert_array_equal(cnt, (pk["index"] >= 0).sum(axis=2))


def test_single_row_setters_and_device_tensors(gpu_device):
    import torch
    n, nd, nt = 2048, 3, 4
    rng = np.random.default_rng(12)
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    for i in range(nd):
        f.set_data(D[i], index=i)
    f.set_templates(torch.from_numpy(T).to("cuda"))
    a = f.run(binsize=256).copy()
    b = run(D, T, 256)
    np.testing.assert_array_equal(a["index"], b["index"])
    np.testing.assert_array_equal(a["value"], b["value"])


def test_many_bins_split_across_launches(gpu_device):
    n = 16384
    rng = np.random.default_rng(13)
    D, T = cnoise(rng, (1, n)), cnoise(rng, (2, n))
    pk = run(D, T, 3, 0.0)
    assert_peaks_match(D, T, pk["index"], pk["value"], 3, 0.0, context="16384 bs=3")


# ------------------------------------------------- four-step lengths

def test_config2b_workload(gpu_device):
    """n = 65536, 4 data x 128 templates, binsize 1024 (64 bins), threshold 5.5
    on sigma-normalised |z|; 5% of pairs carry their template at SNR 8."""
    n, nd, nt = 65536, 4, 128
    rng = np.random.default_rng(20261202 + 22)
    t_time = rng.standard_normal((nt, n))
    T = np.fft.fft(t_time)
    T = (T / np.sqrt(n * np.sum(np.abs(T) ** 2, axis=1, keepdims=True))).astype(np.complex64)
    x = rng.standard_normal((nd, n))
    inj = rng.random((nd, nt)) < 0.05
    lag = {}
    for d, t in zip(*np.nonzero(inj)):
        s = int(rng.integers(0, n))
        lag[(d, t)] = s
        x[d] += np.roll(t_time[t], s) * 8.0 / np.linalg.norm(t_time[t])
    D = np.fft.fft(x).astype(np.complex64)
    pk = run(D, T, 1024, 5.5)
    assert pk.shape == (nd, nt, 64)
    assert_peaks_match(D, T, pk["index"], pk["value"], 1024, 5.5, context="config 2b", pair_chunk=64)
    for (d, t), s in lag.items():
        assert pk["index"][d, t, s // 1024] == s


@pytest.mark.parametrize("n", [32768, 65536])
@pytest.mark.parametrize("binsize,window", [(None, None), (1000, (37, 30000)), (7, (100, 20000)),
                                            (4096, (0, 32768))])
def test_four_step_bins_and_windows(n, binsize, window, gpu_device):
    rng = np.random.default_rng(n + (binsize or 0))
    D, T = cnoise(rng, (2, n)), cnoise(rng, (3, n))
    pk = run(D, T, binsize, 0.0, window)
    bs = n if binsize is None else binsize
    assert_peaks_match(D, T, pk["index"], pk["value"], bs, 0.0, window,
                       context="n=%d bs=%s w=%s" % (n, binsize, window))


def test_four_step_known_lags_and_subranges(gpu_device):
    n = 65536
    rng = np.random.default_rng(9)
    x = cnoise(rng, (3, n))
    lags = [0, 255, 256, 40000, n - 1]
    T = np.stack([(x[0] * np.exp(2j * np.pi * np.arange(n) * g / n)).astype(np.complex64)
                  for g in lags])
    f = mf.MatchedFilter(n, 3, len(lags), device="gpu")
    f.set_data(x)
    f.set_templates(T)
    full = f.run(binsize=n).copy()
    assert [int(full["index"][0, j, 0]) for j in range(len(lags))] == lags
    sub = f.run(binsize=n, data=(0, 2), templates=(2, 3))
    np.testing.assert_array_equal(sub["index"], full["index"][0:2, 2:5])
    np.testing.assert_array_equal(sub["value"], full["value"][0:2, 2:5])


@pytest.mark.parametrize("n", [131072, 262144, 524288, 1048576])
def test_lengths_beyond_the_reference_gpu_set(n, gpu_device):
    """131072 ... 1048576 through the same four-step split, against the FP64
    oracle, with a planted lag per pair and n/8 bins and 1000-lag bins."""
    rng = np.random.default_rng(n)
    nd, nt = 2, 3
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    f_ = np.arange(n)
    for d in range(nd):
        lag = int(rng.integers(n))
        D[d] += (0.05 * T[d % nt] * np.exp(2j * np.pi * f_ * lag / n)).astype(np.complex64)
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(D)
    f.set_templates(T)
    for binsize, window in ((n // 8, None), (1000, (777, n - 5))):
        kw = {} if window is None else {"window": window}
        got = f.run(binsize=binsize, threshold=0.0, **kw)
        assert_peaks_match(D, T, got["index"], got["value"], binsize, 0.0, window,
                           context="n=%d binsize=%d" % (n, binsize))


def test_records_own_their_page_locked_buffers(gpu_device, monkeypatch):
    """run() records come from pooled page-locked buffers the device writes
    directly (mfh_host_alloc): a result held across calls keeps its values,
    a freed result's buffer serves the next call, the values equal the raw
    arrays (themselves one pooled buffer), and mfh_host_free refuses memory
    it did not hand out."""
    import ctypes
    import gc
    from apogee_amd import _hip
    rng = np.random.default_rng(78)
    n, nd, nt = 1024, 2, 9
    D1, D2, T = cnoise(rng, (nd, n)), cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_templates(T)
    f.set_data(D1)
    a = f.run(binsize=256)
    held = a.copy()
    f.set_data(D2)
    b = f.run(binsize=256)
    np.testing.assert_array_equal(a, held)          # the second call wrote elsewhere
    assert not np.array_equal(a["index"], b["index"])
    idx, val = f.run(binsize=256, raw=True)
    np.testing.assert_array_equal(b["index"], idx)
    np.testing.assert_array_equal(b["value"], val)
    p = b.ctypes.data
    del b
    gc.collect()
    c = f.run(binsize=256)
    assert c.ctypes.data == p                       # the freed buffer, reused
    np.testing.assert_array_equal(c["index"], idx)
    np.testing.assert_array_equal(c["value"], val)
    # raw results: indices and values back to back in one pooled buffer
    idx2, val2 = f.run(binsize=256, raw=True)
    assert val2.ctypes.data == idx2.ctypes.data + idx2.nbytes
    np.testing.assert_array_equal(idx2, idx)
    np.testing.assert_array_equal(val2, val)
    assert _hip.lib().mfh_host_free(ctypes.c_void_p(np.empty(4).ctypes.data)) < 0
    # results above the pool's size limit: plain arrays through the staging copy
    monkeypatch.setattr(_hip, "_POOL_MAX_BYTES", 64)
    d = f.run(binsize=256)
    idx3, val3 = f.run(binsize=256, raw=True)
    assert val3.ctypes.data != idx3.ctypes.data + idx3.nbytes
    np.testing.assert_array_equal(d["index"], idx)
    np.testing.assert_array_equal(d["value"], val)
    np.testing.assert_array_equal(idx3, idx)
    np.testing.assert_array_equal(val3, val)


@pytest.mark.parametrize("n", [4096, 32768, 65536, 262144])
def test_raw_results_written_in_place_equal_the_records(n, gpu_device):
    """raw=True results are written into their pooled page-locked buffer by
    the filter itself, records through the readback kernel: the two agree
    bit for bit at one-pass and four-step lengths, with a window, ragged
    bins and a threshold, counts included."""
    rng = np.random.default_rng(n + 5)
    nd, nt = 2, 3
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    D[1] += 3.0 * T[2]
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(D)
    f.set_templates(T)
    for kw in ({"binsize": n // 8 + 3, "threshold": 0.0, "window": (7, n - 5)},
               {"binsize": n, "threshold": 4.0}):
        rec, cnt = f.run(counts=True, **kw)
        (idx, val), cnt_raw = f.run(counts=True, raw=True, **kw)
        assert val.ctypes.data == idx.ctypes.data + idx.nbytes
        np.testing.assert_array_equal(rec["index"], idx, err_msg=str(kw))
        np.testing.assert_array_equal(rec["value"], val, err_msg=str(kw))
        np.testing.assert_array_equal(cnt, cnt_raw, err_msg=str(kw))
        np.testing.assert_array_equal(cnt, (idx >= 0).sum(axis=2), err_msg=str(kw))


def test_records_equal_the_raw_arrays(gpu_device):
    """run() has the library fill PEAK_DTYPE records directly
    (mfh_flat_peaks_rec, mfh_hier_peaks_rec); they must equal the raw
    (index, value) arrays of the same run bit for bit, counts included:
    flat with ragged bins over a window and with a threshold, hierarchical
    with one bin."""
    rng = np.random.default_rng(77)
    n, nd, nt = 4096, 3, 37
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    D[1] += 4.0 * T[5]
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(D)
    f.set_templates(T)
    for kw in ({"binsize": 300, "threshold": 0.0, "window": (17, 3999)}, {"binsize": n, "threshold": 40.0}):
        rec, cnt = f.run(counts=True, **kw)
        (idx, val), cnt_raw = f.run(counts=True, raw=True, **kw)
        assert rec.dtype == mf.PEAK_DTYPE and rec.shape == idx.shape, kw
        np.testing.assert_array_equal(rec["index"], idx, err_msg=str(kw))
        np.testing.assert_array_equal(rec["value"], val, err_msg=str(kw))
        np.testing.assert_array_equal(cnt, cnt_raw, err_msg=str(kw))
        np.testing.assert_array_equal(cnt, (idx >= 0).sum(axis=2), err_msg=str(kw))
    h = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2, band=512, device="gpu")
    h.set_coarse_threshold(3.0)
    h.set_data(D)
    h.set_templates(T)
    rec, cnt = h.run(binsize=n, threshold=4.5, counts=True)
    (idx, val), cnt_raw = h.run(binsize=n, threshold=4.5, counts=True, raw=True)
    np.testing.assert_array_equal(rec["index"], idx)
    np.testing.assert_array_equal(rec["value"], val)
    np.testing.assert_array_equal(cnt, cnt_raw)


@pytest.mark.parametrize("n", [1024, 4096])
def test_one_bin_over_aligned_windows(n, gpu_device):
    """One bin over a window whose ends are multiples of 64 lags takes the
    wave-key selection (the wave's largest |z|^2, lowest lag first, lags
    outside the window counted as zero) instead of the per-sample scan: the
    peaks must equal the FP64 oracle's, with loud peaks planted at each
    window's first lag, just before it, at its last lag and just after it,
    with and without a threshold."""
    rng = np.random.default_rng(91 + n)
    nd, nt = 3, 21
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    f_ = np.arange(n)
    windows = [(n // 4, 3 * n // 4), (64, n - 64), (0, n // 2), (n // 2, n), (n - 64, n)]
    j = 0
    for ws, we in windows:
        for lag in (ws, ws - 1, we - 1, we):
            if 0 <= lag < n:
                # D_d += a T_t exp(-2 pi i f lag / n): z of (d, t) peaks at `lag`
                D[j % nd] += (3.0 * T[(5 * j) % nt] * np.exp(-2j * np.pi * f_ * lag / n)).astype(np.complex64)
                j += 1
    f = mf.MatchedFilter(n, nd, nt, device="gpu")
    f.set_data(D)
    f.set_templates(T)
    for ws, we in windows:
        for thr in (0.0, 3.0):
            got = f.run(binsize=n, threshold=thr, window=(ws, we))
            assert_peaks_match(D, T, got["index"], got["value"], n, thr, window=(ws, we),
                               context="n %d window %s threshold %s" % (n, (ws, we), thr))
