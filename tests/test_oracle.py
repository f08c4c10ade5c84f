"""Pin the FP64 oracle and the FP32 CPU restatement before trusting them.

Every check here fixes something independently of the code under test: the
O(n^2) definition, closed forms, a brute-force loop, golden fixtures written
from the definition (tests/golden/make_golden.py), finite differences.  Each
is chosen so a plausible mistake -- a conjugate on the wrong side, a sign of
the exponent, an off-by-one bin edge, a non-strict threshold -- fails it.
Mirrors the reference's tests/test_independent.py and tests/test_adversarial.py
ideas, not their code.
"""
import os

import numpy as np
import pytest

from oracle import mf_oracle as O
from oracle import cpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["config1_small", "ragged", "planted", "closed_forms"]


def golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


# ------------------------------------------------------------ definition

@pytest.mark.parametrize("n", [8, 64, 256])
def test_fft_correlation_equals_the_direct_sum(n):
    rng = np.random.default_rng(n)
    D, T = cnoise(rng, (2, n)), cnoise(rng, (3, n))
    a = O.correlate(D[:, None], T[None])
    b = O.direct_correlate(D[:, None], T[None])
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def test_sign_convention_a_shifted_template_peaks_at_its_lag():
    # template = D * exp(+2 pi i f lag / n) peaks at `lag` (test_api.py:84-100)
    n = 256
    rng = np.random.default_rng(1)
    x = cnoise(rng, n)
    f = np.arange(n)
    for lag in (0, 1, 37, n // 2, n - 1):
        T = x * np.exp(2j * np.pi * f * lag / n)
        idx, _, _ = O.flat_peaks(x[None], T[None], n, 0.0)
        assert idx[0, 0, 0] == lag


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden_fixtures(name):
    g = golden(name)
    idx, val, best = O.flat_peaks(g["D"], g["T"], int(g["binsize"]), float(g["threshold"]),
                                  (int(g["ws"]), int(g["we"])))
    scale = np.sqrt(g["best"]).max()
    live = g["best"] > (1e-9 * scale) ** 2          # exact zeros are FP64 noise
    np.testing.assert_array_equal(idx[live], g["idx"][live])
    assert np.abs(val - g["val"]).max() <= 1e-9 * scale


def test_closed_forms():
    g = golden("closed_forms")
    h = g["D"][0].astype(np.complex128)
    n = h.size
    # self-correlation: lag 0, value sum |h|^2 (real)
    assert g["idx"][0, 0, 0] == 0
    assert abs(g["val"][0, 0, 0] - np.sum(np.abs(h) ** 2)) <= 1e-9 * n * n
    # flat spectrum: a delta of height n at lag 0
    assert abs(g["val"][1, 1, 0] - n) <= 1e-9 * n
    assert np.abs(g["val"][1, 1, 1:]).max() <= 1e-9 * n
    # and the oracle agrees on both
    _, val, _ = O.flat_peaks(g["D"], g["T"], 1, 0.0)
    assert abs(val[0, 0, 0] - np.sum(np.abs(h) ** 2)) <= 1e-9 * n * n
    assert abs(val[1, 1, 0] - n) <= 1e-9 * n


# -------------------------------------------------------- selection rule

def test_threshold_is_strict_and_squared():
    z = np.array([0, 3, 4j, 0], dtype=np.complex128)
    # |z| = 4 at lag 2: threshold 4 dismisses, threshold just below reports
    idx, val, _ = O.select(z[None], 4, 4.0, 0, 4)
    assert idx[0, 0] == -1 and val[0, 0] == 0
    idx, val, _ = O.select(z[None], 4, np.nextafter(4.0, 0), 0, 4)
    assert idx[0, 0] == 2 and val[0, 0] == 4j


def test_all_zero_bin_reports_nothing_and_ties_go_to_the_lowest_lag():
    z = np.zeros(8, np.complex128)
    idx, _, _ = O.select(z[None], 8, 0.0, 0, 8)
    assert idx[0, 0] == -1
    z = np.array([1, 2, 2j, -2, 0, 0, 1, 1], np.complex128)
    idx, val, _ = O.select(z[None], 4, 0.0, 0, 8)
    assert list(idx[0]) == [1, 6]                       # equal |z| = 2 at 1, 2, 3
    assert val[0, 0] == 2


def test_nan_never_wins():
    z = np.array([np.nan, 1.0, 0.5], np.complex128)
    idx, _, _ = O.select(z[None], 3, 0.0, 0, 3)
    assert idx[0, 0] == 1


def test_bins_windows_and_ragged_edges_by_brute_force():
    rng = np.random.default_rng(5)
    z = cnoise(rng, 100).astype(np.complex128)
    for binsize, ws, we in [(7, 0, 100), (30, 13, 97), (1, 50, 60), (200, 3, 9)]:
        idx, _, _ = O.select(z[None], binsize, 0.0, ws, we)
        for b in range(idx.shape[-1]):
            lo, hi = ws + b * binsize, min(we, ws + (b + 1) * binsize)
            assert idx[0, b] == lo + int(np.argmax(np.abs(z[lo:hi])))


def test_window_and_nbins_rules():
    assert O.window(100, (-5, 500)) == (0, 100)
    with pytest.raises(ValueError):
        O.window(100, (50, 50))
    assert O.nbins(4096, 512) == 8
    assert O.nbins(4096, 384) == 11
    assert O.nbins(100, 7, 10, 100) == 13
    with pytest.raises(ValueError):
        O.nbins(100, 0)


def test_flop_model():
    # tools/gpu_roofline.py:249-260 of the reference: 282,624 at 4096
    assert O.flops_per_pair(4096) == 282624
    assert O.flops_per_pair(65536) == 5832704


# ------------------------------------------------------- hierarchical

def test_coarse_pass_is_the_band_limited_correlation_on_the_decimated_grid():
    """z_coarse[j] * sqrt(f) == (band-limited full correlation)[j * n/m]."""
    rng = np.random.default_rng(7)
    n, m = 256, 64
    D, T = cnoise(rng, n), cnoise(rng, n)
    ct = O.coarse_templates(T[None], m)[0]
    f = O.band_fraction(T[None], m)[0]
    zc = O.correlate(D[:m], ct)
    band = np.zeros(n, np.complex128)
    band[:m] = D[:m].astype(np.complex128) * np.conj(T[:m].astype(np.complex128))
    zfull = np.fft.ifft(band) * n
    np.testing.assert_allclose(zc * np.sqrt(f), zfull[:: n // m], rtol=1e-5, atol=1e-4)


def test_coarse_window_formula():
    # hmf.c:233-239: widen by one coarse sample on the left, ceil on the right
    assert O.coarse_window(4096, 512, 0, 4096) == (0, 512)
    assert O.coarse_window(4096, 512, 37, 4000) == (3, 500)
    assert O.coarse_window(4096, 512, 500, 501) == (61, 63)
    assert O.coarse_window(4096, 512, 8, 9) == (0, 2)


def test_hierarchical_is_identical_or_absent():
    rng = np.random.default_rng(11)
    n, nd, nt, m = 1024, 4, 6, 128
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    D[1] += 6 * T[2]                                    # one loud pair
    fi, fv, _ = O.flat_peaks(D, T, 256, 40.0)
    for thr in (0.0, 30.0, 60.0, 1e9):
        hi, hv, _, fired = O.hier_peaks(D, T, m, thr, 256, 40.0)
        assert not ((hi >= 0) & (fi < 0)).any()
        np.testing.assert_array_equal(hi[hi >= 0], fi[hi >= 0])
        np.testing.assert_array_equal(hv[hi >= 0], fv[hi >= 0])
        assert (hi[~fired] == -1).all()
    _, _, _, fired = O.hier_peaks(D, T, m, 0.0, 256, 40.0)
    assert fired.all()                                  # gate 0 escalates everything
    _, _, _, fired = O.hier_peaks(D, T, m, 1e9, 256, 40.0)
    assert not fired.any()


def test_band_fraction_with_and_without_reference():
    T = np.zeros((2, 16), np.complex64)
    T[0, :4] = 1
    T[0, 4:] = 1
    T[1, :] = 0
    f = O.band_fraction(T, 4)
    assert f[0] == pytest.approx(0.25) and f[1] == 0.0
    ref = np.arange(16, dtype=np.float64)
    assert O.band_fraction(T, 4, ref) == pytest.approx(6 / 120)
    ct = O.coarse_templates(T, 4)
    np.testing.assert_allclose(ct[0], np.full(4, 2.0, np.complex64))
    assert (ct[1] == 0).all()


# ---------------------------------------------------------- run_series

def test_series_spectra_are_block_ffts_over_n_with_zero_padding():
    rng = np.random.default_rng(3)
    n = 64
    s = cnoise(rng, 150)
    spec = O.series_spectra(s, [0, 32, 120, 200], n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    for b, st in enumerate([0, 32, 120, 200]):
        blk = np.zeros(n, np.complex128)
        seg = s[st:st + n]
        blk[:seg.size] = seg
        np.testing.assert_allclose(spec[b], (blk / n) @ W, atol=1e-12)
    assert np.abs(spec[3]).max() == 0.0


# ------------------------------------------------------------ gradient

def test_peak_value_gradient_against_finite_differences():
    rng = np.random.default_rng(9)
    n, k = 32, 5
    D, T = cnoise(rng, n).astype(np.complex128), cnoise(rng, n).astype(np.complex128)
    a, b = O.peak_value_grad(D, T, k)
    h = 1e-6
    for f in (0, 3, 17):
        e = np.zeros(n)
        e[f] = h
        dz = (O.correlate(D + e, T)[k] - O.correlate(D - e, T)[k]) / (2 * h)
        assert abs(dz - a[f]) < 1e-6 * abs(a[f]) + 1e-9
        # z depends on T through conj(T): d z / d Re T[f] = b_f, d z / d Im T[f] = -i b_f
        dz = (O.correlate(D, T + e)[k] - O.correlate(D, T - e)[k]) / (2 * h)
        assert abs(dz - b[f]) < 1e-6 * abs(b[f]) + 1e-9
        dz = (O.correlate(D, T + 1j * e)[k] - O.correlate(D, T - 1j * e)[k]) / (2 * h)
        assert abs(dz - (-1j) * b[f]) < 1e-6 * abs(b[f]) + 1e-9


def _peak_loss(D, T, idx, c):
    """A real loss of the peak values: sum Re(c z) + |z|^2 at the stored lags."""
    n = D.shape[-1]
    f = np.arange(n)
    L = 0.0
    for d, t, b in zip(*np.nonzero(idx >= 0)):
        z = np.sum(D[d] * np.conj(T[t]) * np.exp(2j * np.pi * f * idx[d, t, b] / n))
        L += (c[d, t, b] * z).real + abs(z) ** 2
    return L


def _loss_grad(D, T, idx, c):
    """dL/dRe z + i dL/dIm z of _peak_loss: conj(c) + 2 z."""
    z = np.zeros(idx.shape, np.complex128)
    n = D.shape[-1]
    f = np.arange(n)
    for d, t, b in zip(*np.nonzero(idx >= 0)):
        z[d, t, b] = np.sum(D[d] * np.conj(T[t]) * np.exp(2j * np.pi * f * idx[d, t, b] / n))
    return np.where(idx >= 0, np.conj(c) + 2 * z, 0)


def test_peak_grads_against_finite_differences():
    """Real and imaginary parts of every input, central differences in FP64."""
    rng = np.random.default_rng(21)
    nd, nt, n, nb = 2, 3, 16, 2
    D = cnoise(rng, (nd, n)).astype(np.complex128)
    T = cnoise(rng, (nt, n)).astype(np.complex128)
    idx = rng.integers(0, n, size=(nd, nt, nb))
    idx[0, 1, 1] = -1                      # an empty bin contributes nothing
    c = cnoise(rng, (nd, nt, nb)).astype(np.complex128)
    gD, gT = O.peak_grads(D, T, idx, _loss_grad(D, T, idx, c))
    h = 1e-6
    for X, G, which in ((D, gD, "D"), (T, gT, "T")):
        for r in range(X.shape[0]):
            for f in (0, 5, 15):
                for unit in (1.0, 1j):
                    Xp, Xm = X.copy(), X.copy()
                    Xp[r, f] += h * unit
                    Xm[r, f] -= h * unit
                    args_p = (Xp, T) if which == "D" else (D, Xp)
                    args_m = (Xm, T) if which == "D" else (D, Xm)
                    fd = (_peak_loss(*args_p, idx, c) - _peak_loss(*args_m, idx, c)) / (2 * h)
                    got = G[r, f].real if unit == 1.0 else G[r, f].imag
                    assert abs(fd - got) <= 1e-6 * (abs(got) + 1.0), (which, r, f, unit, fd, got)


def test_peak_grads_against_torch_autograd():
    """The same gradients from PyTorch's complex autograd through n*ifft."""
    import torch
    rng = np.random.default_rng(22)
    nd, nt, n, nb = 3, 2, 64, 4
    D = cnoise(rng, (nd, n)).astype(np.complex128)
    T = cnoise(rng, (nt, n)).astype(np.complex128)
    idx = rng.integers(0, n, size=(nd, nt, nb))
    idx[2, 0, :2] = -1
    c = cnoise(rng, (nd, nt, nb)).astype(np.complex128)
    Dt = torch.tensor(D, requires_grad=True)
    Tt = torch.tensor(T, requires_grad=True)
    z = n * torch.fft.ifft(Dt[:, None, :] * torch.conj(Tt)[None, :, :], dim=-1)
    it = torch.tensor(np.where(idx >= 0, idx, 0))
    zk = torch.gather(z, 2, it)
    m = torch.tensor(idx >= 0)
    ct = torch.tensor(c)
    L = (torch.where(m, (ct * zk).real + zk.abs() ** 2, torch.zeros((), dtype=torch.float64))).sum()
    L.backward()
    gD, gT = O.peak_grads(D, T, idx, _loss_grad(D, T, idx, c))
    np.testing.assert_allclose(gD, Dt.grad.resolve_conj().numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(gT, Tt.grad.resolve_conj().numpy(), rtol=1e-10, atol=1e-10)


# --------------------------------------------- the FP32 CPU restatement

@pytest.mark.parametrize("name", CASES)
def test_cpu_restatement_matches_golden(name, mfcpu):
    from _parity import assert_peaks_match
    g = golden(name)
    idx, val, _ = cpu.flat(g["D"], g["T"], int(g["binsize"]), float(g["threshold"]),
                           (int(g["ws"]), int(g["we"])))
    assert_peaks_match(g["D"], g["T"], idx, val, int(g["binsize"]), float(g["threshold"]),
                       (int(g["ws"]), int(g["we"])), context=name)


@pytest.mark.parametrize("n", [2, 4, 8, 64, 128, 512, 1024, 4096, 16384])
def test_cpu_restatement_every_length(n, mfcpu):
    from _parity import assert_peaks_match
    rng = np.random.default_rng(n)
    D, T = cnoise(rng, (2, n)), cnoise(rng, (3, n))
    bs = max(1, n // 8)
    idx, val, cnt = cpu.flat(D, T, bs, 0.0)
    assert_peaks_match(D, T, idx, val, bs, 0.0, context="n=%d" % n)
    np.testing.assert_array_equal(cnt, (idx >= 0).sum(axis=2))


def test_cpu_restatement_block_spectrum(mfcpu):
    rng = np.random.default_rng(4)
    n = 256
    s = cnoise(rng, 600)
    for st in (0, 100, 500, 700):
        got = cpu.block_spectrum(s, st, n)
        want = O.series_spectra(s, [st], n)[0]
        assert np.abs(got - want).max() <= 1e-6 * max(np.abs(want).max(), 1e-30)
