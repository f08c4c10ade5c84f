"""Coarse gate for a false-dismissal budget, from the reference profile.

Method of python/matchedfilter/gatemodel.py (``gate_for`` :210-227,
``_samples`` :63-132, ``_conditional`` :135-170, ``_nsamp_for`` :173-190):

For a normalised output-power profile q (the reference), the filter output at
lag tau for a unit-SNR signal at lag 0 is A(tau) = sum_k q_k exp(2 pi i k tau /
n), and the output noise covariance between two lags is A(t1 - t2): the noise
correlation IS the signal response.  The coarse stage sees the in-band part
(fraction f of the profile, lags on the decimated grid, normalised by
1/sqrt(f)) and shares its in-band noise with the fine stage.  Drawing that
joint Gaussian noise at a handful of lags around a signal at every sub-step
offset gives samples of (coarse maximum, fine maximum); among the draws the
fine stage keeps (fine >= snr), the fd-quantile of the coarse maximum is the
largest gate that dismisses at most a fraction fd of them.

Noise convention: Re and Im of the output each have unit variance (|z|^2 is
chi-square with 2 degrees of freedom), so SNR 5 means |z| = 5.

The gate is an order statistic of a seeded Monte-Carlo, so the same gate
needs the same random stream and the same arithmetic: FP64 profile sums and
correlation tables, a Cholesky factor rounded to complex64, FP32 normal draws
from ``numpy.random.default_rng(13)`` in the reference's order (in-band block,
then out-of-band block, per offset), and FP32 magnitudes.  The reference's
published gates (its GPU band sweep at n=4096 and its 1e-4 dismissal audit at
n=1024) are reproduced to float32 rounding by the tests.
"""
import numpy as np

_NB = 4             # coarse-grid lags either side of the signal
_W = 3              # fine integer lags either side of the signal offset
_SEED = 13
_ROOT2 = np.sqrt(2.0)
_MEMO = {}          # (profile bytes, n, band, snr, nsamp, seed) -> sorted kept coarse maxima
_MEMO_MAX = 64


def _nsamp_for(fd):
    """Draws so that ~100 land below the fd-quantile (at least 2e4, at most 3e6)."""
    fd = max(float(fd), 1e-6)
    return int(min(max(2.0e4, 200.0 / fd), 3.0e6))


def _validate(n, band, snr, fd):
    if (not isinstance(n, (int, np.integer)) or n < 64
            or not isinstance(band, (int, np.integer)) or band < 1
            or band > n or band & (band - 1) or n % band):
        raise ValueError("n must be divisible by a power-of-two band <= n")
    if not np.isfinite(snr) or snr <= 0:
        raise ValueError("snr must be finite and positive")
    if not np.isfinite(fd) or not 0 < fd < 1:
        raise ValueError("fd must be finite and between zero and one")


class _Tables:
    """Signal-response / noise-covariance tables of one (profile, band)."""

    def __init__(self, prof, n, band):
        self.n = n
        self.f = min(max(float(prof[:band].sum()), 0.0), 1.0)
        self.step = n // band
        inband = np.zeros(n, np.float64)
        if self.f > 0.0:
            inband[:band] = prof[:band] / self.f
        self.full = np.fft.ifft(prof) * n
        self.band = np.fft.ifft(inband) * n
        self.has_out = (1.0 - self.f) > 1e-9
        self.out = (self.full - self.f * self.band) / (1.0 - self.f) if self.has_out else None

    def at(self, table, lags):
        return table[np.asarray(lags) % self.n]

    def factor(self, table, lags):
        """complex64 Cholesky factor of the (Hermitian-symmetrised, jittered)
        covariance between the given lags."""
        cov = self.at(table, lags[:, None] - lags[None, :])
        cov = (cov + cov.conj().T) / 2 + 1e-9 * np.eye(lags.size)
        return np.linalg.cholesky(cov).astype(np.complex64)


def _noise(rng, rows, factor):
    """Unit-variance-per-component correlated noise: FP32 normal draws (real
    block, then imaginary block) coloured by the complex64 factor."""
    k = factor.shape[0]
    re = rng.standard_normal((rows, k), dtype=np.float32)
    im = rng.standard_normal((rows, k), dtype=np.float32)
    white = (re + 1j * im) / np.float32(_ROOT2)
    return (white @ factor.T) * np.float32(_ROOT2)


def samples(power, n, band, snr, nsamp, seed=_SEED):
    """Joint draws ``(coarse_max, fine_max)`` (float32) for a signal of
    strength snr at every sub-step offset, or None for an empty band."""
    prof = np.asarray(power, dtype=np.float64)
    total = prof.sum()
    if total <= 0:
        return None
    prof = prof / total
    tab = _Tables(prof, n, band)
    if tab.f == 0.0:
        return None
    grid = np.arange(-_NB, _NB + 1) * tab.step
    rows = max(nsamp // tab.step, 256)
    rng = np.random.default_rng(seed)
    g_in = np.float32(np.sqrt(tab.f))
    g_out = np.float32(np.sqrt(1.0 - tab.f))
    s_fine = np.float32(snr)
    s_coarse = np.float32(snr * np.sqrt(tab.f))
    coarse, fine = [], []
    for off in range(tab.step):
        lags = np.unique(np.concatenate([np.arange(-_W, _W + 1) + off, grid]))
        on_grid = np.searchsorted(lags, grid)
        n_in = _noise(rng, rows, tab.factor(tab.band, lags))
        n_out = (_noise(rng, rows, tab.factor(tab.out, lags)) if tab.has_out
                 else np.float32(0.0))
        sig_full = tab.at(tab.full, lags - off).astype(np.complex64)
        sig_band = tab.at(tab.band, lags - off).astype(np.complex64)
        fine.append(np.abs(s_fine * sig_full[None, :] + g_in * n_in + g_out * n_out).max(1))
        coarse.append(np.abs(s_coarse * sig_band[None, :] + n_in)[:, on_grid].max(1))
    return np.concatenate(coarse), np.concatenate(fine)


def kept_coarse(power, n, band, snr, nsamp, seed=_SEED):
    """Sorted coarse maxima of the draws the fine stage keeps (fine >= snr)."""
    p = np.asarray(power, dtype=np.float64)
    if (p.shape != (n,) or not np.isfinite(p).all() or (p < 0).any()
            or not np.isfinite(p.sum()) or p.sum() <= 0):
        raise ValueError("power must be a finite nonnegative length-n profile with positive sum")
    p = p / p.sum()
    key = (p.tobytes(), int(n), int(band), float(snr), int(nsamp), int(seed))
    if key in _MEMO:
        return _MEMO[key]
    got = samples(p, n, band, snr, nsamp, seed)
    kept = None
    if got is not None:
        c, fmax = got
        kept = np.sort(c[fmax >= snr])
    if len(_MEMO) >= _MEMO_MAX:
        _MEMO.pop(next(iter(_MEMO)))
    _MEMO[key] = kept
    return kept


def gate_for(power, n, band, snr, fd):
    """Largest coarse gate whose modelled dismissal meets ``fd`` (None if the
    draws cannot resolve the budget)."""
    _validate(n, band, snr, fd)
    nsamp = _nsamp_for(fd)
    if fd * nsamp < 8:
        return None
    kept = kept_coarse(power, n, band, snr, nsamp)
    if kept is None or not len(kept):
        return None
    i = int(np.floor(float(fd) * len(kept)))
    if i < 8:
        return None
    return float(kept[i])


def dismissal(power, n, band, snr, gate, fd_hint=1e-3):
    """Modelled false-dismissal rate at ``gate``."""
    _validate(n, band, snr, fd_hint)
    if not np.isfinite(gate) or gate < 0:
        raise ValueError("gate must be finite and nonnegative")
    kept = kept_coarse(power, n, band, snr, _nsamp_for(fd_hint))
    if kept is None or not len(kept):
        return None
    return float(np.searchsorted(kept, float(gate)) / len(kept))


def noise_exceedance(power, n, band, gate):
    """Modelled fraction of noise-only pairs whose coarse maximum reaches
    ``gate``: the refine rate the gate costs on pure noise.

    The coarse series has band lags whose noise is correlated through the
    in-band profile; it behaves like ``m`` independent lags, with m the
    profile's effective number of in-band bins B = (sum q)^2 / sum q^2 (at
    most band), each with |z|^2 exponential of mean 2 (the unit-variance
    convention above): P = 1 - (1 - exp(-gate^2 / 2))^m.  An approximation
    for ranking bands, not a calibration: the gate itself comes from
    ``gate_for``.
    """
    _validate(n, band, 1.0, 0.5)
    q = np.asarray(power, dtype=np.float64)[:band]
    q = np.where(q > 0, q, 0.0)
    if q.sum() <= 0:
        return 0.0
    m = min(float(band), max(1.0, q.sum() ** 2 / np.square(q).sum()))
    tail = np.exp(-0.5 * float(gate) ** 2)
    return float(-np.expm1(m * np.log1p(-tail))) if tail < 1 else 1.0
