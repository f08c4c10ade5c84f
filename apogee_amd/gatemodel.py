"""Coarse gate for a false-dismissal budget, from the reference profile.

Restates the method of python/matchedfilter/gatemodel.py (``gate_for`` :210,
``_samples`` :63-132, ``_conditional`` :135-170, ``_nsamp_for`` :173-190) in
code of its own:

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

Differences from the reference: FP64 throughout (the reference draws in
FP32), and this module's own seeded stream.  The exact gate values therefore
differ within the Monte-Carlo error; they are pinned by invariants and by
measured dismissal in the tests, not by golden values.
"""
import numpy as np

_NB = 4      # coarse grid lags either side of the signal
_W = 3       # fine integer lags either side
_SEED = 13


def _nsamp_for(fd):
    """Draws so that ~100 land below the fd-quantile (at least 2e4, at most 3e6)."""
    fd = max(float(fd), 1e-6)
    return int(min(max(2.0e4, 200.0 / fd), 3.0e6))


def _validate(n, band, snr, fd):
    if (n < 64 or band < 1 or band > n or band & (band - 1) or n % band):
        raise ValueError("n must be divisible by a power-of-two band <= n")
    if not np.isfinite(snr) or snr <= 0:
        raise ValueError("snr must be finite and positive")
    if not np.isfinite(fd) or not 0 < fd < 1:
        raise ValueError("fd must be finite and between zero and one")


def _chol(cov):
    cov = (cov + cov.conj().T) / 2 + 1e-9 * np.eye(cov.shape[0])
    return np.linalg.cholesky(cov)


def samples(power, n, band, snr, nsamp, seed=_SEED):
    """Joint draws ``(coarse_max, fine_max)`` for a signal of strength snr."""
    p = np.asarray(power, dtype=np.float64)
    tot = p.sum()
    if not np.isfinite(tot) or tot <= 0:
        return None
    p = p / tot
    f = float(min(max(p[:band].sum(), 0.0), 1.0))
    if f == 0.0:
        return None
    step = n // band
    qb = np.zeros(n)
    qb[:band] = p[:band] / f
    A_full = np.fft.ifft(p) * n            # signal response / noise covariance, full band
    A_band = np.fft.ifft(qb) * n           # the same for the in-band part
    oob = (1.0 - f) > 1e-9
    A_out = (A_full - f * A_band) / (1.0 - f) if oob else None
    at = lambda tab, d: tab[np.asarray(d) % n]
    grid = np.arange(-_NB, _NB + 1) * step
    rng = np.random.default_rng(seed)
    per_off = max(nsamp // step, 256)
    coarse, fine = [], []
    for off in range(step):
        lags = np.unique(np.concatenate([np.arange(-_W, _W + 1) + off, grid]))
        gi = np.searchsorted(lags, grid)
        dl = lags[:, None] - lags[None, :]
        Lb = _chol(at(A_band, dl))
        w = (rng.standard_normal((per_off, lags.size))
             + 1j * rng.standard_normal((per_off, lags.size))) / np.sqrt(2.0)
        n_in = (w @ Lb.T) * np.sqrt(2.0)
        if oob:
            Lo = _chol(at(A_out, dl))
            w = (rng.standard_normal((per_off, lags.size))
                 + 1j * rng.standard_normal((per_off, lags.size))) / np.sqrt(2.0)
            n_out = (w @ Lo.T) * np.sqrt(2.0)
        else:
            n_out = 0.0
        s_full = at(A_full, lags - off)
        s_band = at(A_band, lags - off)
        fine.append(np.abs(snr * s_full[None, :] + np.sqrt(f) * n_in
                           + np.sqrt(1.0 - f) * n_out).max(axis=1))
        coarse.append(np.abs(snr * np.sqrt(f) * s_band[None, :] + n_in)[:, gi].max(axis=1))
    return np.concatenate(coarse), np.concatenate(fine)


def kept_coarse(power, n, band, snr, nsamp, seed=_SEED):
    """Sorted coarse maxima of the draws the fine stage keeps (fine >= snr)."""
    got = samples(power, n, band, snr, nsamp, seed)
    if got is None:
        return None
    c, fmax = got
    return np.sort(c[fmax >= snr])


def gate_for(power, n, band, snr, fd):
    """Largest coarse gate whose modelled dismissal meets ``fd`` (None if the
    draws cannot resolve the budget)."""
    _validate(n, band, snr, fd)
    nsamp = _nsamp_for(fd)
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
