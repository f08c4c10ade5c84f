"""Write the golden fixtures of tests/golden/*.npz.

Every expected output here comes from the correlation's DEFINITION -- a
direct O(n^2) sum in float64, z[k] = sum_f D[f] conj(T[f]) exp(+2 pi i f k/n)
-- and the selection rule written out as a plain loop (strict '>' against
threshold^2, best |z|^2 per bin, lowest lag among ties, -1 / 0 for empty
bins).  No FFT and no code of the oracle is used, so the fixtures pin the
oracle and the device path independently.

Inputs follow the synthetic recipes of the benchmark configurations, at
sizes the direct sum finishes in seconds:
  * ``config1_small``: templates = FFT of seeded standard-normal real series;
    data d = 0.5 * roll(template[d mod nt], shift) + N(0, 1) real noise, FFT'd
    (the CPU-parity configuration, scaled down);
  * ``ragged``: complex Gaussian spectra, a window that is not a multiple of
    the non-power-of-two binsize, and a threshold that empties some bins;
  * ``planted``: a template and data holding it at known lags (phase ramps),
    with one lag on a bin edge;
  * ``closed_forms``: self-correlation (peak n*energy... at lag 0 with value
    sum |h|^2) and a flat spectrum (a delta of height n at lag 0).

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def direct(D, T):
    n = D.shape[-1]
    k = np.arange(n)
    w = np.exp(2j * np.pi * np.outer(k, k) / n)
    return (D.astype(np.complex128) * np.conj(T.astype(np.complex128))) @ w


def select_loop(z, binsize, threshold, ws, we):
    nb = -(-(we - ws) // binsize)
    seed = threshold * threshold if threshold > 0 else 0.0
    idx = np.full(nb, -1, np.int64)
    val = np.zeros(nb, np.complex128)
    best = np.full(nb, seed)
    for k in range(ws, we):
        b = (k - ws) // binsize
        m2 = z[k].real ** 2 + z[k].imag ** 2
        if m2 > best[b]:
            best[b] = m2
            idx[b] = k
            val[b] = z[k]
    best[idx < 0] = 0.0
    return idx, val, best


def expected(D, T, binsize, threshold, ws, we):
    nd, nt = D.shape[0], T.shape[0]
    nb = -(-(we - ws) // binsize)
    idx = np.empty((nd, nt, nb), np.int64)
    val = np.empty((nd, nt, nb), np.complex128)
    best = np.empty((nd, nt, nb))
    for d in range(nd):
        for t in range(nt):
            z = direct(D[d], T[t])
            idx[d, t], val[d, t], best[d, t] = select_loop(z, binsize, threshold, ws, we)
    return idx, val, best


def save(name, D, T, binsize, threshold, ws, we):
    idx, val, best = expected(D, T, binsize, threshold, ws, we)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), D=D, T=T,
                        binsize=binsize, threshold=threshold, ws=ws, we=we,
                        idx=idx, val=val, best=best)
    print(name, D.shape, T.shape, "bins", idx.shape[-1],
          "reported", int((idx >= 0).sum()), "of", idx.size)


def main():
    # config-1 recipe at n=256 (16 data x 8 templates, 8 bins, no threshold)
    rng = np.random.default_rng(20261202 + 1)
    n, nd, nt = 256, 16, 8
    t_time = rng.standard_normal((nt, n))
    T = np.fft.fft(t_time).astype(np.complex64)
    x = np.empty((nd, n))
    for d in range(nd):
        s = int(rng.integers(0, n))
        x[d] = 0.5 * np.roll(t_time[d % nt], s) + rng.standard_normal(n)
    D = np.fft.fft(x).astype(np.complex64)
    save("config1_small", D, T, 32, 0.0, 0, n)

    # ragged bins, offset window, threshold that empties some bins
    rng = np.random.default_rng(20261202 + 11)
    n, nd, nt = 128, 3, 5
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    save("ragged", D, T, 37, 30.0, 5, 120)

    # planted lags, one on a bin edge; amplitude 10 over unit-norm templates
    rng = np.random.default_rng(20261202 + 12)
    n, nt = 512, 2
    h = rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    lags = [0, 63, 64, 511]
    f = np.arange(n)
    D = np.stack([10.0 * h[i % nt] * np.exp(-2j * np.pi * lag * f / n)
                  for i, lag in enumerate(lags)]).astype(np.complex64)
    save("planted", D, h.astype(np.complex64), 64, 1.0, 0, n)

    # closed forms: self-correlation and a flat spectrum (binsize 1: every lag)
    rng = np.random.default_rng(20261202 + 13)
    n = 64
    h = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    D = np.stack([h, np.ones(n, np.complex64)])
    T = np.stack([h, np.ones(n, np.complex64)])
    save("closed_forms", D, T, 1, 0.0, 0, n)


if __name__ == "__main__":
    main()
