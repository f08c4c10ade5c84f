"""FP32 accuracy of the device filter against an FP64 correlation, swept over
injected SNR and transform length.

Follows the reference's tests/test_precision.py and its sweep
(python/matchedfilter/precision.py:42-96, seed ``20240921 + n``): pure-noise
data with a unit-norm template injected by a phase ramp (an exact circular
shift) at a random lag, one bin over the whole transform, threshold 0.  Two
quantities are kept apart:

* the magnitude error at the lag the device reported, relative to the pair's
  FP64 maximum -- arithmetic only, bounded at 64 float32 epsilons;
* whether the reported lag is the FP64 argmax -- over noise plus an injection
  exact ties do not occur, so agreement must be complete.
"""
import numpy as np
import pytest

import apogee_amd as mf

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -24
SNRS = (0.0, 4.0, 5.0, 6.0, 8.0, 12.0, 20.0, 50.0, 200.0, 1000.0)
NS = (1024, 4096, 16384)
TRIALS, BATCH = 96, 32


def _batch(f, n, snr, rng):
    h = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    h /= np.linalg.norm(h)
    d = (rng.standard_normal((BATCH, n)) + 1j * rng.standard_normal((BATCH, n))).astype(np.complex64)
    lags = rng.integers(0, n, size=BATCH)
    if snr:
        k = np.arange(n)
        for j, lag in enumerate(lags):
            d[j] += (snr * h * np.exp(-2j * np.pi * lag * k / n)).astype(np.complex64)
    f.set_data(d)
    f.set_templates(h[None])
    pk = f.run(binsize=n, threshold=0.0)
    ref = np.abs(np.fft.ifft(d.astype(np.complex128) * np.conj(h.astype(np.complex128))[None]) * n)
    idx = pk["index"][:, 0, 0]
    mag = np.abs(pk["value"][:, 0, 0]).astype(np.float64)
    assert (idx >= 0).all()
    rows = np.arange(BATCH)
    scale = ref.max(axis=1)
    rel = np.abs(mag - ref[rows, idx]) / scale
    hits = int((idx == ref.argmax(axis=1)).sum())
    return rel, hits


@pytest.fixture(scope="module")
def rows(gpu_device):
    out = []
    for n in NS:
        rng = np.random.default_rng(20240921 + n)
        f = mf.MatchedFilter(n, BATCH, 1, device="gpu")
        for snr in SNRS:
            rel, hits = [], 0
            for _ in range(TRIALS // BATCH):
                r, h = _batch(f, n, snr, rng)
                rel.append(r)
                hits += h
            a = np.concatenate(rel)
            out.append(dict(n=n, snr=snr, rel_median=float(np.median(a)),
                            rel_max=float(a.max()), index_agreement=hits / TRIALS))
            print("n=%6d snr=%7.1f median %.2e max %.2e agree %.3f"
                  % (n, snr, out[-1]["rel_median"], out[-1]["rel_max"], out[-1]["index_agreement"]))
    return out


def test_magnitude_error_within_a_few_float32_ulps(rows):
    for r in rows:
        assert r["rel_max"] < 64 * EPS32, r


def test_error_does_not_grow_with_signal_strength(rows):
    for n in NS:
        at = {r["snr"]: r["rel_median"] for r in rows if r["n"] == n}
        assert at[max(at)] < 4 * max(at[min(at)], EPS32 / 4), (n, at)


def test_error_grows_no_faster_than_the_transform(rows):
    worst = {n: max(r["rel_max"] for r in rows if r["n"] == n) for n in NS}
    assert worst[NS[-1]] < 4 * worst[NS[0]], worst


def test_reported_lag_is_the_fp64_argmax(rows):
    for r in rows:
        assert r["index_agreement"] == 1.0, r


def test_pure_noise_is_as_accurate_as_a_loud_signal(rows):
    for n in NS:
        noise = next(r for r in rows if r["n"] == n and r["snr"] == 0.0)
        loud = max((r for r in rows if r["n"] == n), key=lambda r: r["snr"])
        assert noise["rel_max"] < 4 * loud["rel_max"], (noise, loud)
