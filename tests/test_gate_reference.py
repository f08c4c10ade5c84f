"""The coarse gate model against the gates the reference published.

tests/golden/gate_reference.json holds the reference's own gate values (its
GPU band sweep at n=4096 and its 1e-4 dismissal audit at n=1024; the fixture
script names the files).  Each gate is an order statistic of the model's
seeded Monte-Carlo, so matching them to float32 rounding pins the profile
normalisation, the correlation tables, the draw stream and order, the FP32
arithmetic and the quantile index all at once; a 2% gate shift fails by four
orders of magnitude.
"""
import json
import os

import numpy as np
import pytest

from apogee_amd import gatemodel as G

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "gate_reference.json")) as _fh:
    CASES = json.load(_fh)["cases"]


def test_api_inspiral(n, exponent=-7 / 3.0, knee_frac=0.0150):
    """The reference tests' |h|^2/S profile (tests/test_api.py:26-43)."""
    p = np.zeros(n, dtype=np.float32)
    k = np.arange(1, n // 2).astype(np.float64)
    p[1:n // 2] = (k ** exponent / ((knee_frac * n / k) ** 4 + 1.0)).astype(np.float32)
    return p / p.sum()


test_api_inspiral.__test__ = False          # a helper, not a test


def benchmark_inspiral(n, frac=0.895, fmax_frac=0.125, knee_frac=0.0150):
    """The reference benchmark's profile (python/matchedfilter/benchmark.py:336-383):
    knee-limited f^-7/3 with a fixed fraction `frac` of the power below n*fmax_frac."""
    k = np.arange(1, n // 2, dtype=np.float64)
    knee = max(2.0, knee_frac * n)
    p = k ** (-7.0 / 3.0) / ((knee / k) ** 4 + 1.0)
    cut = max(2, int(n * fmax_frac))
    lo, hi = p[:cut - 1], p[cut - 1:]
    if hi.size and hi.sum() > 0:
        hi *= (lo.sum() * (1 - frac) / frac) / hi.sum()
    out = np.zeros(n, np.float32)
    out[1:n // 2] = (p / p.sum()).astype(np.float32)
    return out


PROFILES = {"test_api_inspiral": test_api_inspiral, "benchmark_inspiral": benchmark_inspiral}


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-n%d-b%d-fd%g" % (c["profile"], c["n"], c["band"], c["fd"]))
def test_gate_matches_reference(case):
    p = PROFILES[case["profile"]](case["n"])
    g = G.gate_for(p, case["n"], case["band"], case["snr"], case["fd"])
    assert g is not None
    assert g == pytest.approx(case["gate"], rel=1e-6), (g, case)


def test_fixture_discriminates_profiles():
    """A 2% change of the profile's exponent moves the gate far outside the
    tolerance above, so the match is not an artefact of an insensitive model."""
    case = next(c for c in CASES if c["band"] == 512 and c["fd"] == 1e-2)
    p = test_api_inspiral(4096, exponent=-7 / 3.0 * 1.02)
    g = G.gate_for(p, 4096, 512, 5.5, 1e-2)
    assert abs(g / case["gate"] - 1) > 1e-4
