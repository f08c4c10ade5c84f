"""Matrix-core coarse screen (gfx950 MFMA).

Pins (1) the fragment layouts of the four MFMA instructions the screen uses,
exactly, with small-integer matrices that fp16 and bf16 represent exactly;
(2) the coarse statistic against the FP32 screen, whose relative error bounds
the gate widening; (3) that the widened fp16/bf16 screen keeps every pair the
FP32 screen keeps, and the hierarchical one-sided guarantee on top of it.
"""
import numpy as np
import pytest

import apogee_amd as mf
from apogee_amd import _hip
from apogee_amd.hierarchical import _COARSE_PRECISIONS
from oracle import mf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant,k", [(0, 32), (1, 32), (2, 16), (3, 16)])
def test_mfma_fragment_layouts(variant, k, gpu_device):
    rng = np.random.default_rng(variant)
    A = rng.integers(-8, 9, (16, k)).astype(np.float32)
    B = rng.integers(-8, 9, (k, 16)).astype(np.float32)
    C = _hip.selftest_mfma(0, variant, A, B)
    np.testing.assert_array_equal(C, A @ B)


def inspiral_power(n, exponent=-7 / 3.0, knee_frac=0.0150):
    p = np.zeros(n, dtype=np.float32)
    kk = np.arange(1, n // 2).astype(np.float64)
    p[1:n // 2] = (kk ** exponent / ((knee_frac * n / kk) ** 4 + 1.0)).astype(np.float32)
    return p / p.sum()


def bank(kind, rng, n, nt):
    if kind == "noise":
        return (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    power = inspiral_power(n)
    H = np.sqrt(power)[None, :] * np.exp(2j * np.pi * rng.random((nt, n)))
    return (H / np.linalg.norm(H, axis=1, keepdims=True)).astype(np.complex64)


@pytest.mark.parametrize("kind", ["noise", "inspiral"])
@pytest.mark.parametrize("window", [None, (37, 4000), (2048, 4096)])
def test_coarse_statistic_error_is_within_the_widening(kind, window, gpu_device):
    n, nd, nt = 4096, 16, 64
    rng = np.random.default_rng(7)
    T = bank(kind, rng, n, nt)
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    for d in range(0, nd, 3):                    # loud pairs as well as noise pairs
        D[d] += (20 * np.linalg.norm(D[d]) / np.sqrt(n) * T[d % nt] * np.sqrt(n)).astype(np.complex64)
    h = mf.HierarchicalFilter(n, nd, nt, band=512, device="gpu")
    h.set_templates(T)
    h.set_data(D)
    ref = h.coarse_statistic("fp32", window=window).astype(np.float64)
    # the FP32 screen itself against the FP64 oracle
    start, end = O.window(n, window)
    cs, ce = O.coarse_window(n, 512, start, end)
    ct = O.coarse_templates(T, 512)
    z = O.correlate(D[:, None, :512], ct[None])
    exact = np.abs(z[..., cs:ce]).max(axis=-1)
    assert np.abs(ref - exact).max() <= 2e-6 * 9 * exact.max()
    for prec in ("fp16", "bf16"):
        got = h.coarse_statistic(prec, window=window).astype(np.float64)
        rel = np.abs(got - ref) / ref
        delta = _COARSE_PRECISIONS[prec][1]
        print("%s %s window=%s: max rel err %.3g (median %.3g), widening %.3g"
              % (kind, prec, window, rel.max(), np.median(rel), delta))
        assert rel.max() <= delta / 2


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_mfma_screen_keeps_the_fp32_survivors(prec, gpu_device):
    n, nd, nt = 4096, 16, 256
    rng = np.random.default_rng(21)
    power = inspiral_power(n)
    T = bank("inspiral", rng, n, nt)
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    for d in range(nd):
        D[d] += (6.0 * T[(7 * d) % nt] * np.exp(2j * np.pi * np.arange(n) * (100 + 37 * d) / n)
                 ).astype(np.complex64)
    flat = mf.MatchedFilter(n, nd, nt, device="gpu")
    h32 = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2, band=512, device="gpu")
    hm = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2, band=512, device="gpu")
    hm.set_coarse_precision(prec)
    for o in (flat, h32, hm):
        o.set_data(D)
        o.set_templates(T)
    for o in (h32, hm):
        o.set_reference(power)
    a = flat.run(binsize=1024, threshold=5.0)
    b32 = h32.run(binsize=1024, threshold=5.0)
    bm = hm.run(binsize=1024, threshold=5.0)
    # one-sided against flat
    fired = bm["index"] >= 0
    assert not (fired & (a["index"] < 0)).any()
    np.testing.assert_array_equal(a["index"][fired], bm["index"][fired])
    np.testing.assert_array_equal(a["value"][fired], bm["value"][fired])
    # superset of the FP32 screen's reports
    assert not ((b32["index"] >= 0) & (bm["index"] < 0)).any()
    assert hm.stats[1] >= h32.stats[1]
    assert hm.refine_rate < 0.8


@pytest.mark.parametrize("lds", [0, 1])
@pytest.mark.parametrize("rpw,tpw", [(1, 1), (1, 3), (1, 8), (2, 2), (2, 5), (2, 16), (1, 64), (2, 64)])
def test_every_coarse_launch_shape_stays_within_the_widening(rpw, tpw, lds, gpu_device):
    """The rows and templates per wave and where the DFT constants live
    (autotuned) change only the work split: every shape keeps the statistic
    within half the widening of the FP32 screen, including ragged edges
    (nd = 7 rows, nt = 101 templates) and a sub-range of templates; the two
    constant placements agree to well within the widening."""
    n, nd, nt = 4096, 7, 101
    rng = np.random.default_rng(40 + rpw * 100 + tpw)
    T = bank("inspiral", rng, n, nt)
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    D[3] += (15 * T[9] * np.sqrt(n) * np.linalg.norm(D[3]) / np.sqrt(n)).astype(np.complex64)
    h = mf.HierarchicalFilter(n, nd, nt, band=512, device="gpu")
    h.set_templates(T)
    h.set_data(D)
    h._gpu.set_tuning("coarse_lds", lds)
    h._gpu.set_tuning("coarse_rpw", rpw)
    h._gpu.set_tuning("coarse_tpw", tpw)
    for rng_t in (None, (13, 77)):
        ref = h.coarse_statistic("fp32", templates=rng_t).astype(np.float64)
        for prec in ("fp16", "bf16"):
            got = h.coarse_statistic(prec, templates=rng_t).astype(np.float64)
            assert got.shape == ref.shape
            rel = np.abs(got - ref) / ref
            assert rel.max() <= _COARSE_PRECISIONS[prec][1] / 2, (prec, rng_t, rel.max())
            if lds:
                # the compiler may fuse the twiddle into the fp16 conversion
                # differently in the two placements (v_fma_mix or not): the
                # statistics agree to within a few fp16 roundings
                h._gpu.set_tuning("coarse_lds", 0)
                base = h.coarse_statistic(prec, templates=rng_t).astype(np.float64)
                h._gpu.set_tuning("coarse_lds", 1)
                assert (np.abs(got - base) / base).max() <= _COARSE_PRECISIONS[prec][1] / 4
