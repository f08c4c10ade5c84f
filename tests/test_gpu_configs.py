"""Full-size parity of the configurations the benchmark times, and of every
launch shape the autotuner may pick.

* Configuration 2a (the headline): n = 4096, 16 x 1024 pairs, one bin,
  threshold 5.5, built by bench.workload() exactly as timed.  The device
  peaks against the FP64 oracle over all 16384 pairs, the injected lags
  found where the oracle finds them, and every tuned data-rows-per-workgroup
  candidate bit-identical to the tuned launch.
* Configuration 3: the chirp bank (bench.chirp_bank), 16 x 1024, band 512, in
  the gate model's units, on the 1%-injected SNR-8 batch of the benchmark,
  for the FP32, fp16 and bf16 screens: the coarse statistic against the FP64
  coarse correlation within the screen's widening, the gate decisions a
  superset of FP64's, the one-sided guarantee against the device flat filter,
  and the injected detections kept.
* Autotuner candidates: every flat dpw at every tuned shape, every
  four-step chunk size at both four-step shapes, every refine grid, and
  every persistent-queue grid (the MFMA rows x templates per wave are
  covered in test_gpu_mfma.py).
"""
import numpy as np
import pytest

import apogee_amd as mf
import bench
from apogee_amd.autotune import BIG_SHAPES, CANDIDATES, FLAT_SHAPES
from apogee_amd.hierarchical import _COARSE_PRECISIONS
from oracle import mf_oracle as O
from _parity import assert_peaks_match, tol

pytestmark = pytest.mark.gpu


def flat_run(D, T, binsize, threshold, dpw=None, window=None):
    f = mf.MatchedFilter(D.shape[1], D.shape[0], T.shape[0], device="gpu")
    if dpw is not None:
        f._gpu.set_tuning("dpw", dpw)
    f.set_data(D)
    f.set_templates(T)
    kw = {} if window is None else {"window": window}
    return f.run(binsize=binsize, threshold=threshold, **kw).copy()


@pytest.fixture(scope="module")
def headline(gpu_device):
    D, T, pairs, lags = bench.workload()
    return D, T, pairs, lags, flat_run(D, T, bench.N, bench.THRESHOLD)


def test_headline_config_2a_against_the_oracle(headline):
    D, T, pairs, lags, pk = headline
    assert pk.shape == (16, 1024, 1)
    excused = assert_peaks_match(D, T, pk["index"], pk["value"], bench.N, bench.THRESHOLD,
                                 context="config 2a", pair_chunk=2048)
    # the injected lags: wherever the FP64 filter puts the pair's peak at the
    # injected lag above the threshold, the device reports exactly that lag
    ridx = np.empty((16, 1024), np.int64)
    for t0 in range(0, 1024, 128):
        ridx[:, t0:t0 + 128] = O.flat_peaks(D, T[t0:t0 + 128], bench.N, bench.THRESHOLD)[0][..., 0]
    at_lag = pairs & (ridx == lags)
    assert at_lag.sum() > 0.95 * pairs.sum(), (at_lag.sum(), pairs.sum())
    np.testing.assert_array_equal(pk["index"][..., 0][at_lag], lags[at_lag])
    found = (pk["index"][..., 0] >= 0) & pairs
    assert found.sum() == ((ridx >= 0) & pairs).sum()
    print("config 2a: %d injected, %d at their lag, %d near-ties excused"
          % (pairs.sum(), at_lag.sum(), excused))


@pytest.mark.parametrize("dpw", CANDIDATES["flat"]["dpw"])
def test_headline_every_dpw_is_bit_identical(dpw, headline):
    D, T, _, _, pk = headline
    got = flat_run(D, T, bench.N, bench.THRESHOLD, dpw=dpw)
    np.testing.assert_array_equal(got["index"], pk["index"])
    np.testing.assert_array_equal(got["value"], pk["value"])


@pytest.mark.parametrize("shape", FLAT_SHAPES, ids=lambda s: "n%d_%dx%d_bs%d" % s)
def test_every_tuned_flat_shape_and_dpw(shape, gpu_device):
    """Each flat shape of the tuning table at every dpw candidate: identical
    bits across candidates, and the FP64 oracle on a slice of templates."""
    n, nd, nt, bs = shape
    rng = np.random.default_rng(n + nd + nt)
    nt_run = min(nt, 256)              # every dpw tile shape is exercised at 256 templates
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    T = (rng.standard_normal((nt_run, n)) + 1j * rng.standard_normal((nt_run, n))).astype(np.complex64)
    D[0] += 3 * T[1]
    base = None
    for dpw in CANDIDATES["flat"]["dpw"]:
        got = flat_run(D, T, bs, 0.0, dpw=dpw)
        if base is None:
            base = got
            assert_peaks_match(D, T[:8], got["index"][:, :8], got["value"][:, :8], bs, 0.0,
                               context="shape %s" % (shape,))
        else:
            np.testing.assert_array_equal(got["index"], base["index"], err_msg="dpw %d" % dpw)
            np.testing.assert_array_equal(got["value"], base["value"], err_msg="dpw %d" % dpw)


@pytest.mark.parametrize("shape", BIG_SHAPES, ids=lambda s: "n%d_%dx%d_bs%d" % s)
def test_every_four_step_chunk_is_bit_identical(shape, gpu_device):
    """Each four-step shape of the tuning table at every chunk size (the
    pairs whose intermediate is held at once: one chunk, or several with a
    ragged last one): identical bits, and the FP64 oracle on a slice, with
    a threshold so that the row kernel's below-seed waves skip selection."""
    n, nd, nt, bs = shape
    rng = np.random.default_rng(n + nt)
    D = (rng.standard_normal((nd, n)) + 1j * rng.standard_normal((nd, n))).astype(np.complex64)
    T = (rng.standard_normal((nt, n)) + 1j * rng.standard_normal((nt, n))).astype(np.complex64)
    T /= np.sqrt(np.sum(np.abs(T) ** 2, axis=1, keepdims=True))   # Re z, Im z of unit variance
    D[1] += 8.0 * T[5]                                               # |z| = 8 at lag 0 of pair (1, 5)
    thr = 4.0   # |z|^2 > 16 in noise: about 3e-4 per lag, so some bins are set and most waves skip
    base = None
    for mb in CANDIDATES["big"]["big_chunk_mb"]:
        f = mf.MatchedFilter(n, nd, nt, device="gpu")
        f._gpu.set_tuning("big_chunk_mb", mb)
        f.set_data(D)
        f.set_templates(T)
        got = f.run(binsize=bs, threshold=thr).copy()
        if base is None:
            base = got
            assert (got["index"] >= 0).any() and (got["index"] < 0).any()
            assert_peaks_match(D, T[:8], got["index"][:, :8], got["value"][:, :8], bs, thr,
                               context="shape %s" % (shape,))
        else:
            np.testing.assert_array_equal(got["index"], base["index"], err_msg="chunk %d MiB" % mb)
            np.testing.assert_array_equal(got["value"], base["value"], err_msg="chunk %d MiB" % mb)


# ------------------------------------------------------------ configuration 3

@pytest.fixture(scope="module")
def chirp(gpu_device):
    rng = np.random.default_rng(bench.SEED + 1)
    T = bench.chirp_bank(rng)
    power = np.mean(np.abs(T.astype(np.complex128)) ** 2, axis=0).astype(np.float32)
    D, inj, lags = bench.hier_workload(bench.SEED + 100, T, snr=8.0, frac=0.01)
    flat = flat_run(D, T, bench.N, 5.5)
    # FP64 coarse statistic of the whole batch: D[:512] against T[:512]/sqrt(f)
    f = O.band_fraction(T, 512, reference=power)
    ct = O.coarse_templates(T, 512, reference=power)
    exact = np.empty((16, 1024))
    for t0 in range(0, 1024, 256):
        z = O.correlate(D[:, None, :512], ct[None, t0:t0 + 256])
        exact[:, t0:t0 + 256] = np.abs(z).max(axis=-1)
    return T, power, D, inj, lags, flat, exact, f


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
def test_config3_chirp_bank_per_screen(prec, chirp):
    T, power, D, inj, lags, flat, exact, f = chirp
    h = mf.HierarchicalFilter(bench.N, 16, 1024, snr=5.5, fd=1e-2, band=512, device="gpu")
    h.set_reference(power)
    h.set_coarse_precision(prec)
    h.set_templates(T)
    h.set_data(D)
    thr = h.coarse_threshold
    delta = _COARSE_PRECISIONS[prec][1]
    # the coarse statistic against FP64
    got = h.coarse_statistic(prec).astype(np.float64)
    rel = np.abs(got - exact) / exact
    bound = max(delta / 2, 4 * tol(512))
    print("%s: coarse statistic max rel err %.3g (bound %.3g), gate %.4f" % (prec, rel.max(), bound, thr))
    assert rel.max() <= bound
    # gate decisions: every pair FP64 keeps (with an FP32 rounding margin) the
    # device keeps; the device gate is thr / (1 + delta)
    dev_keep = got >= thr / (1 + delta)
    fp64_keep = exact >= thr * (1 + 4 * tol(512))
    assert not (fp64_keep & ~dev_keep).any()
    pk = h.run(binsize=bench.N, threshold=5.5)
    # one-sided: nothing invented, every report the flat filter's exact bits
    fired = pk["index"] >= 0
    assert not (fired & (flat["index"] < 0)).any()
    np.testing.assert_array_equal(pk["index"][fired], flat["index"][fired])
    np.testing.assert_array_equal(pk["value"][fired], flat["value"][fired])
    # the refined pairs are the device gate's survivors (pairs within FP32
    # rounding of the gate may fall either side)
    edge = np.abs(got - thr / (1 + delta)) <= 1e-5 * thr
    assert int((dev_keep & ~edge).sum()) <= h.stats[1] <= int((dev_keep | edge).sum()), h.stats
    # injected SNR-8 signals the flat filter detects are kept
    det = (flat["index"][..., 0] >= 0) & inj
    omitted = det & ~fired[..., 0]
    print("%s: %d injected, %d detected flat, %d omitted, refine rate %.4f"
          % (prec, inj.sum(), det.sum(), omitted.sum(), h.refine_rate))
    assert omitted.sum() <= max(1, 0.01 * det.sum())


@pytest.mark.parametrize("grid", CANDIDATES["hier"]["refine_grid"])
def test_every_refine_grid_is_bit_identical(grid, chirp):
    """Refine launches of every tuned grid size give the same bits, with a
    low gate so about a quarter of the pairs are refined."""
    T, power, D, _, _, flat, exact, _ = chirp
    gate = float(np.quantile(exact, 0.75))
    outs = []
    for g in (CANDIDATES["hier"]["refine_grid"][0], grid):
        h = mf.HierarchicalFilter(bench.N, 16, 1024, band=512, device="gpu")
        h.set_reference(power)
        h.set_coarse_threshold(gate)
        h._gpu.set_tuning("refine_grid", g)
        h.set_templates(T)
        h.set_data(D)
        outs.append(h.run(binsize=512, threshold=3.0).copy())
        assert 0.15 < h.refine_rate < 0.35
    np.testing.assert_array_equal(outs[0]["index"], outs[1]["index"])
    np.testing.assert_array_equal(outs[0]["value"], outs[1]["value"])


def test_every_queue_grid_is_bit_identical(gpu_device):
    """The persistent refine queue at every tuned workgroups-per-CU value:
    identical results, and identical to the per-request flat path."""
    from apogee_amd.autotune import uneven_workload
    pool, plan = uneven_workload(300, 77)
    results = []
    for gpc in CANDIDATES["queue"]["grid_per_cu"]:
        q = mf.RefineQueue(device="gpu")
        q.set_tuning("grid_per_cu", gpc)
        q.plan(pool, **plan)
        results.append(q.run())
    for other in results[1:]:
        for a, b in zip(results[0], other):
            np.testing.assert_array_equal(a["index"], b["index"])
            np.testing.assert_array_equal(a["value"], b["value"])
    host = pool.cpu().numpy()
    for r in range(0, 300, 37):
        L, c = int(plan["n"][r]), int(plan["candidates"][r])
        d0, t0 = int(plan["data_offset"][r]), int(plan["template_offset"][r])
        ref = flat_run(host[d0:d0 + L][None], host[t0:t0 + c * L].reshape(c, L), L // 8, 0.0)
        np.testing.assert_array_equal(results[0][r]["index"], ref["index"][0])
        np.testing.assert_array_equal(results[0][r]["value"], ref["value"][0])
