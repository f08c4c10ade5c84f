"""HierarchicalFilter on the MI355X.

Asserted against the one-sided guarantee (the reference's
tests/test_hierarchical_matrix.py idea): every reported peak is bit-identical
to the flat filter's on the same device, only omissions are allowed; plus the
coarse decisions against the FP64 oracle, the measured omission rate against
the budget, and that the screen actually skips work on noise.
"""
import numpy as np
import pytest

import apogee_amd as mf
from oracle import mf_oracle as O

pytestmark = pytest.mark.gpu

PROFILE = None


def inspiral_power(n, exponent=-7 / 3.0, knee_frac=0.0150):
    p = np.zeros(n, dtype=np.float32)
    k = np.arange(1, n // 2).astype(np.float64)
    p[1:n // 2] = (k ** exponent / ((knee_frac * n / k) ** 4 + 1.0)).astype(np.float32)
    return p / p.sum()


def template_with_power(n, power, rng=None):
    h = np.sqrt(power).astype(np.complex64)
    if rng is not None:   # random phases keep the bank's templates distinct
        h = (h * np.exp(2j * np.pi * rng.random(n))).astype(np.complex64)
    return h / np.linalg.norm(h)


def noise(shape, rng):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def loud_case(n, nd, nt, seed=11, snr=12.0):
    rng = np.random.default_rng(seed)
    power = inspiral_power(n)
    H = np.stack([template_with_power(n, power, rng) for _ in range(nt)])
    D = noise((nd, n), rng)
    for d in range(nd):
        lag = 300 + 17 * d
        D[d] += (snr * H[d % nt] * np.exp(2j * np.pi * np.arange(n) * lag / n)).astype(np.complex64)
    return power, D, H


def build(n, nd, nt, power, snr=5.5, fd=1e-2, band=None):
    flat = mf.MatchedFilter(n, nd, nt, device="gpu")
    hier = mf.HierarchicalFilter(n, nd, nt, snr=snr, fd=fd, band=band, device="gpu")
    hier.set_reference(power)
    return flat, hier


def assert_one_sided(a, b, context=""):
    fired = b["index"] >= 0
    assert not (fired & (a["index"] < 0)).any(), "invented a peak %s" % context
    np.testing.assert_array_equal(a["index"][fired], b["index"][fired], err_msg=context)
    np.testing.assert_array_equal(a["value"][fired], b["value"][fired], err_msg=context)
    return fired


def run_both(n, nd, nt, binsize, threshold, window, seed=11, snr=12.0):
    power, D, H = loud_case(n, nd, nt, seed=seed, snr=snr)
    flat, hier = build(n, nd, nt, power)
    for o in (flat, hier):
        o.set_data(D)
        o.set_templates(H)
    kw = {} if window is None else {"window": window}
    return flat.run(binsize=binsize, threshold=threshold, **kw), \
        hier.run(binsize=binsize, threshold=threshold, **kw), hier


@pytest.mark.parametrize("binsize", [1, 3, 64, 100, 257, 512, 1024, 4096])
def test_binsize(binsize, gpu_device):
    a, b, _ = run_both(4096, 2, 4, binsize, 5.5, None)
    assert_one_sided(a, b, "binsize=%d" % binsize)


@pytest.mark.parametrize("window", [(0, 4096), (0, 4000), (37, 4000), (1, 4095), (500, 501),
                                    (2048, 4096)])
def test_window_alignment(window, gpu_device):
    a, b, _ = run_both(4096, 2, 4, 128, 5.5, window)
    assert_one_sided(a, b, "window=%s" % (window,))


@pytest.mark.parametrize("binsize,window", [(300, (37, 4000)), (7, (13, 3999)), (8192, (100, 900))])
def test_awkward_combinations(binsize, window, gpu_device):
    a, b, _ = run_both(4096, 2, 4, binsize, 5.5, window)
    assert_one_sided(a, b, "binsize=%d window=%s" % (binsize, window))


@pytest.mark.parametrize("n", [1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072])
def test_every_transform_length(n, gpu_device):
    a, b, _ = run_both(n, 2, 4, n // 8, 5.5, None)
    assert_one_sided(a, b, "n=%d" % n)


@pytest.mark.parametrize("threshold", [3.0, 5.5, 8.0, 20.0])
def test_thresholds(threshold, gpu_device):
    a, b, _ = run_both(4096, 2, 4, 512, threshold, None)
    fired = assert_one_sided(a, b, "threshold=%.1f" % threshold)
    if threshold >= 20.0:
        assert not fired.any()


def test_loud_signals_survive_and_noise_is_screened(gpu_device):
    a, b, hier = run_both(4096, 4, 4, 1024, 5.5, None)
    loud = np.zeros((4, 4), bool)
    for d in range(4):
        loud[d, d % 4] = True
    assert (b["index"][loud] >= 0).any(axis=-1).all(), "a 12-sigma signal was dismissed"
    n, nd, nt = 4096, 4, 16
    rng = np.random.default_rng(5)
    power = inspiral_power(n)
    H = np.stack([template_with_power(n, power, rng) for _ in range(nt)])
    # band pinned: at 64 pairs the automatic choice may rightly prefer a
    # narrow band that refines more (refining 64 pairs costs less than the
    # wider screen), and this checks that a screen does skip noise
    _, hier = build(n, nd, nt, power, band=512)
    hier.set_data(noise((nd, n), rng))
    hier.set_templates(H)
    hier.run(binsize=n, threshold=5.5)
    assert hier.refine_rate < 0.5, hier.refine_rate


def test_coarse_decisions_match_the_fp64_oracle(gpu_device):
    """The device gate reproduces the oracle's fired set except where the
    FP64 coarse maximum sits within 1e-4 (relative) of the gate."""
    n, nd, nt, band = 4096, 8, 32, 512
    power, D, H = loud_case(n, nd, nt, seed=3, snr=7.0)
    hier = mf.HierarchicalFilter(n, nd, nt, band=band, device="gpu")
    hier.set_reference(power)
    gate = 4.0
    hier.set_coarse_threshold(gate)
    hier.set_data(D)
    hier.set_templates(H)
    got = hier.run(binsize=n, threshold=0.0)
    f = O.band_fraction(H, band, power)
    ct = (H[:, :band].astype(np.complex128) / np.sqrt(f)).astype(np.complex64)
    zc = O.correlate(D[:, None, :band], ct[None, :, :])
    cmax = np.abs(zc).max(axis=-1)
    fired_oracle = cmax >= gate
    fired_dev = got["index"][..., 0] >= 0
    near = np.abs(cmax - gate) <= 1e-4 * gate
    assert ((fired_oracle == fired_dev) | near).all()
    assert hier.stats == (nd * nt, int(fired_dev.sum()))


def test_explicit_threshold_zero_refines_everything(gpu_device):
    n = 4096
    power, D, H = loud_case(n, 8, 4, seed=5, snr=6.0)
    rates = {}
    for thr in (0.0, 3.5, 4.5):
        hf = mf.HierarchicalFilter(n, 8, 4, snr=5.0, fd=1e-3, band=512, taps=8, device="gpu")
        hf.set_reference(power)
        hf.set_coarse_threshold(thr)
        hf.set_templates(H)
        hf.set_data(D)
        hf.run(binsize=n, threshold=5.0)
        rates[thr] = hf.refine_rate
    assert rates[0.0] == pytest.approx(1.0)
    assert rates[3.5] >= rates[4.5]


@pytest.mark.parametrize("band", [128, 256, 512])
def test_gate_holds_under_scaling_and_bank_partition(band, gpu_device):
    """The reference's tests/test_calibration_invariants.py:19-69: with a
    coarse gate in a wide gap of the FP64 coarse maxima (so rounding cannot
    move a decision), the hierarchical filter reports exactly the flat peaks
    of the admitted pairs, whatever the template order, the partition of the
    bank over filters, a reciprocal scaling of data and templates or the
    amplitude of the reference; a changed reference rescales the coarse
    templates, moving the decisions to the new band fraction."""
    n, nd, nt = 1024, 12, 4
    rng = np.random.default_rng(810 + band)
    data = (rng.normal(size=(nd, n)) + 1j * rng.normal(size=(nd, n))).astype(np.complex64)
    templates = (rng.normal(size=(nt, n)) + 1j * rng.normal(size=(nt, n))).astype(np.complex64)
    templates *= np.exp(-np.arange(n)[None, :] / np.array([80, 160, 320, 640])[:, None])
    templates /= np.linalg.norm(templates, axis=1)[:, None]
    reference = np.exp(-np.arange(n) / 240).astype(np.float32)
    fraction = reference[:band].astype(float).sum() / reference.astype(float).sum()
    product = data[:, None, :band].astype(np.complex128) * np.conj(templates[None, :, :band])
    coarse = np.abs(np.fft.ifft(product, axis=-1) * band).max(axis=-1) / np.sqrt(fraction)
    ordered = np.sort(coarse.ravel())
    lo, hi = len(ordered) // 3, 2 * len(ordered) // 3
    gap = lo + int(np.argmax(np.diff(ordered)[lo:hi]))
    threshold = (ordered[gap] + ordered[gap + 1]) / 2
    assert ordered[gap + 1] - ordered[gap] > 0.01 * threshold
    admitted = coarse >= threshold
    assert admitted.any() and (~admitted).any()
    flat = mf.MatchedFilter(n, nd, nt, device="gpu")
    flat.set_data(data)
    flat.set_templates(templates)
    expected = flat.run().copy()
    expected["index"][~admitted] = -1
    expected["value"][~admitted] = 0
    for scale, refscale, chunks in [(1.0, 1.0, [(0, 4)]), (4.0, 8.0, [(0, 1), (1, 4)])]:
        order = np.arange(nt)[::-1]
        for begin, end in chunks:
            cols = order[begin:end]
            f = mf.HierarchicalFilter(n, nd, len(cols), band=band, device="gpu")
            f.set_coarse_threshold(float(threshold))
            f.set_reference(reference * refscale)
            f.set_templates(templates[cols] / scale)
            f.set_data(data * scale)
            got = f.run().copy()
            np.testing.assert_array_equal(got["index"], expected["index"][:, cols])
            np.testing.assert_array_equal(got["value"], expected["value"][:, cols])
            assert f.stats == (nd * len(cols), int(admitted[:, cols].sum()))
            newref = reference.copy()
            newref[:band] *= 0.5
            f.set_reference(newref)
            newfraction = newref[:band].astype(float).sum() / newref.astype(float).sum()
            newmask = coarse[:, cols] * np.sqrt(fraction / newfraction) >= threshold
            changed = f.run()
            np.testing.assert_array_equal(changed["index"] >= 0, np.broadcast_to(newmask[..., None],
                                                                                  changed["index"].shape))


def test_reference_update_recalibrates_a_reused_filter(gpu_device):
    """The reference's tests/test_series_lifecycle.py:120-142: a flat spectrum
    of which only the band survives in the data gives a coarse maximum of
    band / (n sqrt(f)): 0.5 at band fraction 1/4 and 0.293 at 8/11.  With the
    gate between them, each new reference must reach the executed gate."""
    n, band = 2048, 512
    f = mf.HierarchicalFilter(n, band=band, snr=5.5, fd=1e-2, device="gpu")
    f.set_coarse_threshold(0.35)
    h = np.full((1, n), 1.0 / np.sqrt(n), np.complex64)
    d = h.copy()
    d[:, band:] = 0
    f.set_templates(h)
    f.set_data(d)
    for low_power in (1.0, 8.0, 1.0):
        ref = np.ones(n, np.float32)
        ref[:band] = low_power
        f.set_reference(ref)
        peaks = f.run(binsize=n)
        assert (peaks["index"] >= 0).all() == (low_power == 1.0)
        fraction = ref[:band].astype(float).sum() / ref.astype(float).sum()
        assert f.coarse_statistic("fp32")[0, 0] == pytest.approx(band / (n * np.sqrt(fraction)), rel=1e-6)


@pytest.mark.parametrize("band,prec", [(64, "fp32"), (128, "fp32"), (256, "fp32"), (512, "fp32"),
                                       (1024, "fp32"), (2048, "fp32"), (512, "fp16"), (512, "bf16")])
def test_ties_through_coarse_compaction_and_refinement(band, prec, gpu_device):
    """The reference's tests/test_gpu_peak_ties.py:32-51: a delta at band/4,
    inside the coarse band, ties |z| = 1 at every coarse and every full lag.
    Every pair clears a coarse gate of 0.5, is compacted and refined, and the
    reported value is the full correlation at the reported lag."""
    n = 4096
    bank = np.zeros((4, n), np.complex64)
    bank[:, band // 4] = 1
    f = mf.HierarchicalFilter(n, 4, 4, band=band, device="gpu")
    f.set_coarse_threshold(0.5)
    if prec != "fp32":
        f.set_coarse_precision(prec)
    f.set_data(bank)
    f.set_templates(bank)
    for _ in range(3):
        out = f.run()
        assert (out["index"] >= 0).all()
        expected = np.exp(2j * np.pi * (band // 4) * out["index"] / n)
        np.testing.assert_allclose(out["value"], expected, atol=3e-5)
    assert f.stats == (48, 48)


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
def test_omission_rate_meets_the_budget(prec, gpu_device):
    """Measured false dismissal of SNR-5.5 injections against fd = 1e-2, for
    the FP32 screen and the widened matrix-core screens (same trials)."""
    n, trials, snr, fd = 4096, 4096, 5.5, 1e-2
    rng = np.random.default_rng(13)
    power = inspiral_power(n)
    H = template_with_power(n, power)
    nd = 512
    flat, hier = build(n, nd, 1, power, snr=snr, fd=fd)
    hier.set_coarse_precision(prec)
    flat.set_templates(H[None, :])
    hier.set_templates(H[None, :])
    ph = np.exp(2j * np.pi * np.arange(n) / n)
    detected = omitted = 0
    for b0 in range(0, trials, nd):
        D = noise((nd, n), rng)
        for i in range(nd):
            D[i] += (snr * H * ph ** ((37 * (b0 + i)) % n)).astype(np.complex64)
        flat.set_data(D)
        hier.set_data(D)
        a = flat.run(binsize=n, threshold=snr)["index"][:, 0, 0]
        b = hier.run(binsize=n, threshold=snr)["index"][:, 0, 0]
        detected += int((a >= 0).sum())
        omitted += int(((a >= 0) & (b < 0)).sum())
    assert detected > 1000
    rate = omitted / detected
    print("%s fd budget %.3g: omitted %d of %d detections (%.4f)" % (prec, fd, omitted, detected, rate))
    assert rate <= 2 * fd, rate


def test_sub_ranges_match_slices(gpu_device):
    n = 4096
    power, D, H = loud_case(n, 6, 5, seed=9)
    hier = mf.HierarchicalFilter(n, 6, 5, band=512, device="gpu")
    hier.set_reference(power)
    hier.set_coarse_threshold(3.0)
    hier.set_data(D)
    hier.set_templates(H)
    full = hier.run(binsize=512, threshold=5.0).copy()
    sub = hier.run(binsize=512, threshold=5.0, data=(2, 3), templates=(1, 3))
    np.testing.assert_array_equal(sub["index"], full["index"][2:5, 1:4])
    np.testing.assert_array_equal(sub["value"], full["value"][2:5, 1:4])


@pytest.mark.parametrize("graphs", [0, 1])
def test_repeated_calls_alternate_counters_and_graphs(graphs, gpu_device):
    """Consecutive calls alternate two survivor counters and, with graphs on,
    identical calls replay one captured graph per counter parity.  A run of identical calls,
    new data between replays and switches of the coarse precision must each
    give exactly what a fresh filter gives on the same inputs, survivor
    counts included."""
    n, nd, nt = 4096, 16, 64
    rng = np.random.default_rng(23)
    power = inspiral_power(n)
    H = np.stack([template_with_power(n, power, rng) for _ in range(nt)])
    datasets = []
    for s in range(2):
        D = noise((nd, n), rng)
        for d in range(0, nd, 2 + s):     # different survivor sets per data set
            lag = 100 + 31 * d + 7 * s
            D[d] += (9.0 * H[(5 * d + s) % nt] * np.exp(2j * np.pi * np.arange(n) * lag / n)).astype(np.complex64)
        datasets.append(D)

    def fresh(prec, D):
        h = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2, band=512, device="gpu")
        h.set_reference(power)
        h.set_coarse_precision(prec)
        h.set_templates(H)
        h.set_data(D)
        out = h.run(binsize=1024, threshold=5.0)
        return out["index"].copy(), out["value"].copy(), h.stats[1]

    expected = {(p, i): fresh(p, D) for p in ("fp16", "bf16", "fp32") for i, D in enumerate(datasets)}
    h = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2, band=512, device="gpu")
    h.set_reference(power)
    h.set_templates(H)
    h._gpu.set_tuning("graphs", graphs)
    steps = [("fp16", 0)] * 5 + [("fp16", 1)] * 3 + [("fp32", 0)] * 3 + [("bf16", 1)] * 4 + \
            [("fp16", 0), ("bf16", 0), ("fp16", 1)] * 2
    for step, (prec, i) in enumerate(steps):
        h.set_coarse_precision(prec)
        h.set_data(datasets[i])
        before = h.stats[1]
        out = h.run(binsize=1024, threshold=5.0)
        idx, val, refined = expected[(prec, i)]
        ctx = "step %d: %s on data set %d" % (step, prec, i)
        np.testing.assert_array_equal(out["index"], idx, err_msg=ctx)
        np.testing.assert_array_equal(out["value"], val, err_msg=ctx)
        assert h.stats[1] - before == refined, ctx
        assert 0 < refined < nd * nt, ctx


@pytest.mark.parametrize("prec", ["fp16", "fp32"])
def test_refine_grid_sized_from_a_stale_count_drains_every_survivor(prec, gpu_device):
    """The refine launch is sized from the previous call's survivor count: a
    call that refines every pair right after one that refined none (the
    smallest grid) must still drain the whole list, and the reverse must
    give the noise answer; each against the flat filter on the same data."""
    n, nd, nt = 4096, 16, 1024
    rng = np.random.default_rng(41)
    power = inspiral_power(n)
    H = np.stack([template_with_power(n, power, rng) for _ in range(nt)])
    D = noise((nd, n), rng)
    flat, hier = build(n, nd, nt, power, band=512)
    hier.set_coarse_precision(prec)
    for f in (flat, hier):
        f.set_templates(H)
        f.set_data(D)
    ref = flat.run(binsize=n, threshold=4.5)
    for step, gate in enumerate((None, 0.0, None, 0.0)):
        hier.set_coarse_threshold(gate)
        before = hier.stats[1]
        out = hier.run(binsize=n, threshold=4.5)
        refined = hier.stats[1] - before
        ctx = "step %d, %s, gate %s" % (step, prec, gate)
        if gate == 0.0:
            assert refined == nd * nt, ctx
            np.testing.assert_array_equal(out["index"], ref["index"], err_msg=ctx)
            np.testing.assert_array_equal(out["value"], ref["value"], err_msg=ctx)
        else:
            assert refined < nd * nt // 8, ctx
            assert_one_sided(ref, out, ctx)


@pytest.mark.parametrize("prec", ["fp16", "bf16", "fp32"])
def test_one_launch_screen_and_refine_matches_two_launches(prec, gpu_device):
    """At n = 4096 the hierarchical call runs in fewer launches (tuning key
    coarse_fuse): with the matrix-core screen and one bin, the screen's
    workgroups refine their own survivors; with the FP32 screen, one launch
    gates and refines after the screen (one bin, or bins on 64-lag
    boundaries).  The separate launches (screen, gate, survivor list,
    list-driven refine) run the same arithmetic in another order of launches,
    so every output must be bit-identical, and so must the refinement count:
    with the automatic gate (a few survivors), a gate a quarter of the pairs
    pass, and every pair refined; over splits of rows and templates per wave
    (data rows not a multiple of a workgroup's, templates not of a wave's),
    one bin and 64-lag bins, the whole transform, an aligned window and an
    unaligned one (which keeps the separate launches for bins)."""
    n, nd, nt = 4096, 12, 200
    power, D, H = loud_case(n, nd, nt, seed=43, snr=12.0)
    hier = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2, band=512, device="gpu")
    hier.set_reference(power)
    hier.set_coarse_precision(prec)
    hier.set_templates(H)
    hier.set_data(D)
    ctx = hier._gpu
    ctx.set_tuning("coarse_lds", 1)
    cmax = np.sort(hier.coarse_statistic().ravel())
    gates = (None, float(cmax[int(0.75 * cmax.size)]), 0.0)
    splits = ((1, 4), (2, 2), (1, 16), (2, 16), (1, 1)) if prec != "fp32" else ((1, 4),)
    for rpw, tpw in splits:
        ctx.set_tuning("coarse_rpw", rpw)
        ctx.set_tuning("coarse_tpw", tpw)
        for gate in gates:
            hier.set_coarse_threshold(gate)
            for binsize in (n, 512):
                for window in (None, (128, 3968), (100, 3900)):
                    got = []
                    for fuse in (1, 0):
                        ctx.set_tuning("coarse_fuse", fuse)
                        before = hier.stats[1]
                        out = hier.run(binsize=binsize, threshold=4.5, window=window)
                        got.append((out, hier.stats[1] - before))
                    (a, ra), (b, rb) = got
                    label = "%s, %d x %d per wave, gate %s, binsize %d, window %s" % (
                        prec, rpw, tpw, gate, binsize, window)
                    print(label, "refined", ra, flush=True)
                    assert ra == rb, label
                    if gate == 0.0:
                        assert ra == nd * nt, label
                    np.testing.assert_array_equal(a["index"], b["index"], err_msg=label)
                    np.testing.assert_array_equal(a["value"], b["value"], err_msg=label)
