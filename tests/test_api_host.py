"""The Python surface on a host double of the device (no GPU needed).

Mirrors the reference's API contract (python/matchedfilter/__init__.py:176-686,
:1055-1461; tests/test_api.py, tests/test_device_parity.py:31-140): shapes
and errors, sub-range = slice, raw = structured, counts, readiness checks,
hierarchical configuration, stats and the gate model's invariants.
"""
import numpy as np
import pytest

import apogee_amd as mf
from apogee_amd import device as mfdev
from oracle import mf_oracle as O
from _fake import FakeContext


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(mf.MatchedFilter, "_context_factory", FakeContext)
    monkeypatch.setattr(mfdev, "devices", lambda: [mfdev.Device("gpu", 0, "fake gfx950")])
    FakeContext.calls = []
    return FakeContext


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_shapes_and_errors(fake):
    with pytest.raises(ValueError):
        mf.MatchedFilter(1000, 1, 1)
    with pytest.raises(ValueError):
        mf.MatchedFilter(1024, 0, 1)
    f = mf.MatchedFilter(1024, 2, 3)
    with pytest.raises(ValueError):
        f.set_data(np.zeros((2, 100), np.complex64))
    with pytest.raises(IndexError):
        f.set_data(np.zeros(1024, np.complex64), index=2)
    with pytest.raises(ValueError):
        f.set_data(np.zeros(100, np.complex64), index=0)
    with pytest.raises(ValueError, match="no data"):
        f.run()
    f.set_data(np.zeros((2, 1024), np.complex64))
    with pytest.raises(ValueError, match="no templates"):
        f.run()
    f.set_templates(np.zeros((3, 1024), np.complex64))
    with pytest.raises(ValueError):
        f.run(binsize=0)
    with pytest.raises(ValueError):
        f.run(window=(500, 500))
    with pytest.raises(ValueError):
        f.run(data=(1, 2))
    with pytest.raises(ValueError):
        f.nbins(0)
    assert f.nbins(100) == 11
    assert f.nbins(100, window=(10, 1000)) == 10


def test_device_strings(fake):
    assert str(mfdev.parse(None)) == "gpu:0"
    assert str(mfdev.parse("gpu:0")) == "gpu:0"
    assert mfdev.parse("auto") == "gpu:0"
    with pytest.raises(ValueError):
        mfdev.parse("cpu")
    with pytest.raises(ValueError):
        mfdev.parse("tpu")
    with pytest.raises(ValueError):
        mfdev.parse("gpu:7")
    with pytest.raises(TypeError):
        mfdev.parse(3)


def test_results_layout_subranges_raw_counts(fake):
    rng = np.random.default_rng(3)
    n, nd, nt = 1024, 4, 3
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    f = mf.MatchedFilter(n, nd, nt)
    f.set_data(D)
    f.set_templates(T)
    full = f.run(binsize=256, threshold=0.0, window=(100, 900))
    assert full.dtype == mf.PEAK_DTYPE and full.shape == (nd, nt, 4)
    ridx, rval, _ = O.flat_peaks(D, T, 256, 0.0, (100, 900))
    np.testing.assert_array_equal(full["index"], ridx)
    sub = f.run(binsize=256, threshold=0.0, window=(100, 900), data=(1, 2), templates=(1, 2))
    np.testing.assert_array_equal(sub["index"], full["index"][1:3, 1:3])
    idx, val = f.run(binsize=256, threshold=0.0, window=(100, 900), raw=True)
    np.testing.assert_array_equal(idx, full["index"])
    np.testing.assert_array_equal(val, full["value"])
    pk, cnt = f.run(binsize=128, threshold=60.0, counts=True)
    np.testing.assert_array_equal(cnt, (pk["index"] >= 0).sum(axis=2))
    # the device saw exactly the requested sub-range
    assert ("peaks", 1, 2, 1, 2, 256, 0.0, 100, 900) in fake.calls


def test_window_is_clamped(fake):
    rng = np.random.default_rng(4)
    f = mf.MatchedFilter(512, 1, 1)
    f.set_data(cnoise(rng, (1, 512)))
    f.set_templates(cnoise(rng, (1, 512)))
    pk = f.run(binsize=1000, window=(-50, 5000))
    assert pk.shape == (1, 1, 1)
    assert fake.calls[-1][-2:] == (0, 512)


def test_single_rows_track_readiness(fake):
    n = 512
    f = mf.MatchedFilter(n, 3, 2)
    f.set_templates(np.ones((2, n), np.complex64))
    f.set_data(np.ones(n, np.complex64), index=0)
    f.run(data=(0, 1))                     # row 0 is ready
    with pytest.raises(ValueError, match="no data"):
        f.run()                            # rows 1, 2 are not
    f.set_data(np.ones(n, np.complex64), index=1)
    f.set_data(np.ones(n, np.complex64), index=2)
    f.run()
    # a whole bank stays ready when single rows are replaced afterwards
    g = mf.MatchedFilter(n, 3, 2)
    g.set_templates(np.ones((2, n), np.complex64))
    g.set_data(np.ones((3, n), np.complex64))
    g.set_data(np.zeros(n, np.complex64), index=1)
    g.set_templates(np.zeros(n, np.complex64), index=0)
    g.run()


def test_device_arrays_other_than_rocm_are_refused(fake):
    class Foreign:
        def __dlpack_device__(self):
            return (2, 0)                  # CUDA
    f = mf.MatchedFilter(512, 1, 1)
    with pytest.raises(TypeError):
        f.set_data(Foreign())


# ----------------------------------------------------------- hierarchical

def inspiral_power(n, exponent=-7 / 3.0, knee_frac=0.0150):
    p = np.zeros(n, dtype=np.float32)
    k = np.arange(1, n // 2).astype(np.float64)
    p[1:n // 2] = (k ** exponent / ((knee_frac * n / k) ** 4 + 1.0)).astype(np.float32)
    return p / p.sum()


def test_hierarchical_validation_and_config(fake):
    with pytest.raises(ValueError):
        mf.HierarchicalFilter(4096, band=100)
    with pytest.raises(ValueError):
        mf.HierarchicalFilter(4096, band=4096)
    with pytest.raises(ValueError):
        mf.HierarchicalFilter(4096, band=512, taps=3)
    h = mf.HierarchicalFilter(4096, 2, 2)
    with pytest.raises(ValueError):
        h.set_coarse_threshold(3.0)        # needs an explicit band
    with pytest.raises(ValueError):
        h.config                           # automatic band needs a reference
    h.set_reference(inspiral_power(4096))
    band, taps = h.config
    assert taps == 8 and h.cost_table == "gfx950"
    assert 64 <= band < 4096 and band & (band - 1) == 0
    with pytest.raises(ValueError):
        h.set_reference(np.ones(10))
    with pytest.raises(ValueError):
        h.set_reference(-np.ones(4096))
    p = mf.HierarchicalFilter(4096, 2, 2, band=256, taps=4)
    assert p.config == (256, 4) and p.cost_table is None
    with pytest.raises(ValueError):
        p.set_coarse_threshold(float("nan"))
    p.set_coarse_threshold(2.5)
    assert p.coarse_threshold == 2.5


@pytest.mark.parametrize("band", [512, 1024])
def test_taps_is_inert(fake, band):
    """The reference's tests/test_taps_is_inert.py:15-28: taps is metadata;
    the modelled gate does not move with it."""
    gates = set()
    for taps in (2, 4, 8):
        h = mf.HierarchicalFilter(4096, 2, 2, snr=5.0, fd=1e-2, band=band, taps=taps)
        h.set_reference(inspiral_power(4096))
        assert h.config == (band, taps)
        gates.add(h.coarse_threshold)
    assert len(gates) == 1, gates


def test_hierarchical_run_is_identical_or_absent_and_counts_triggers(fake):
    rng = np.random.default_rng(11)
    n, nd, nt = 1024, 4, 4
    power = inspiral_power(n)
    H = np.stack([(np.sqrt(power) * np.exp(2j * np.pi * rng.random(n))).astype(np.complex64)
                  for _ in range(nt)])
    H /= np.linalg.norm(H, axis=1, keepdims=True)
    D = cnoise(rng, (nd, n))
    D[1] += 12 * H[2]
    flat = mf.MatchedFilter(n, nd, nt)
    hier = mf.HierarchicalFilter(n, nd, nt, snr=5.5, fd=1e-2)
    hier.set_reference(power)
    for o in (flat, hier):
        o.set_data(D)
        o.set_templates(H)
    a = flat.run(binsize=256, threshold=5.5)
    b = hier.run(binsize=256, threshold=5.5)
    fired = b["index"] >= 0
    assert fired[1, 2].any()
    np.testing.assert_array_equal(a["index"][fired], b["index"][fired])
    pairs, trig = hier.stats
    assert pairs == nd * nt and 1 <= trig < nd * nt
    assert hier.refine_rate == trig / pairs
    assert 0 < hier.coarse_threshold < 5.5


# -------------------------------------------------------------- gate model

def test_gate_model_invariants():
    from apogee_amd import gatemodel as G
    import os
    p = np.load(os.path.join(os.path.dirname(__file__), "data", "reference_profile_pycbc.npy"),
                allow_pickle=False)
    g2 = G.gate_for(p, 4096, 512, 5.5, 1e-2)
    g3 = G.gate_for(p, 4096, 512, 5.5, 1e-3)
    assert 0 < g3 < g2 < 5.5                       # a tighter budget lowers the gate
    assert G.gate_for(p, 4096, 512, 6.5, 1e-2) > g2  # a stronger signal raises it
    # the modelled dismissal at the gate is the budget
    assert G.dismissal(p, 4096, 512, 5.5, g2, 1e-2) == pytest.approx(1e-2, abs=2e-3)
    with pytest.raises(ValueError):
        G.gate_for(p, 4096, 500, 5.5, 1e-2)
    with pytest.raises(ValueError):
        G.gate_for(p, 4096, 512, 5.5, 1.5)
    # the coarse signal is at most snr * sqrt(f): the gate sits below it
    f = float(p[:512].sum() / p.sum())
    assert g2 < 5.5 * np.sqrt(f)
    # a profile wholly in band but narrow-peaked in lag scallops on the coarse
    # grid, so its gate is lower than the broadband profile's
    flat = np.zeros(4096)
    flat[:512] = 1.0
    assert G.gate_for(flat, 4096, 512, 5.5, 1e-2) < g2


# --------------------------------------------------------------- series

def test_run_series_contract(fake):
    rng = np.random.default_rng(6)
    n = 256
    f = mf.MatchedFilter(n, 1, 2)
    f.set_templates(cnoise(rng, (2, n)))
    s = cnoise(rng, 2000)
    with pytest.raises(ValueError):
        f.run_series(s, [0, 100], [0], [10])                       # lengths differ
    with pytest.raises(ValueError):
        f.run_series(s, [], [], [])                                # no blocks
    with pytest.raises(ValueError):
        f.run_series(s, [0, 100], [0, 0], [256, 100], binsize=64)  # unequal bin counts
    with pytest.raises(ValueError):
        f.run_series(s, [0], [50], [50])                           # empty window
    with pytest.raises(ValueError):
        f.run_series(s.reshape(2, 1000), [0], [0], [10])           # not 1-D
    with pytest.raises(ValueError):
        f.run_series(s, [-1], [0], [10])                           # negative start
    with pytest.raises(ValueError):
        f.run_series(s, [0], [0], [10], templates=(1, 5))
    out = f.run_series(s.real.astype(np.float32), [0, 300, 1900], [0, 0, 0], [256, 256, 256],
                       binsize=64)
    assert out.shape == (3, 2, 4)
    want = O.series_peaks(s.real.astype(np.complex64), [0, 300, 1900], [0, 0, 0], [256, 256, 256],
                          f._gpu.T, 64, 0.0)
    np.testing.assert_array_equal(out["index"], want[0])
    with pytest.raises(ValueError, match="no data"):
        f.run()                                                    # series consumed the slots


@pytest.mark.parametrize("hier", [False, True])
@pytest.mark.parametrize("case", ["no_blocks", "series_2d", "starts_2d", "window_2d", "negative_start",
                                  "huge_start", "empty_templates", "past_templates", "zero_binsize",
                                  "negative_binsize"])
def test_invalid_series_layout_fails_before_the_device(fake, case, hier):
    """The reference's tests/test_series_lifecycle.py:83-117: every malformed
    series layout raises ValueError in Python and never reaches the device."""
    n = 1024
    f = mf.HierarchicalFilter(n, 1, 1, band=128) if hier else mf.MatchedFilter(n, 1, 1)
    f.set_templates(np.ones((1, n), np.complex64))
    if hier:
        f.set_reference(np.ones(n, np.float32))
    kw = dict(series=np.ones(n, np.complex64), starts=np.array([0]),
              win_start=np.array([0]), win_end=np.array([n]), binsize=n)
    if case == "no_blocks":
        kw.update(starts=[], win_start=[], win_end=[])
    elif case == "series_2d":
        kw["series"] = kw["series"].reshape(2, n // 2)
    elif case == "starts_2d":
        kw["starts"] = np.array([[0]])
    elif case == "window_2d":
        kw["win_end"] = np.array([[n]])
    elif case == "negative_start":
        kw["starts"] = np.array([-1], dtype=np.int64)
    elif case == "huge_start":
        kw["starts"] = np.array([np.iinfo(np.uintp).max], dtype=np.uintp)
    elif case == "empty_templates":
        kw["templates"] = (0, 0)
    elif case == "past_templates":
        kw["templates"] = (1, 2)
    else:
        kw["binsize"] = 0 if case == "zero_binsize" else -1
    with pytest.raises(ValueError):
        f.run_series(**kw)
    assert not [c for c in FakeContext.calls if c[0] == "series"]


def test_autotune_table_lookup(tmp_path, monkeypatch):
    """Nearest entry of the same operation and length, in log (ndata, ntemplates)."""
    import json
    from apogee_amd import autotune
    table = {"entries": [
        {"op": "flat", "n": 4096, "ndata": 16, "ntemplates": 1024, "best": {"dpw": 8}},
        {"op": "flat", "n": 4096, "ndata": 64, "ntemplates": 256, "best": {"dpw": 2}},
        {"op": "flat", "n": 1024, "ndata": 16, "ntemplates": 1024, "best": {"dpw": 16}},
        {"op": "hier", "n": 4096, "ndata": 16, "ntemplates": 1024, "best": {"refine_grid": 512}},
        {"op": "queue", "workload": "uneven", "best": {"grid_per_cu": 1}}]}
    path = tmp_path / "t.json"
    path.write_text(json.dumps(table))
    monkeypatch.setattr(autotune, "TABLE_PATH", str(path))
    monkeypatch.setattr(autotune, "_table", None)
    assert autotune.lookup("flat", 4096, 16, 2048) == {"dpw": 8}
    assert autotune.lookup("flat", 4096, 48, 256) == {"dpw": 2}
    assert autotune.lookup("flat", 1024, 1, 1) == {"dpw": 16}
    assert autotune.lookup("flat", 2048, 16, 1024) in ({"dpw": 8}, {"dpw": 16})
    assert autotune.lookup("hier", 4096, 16, 1024) == {"refine_grid": 512}
    assert autotune.lookup("queue") == {"grid_per_cu": 1}
    assert autotune.lookup("nothing") == {}

    class Sink:
        def __init__(self):
            self.got = {}

        def set_tuning(self, k, v):
            self.got[k] = v
    s = Sink()
    autotune.apply(s, "flat", 4096, 16, 1024)
    assert s.got == {"dpw": 8}
    monkeypatch.setattr(autotune, "TABLE_PATH", str(tmp_path / "missing.json"))
    monkeypatch.setattr(autotune, "_table", None)
    assert autotune.lookup("flat", 4096, 16, 1024) == {}


def synthetic_band_entry(n=4096, screen=None, refine=None):
    """A band-cost entry in the table's format: screen and all-refined
    milliseconds per band at 256 and 16384 pairs."""
    screen = screen or {64: (4.0, 30.0), 128: (4.0, 12.0), 256: (4.0, 8.0), 512: (4.0, 10.0),
                        1024: (4.0, 14.0), 2048: (4.0, 24.0)}
    refine = refine or (2.0, 100.0)
    return {"op": "band", "n": n, "ndata": 16, "ntemplates": 1024, "pairs": [256, 16384],
            "flat_ms": [2.0, 100.0],
            "bands": {str(b): {"screen_ms": list(c), "refine_ms_all": list(refine),
                               **({"screen_ms_fp16": [3.0, 5.0]} if b == 512 else {})}
                      for b, c in screen.items()}}


def test_band_ranking_weighs_screen_cost_against_the_noise_refine_rate():
    """The cheapest measured band wins; a wide refine bill pushes the choice
    to bands whose higher gate refines less; the matrix-core screen's cost
    counts only with an fp16/bf16 screen; pair count interpolates."""
    from apogee_amd.hierarchical import rank_bands
    from apogee_amd import gatemodel as G
    p = inspiral_power(4096)
    ranked = rank_bands(p, 4096, 5.5, 1e-2, 16384, entry=synthetic_band_entry())
    assert ranked and ranked == sorted(ranked, key=lambda c: c["ms"])
    for c in ranked:
        assert c["gate"] == G.gate_for(p, 4096, c["band"], 5.5, 1e-2)
        assert 0.0 <= c["noise_refine_rate"] <= 1.0
    # refine rate falls as the band (and its gate) grows
    rates = {c["band"]: c["noise_refine_rate"] for c in ranked}
    bands = sorted(rates)
    assert all(rates[a] >= rates[b] for a, b in zip(bands, bands[1:]))
    # screen costs alone decide when refining is free
    free = synthetic_band_entry(refine=(0.0, 0.0))
    assert rank_bands(p, 4096, 5.5, 1e-2, 16384, entry=free)[0]["band"] == 256
    # an expensive refine moves the choice to a band that refines less
    dear = synthetic_band_entry(refine=(1e4, 1e7))
    pick = rank_bands(p, 4096, 5.5, 1e-2, 16384, entry=dear)[0]["band"]
    assert rates[pick] <= rates[256]
    # the matrix-core screen at 512 is cheaper than every FP32 screen
    assert rank_bands(p, 4096, 5.5, 1e-2, 16384, "fp16", entry=free)[0]["band"] == 512
    # at 256 pairs every screen costs the same fixed 4 ms: the smallest refine bill wins
    small = rank_bands(p, 4096, 5.5, 1e-2, 256, entry=synthetic_band_entry(refine=(2.0, 2.0)))
    assert small[0]["noise_refine_rate"] == min(c["noise_refine_rate"] for c in small)


def test_band_choice_comes_from_the_measured_table(fake, monkeypatch):
    from apogee_amd import autotune
    monkeypatch.setattr(autotune, "_table", {"entries": [synthetic_band_entry(refine=(0.0, 0.0))]})
    h = mf.HierarchicalFilter(4096, 16, 1024, snr=5.5, fd=1e-2)
    h.set_reference(inspiral_power(4096))
    assert h.config == (256, 8)
    h.set_coarse_precision("fp16")
    assert h.config == (512, 8)
    g = mf.HierarchicalFilter(2048, 2, 2)          # no entry for this length: n / 8
    g.set_reference(inspiral_power(2048))
    assert g.config == (256, 8)


def test_noise_exceedance_model():
    from apogee_amd import gatemodel as G
    n = 4096
    white = np.ones(n)
    # white profile: every one of the band lags independent
    for band, gate in ((512, 4.0), (1024, 5.0)):
        want = 1 - (1 - np.exp(-gate * gate / 2)) ** band
        assert abs(G.noise_exceedance(white, n, band, gate) - want) < 1e-12
    # a narrower profile has fewer independent lags: it exceeds less often
    assert G.noise_exceedance(inspiral_power(n), n, 512, 4.0) < G.noise_exceedance(white, n, 512, 4.0)
    assert G.noise_exceedance(white, n, 512, 0.0) == 1.0


def test_series_layout_table_matches_a_per_block_reading():
    """The (3, nblocks) table check of run_series against a per-block reading
    of the contract: clamped windows, one bin count, the errors and their
    kinds, for uniform and mixed windows, windows past n and a binsize past n."""
    from apogee_amd._series import SeriesLayout, offsets_table
    rng = np.random.default_rng(31)
    n = 1024
    for trial in range(400):
        m = int(rng.integers(1, 9))
        st = rng.integers(0, 5000, m)
        if trial % 2:
            lo = np.full(m, int(rng.integers(0, n + 200)))
            hi = np.full(m, int(rng.integers(0, n + 400)))
        else:
            lo = rng.integers(0, n + 200, m)
            hi = lo + rng.integers(-50, 600, m)
        binsize = int(rng.choice([1, 7, 64, 300, n, 10 * n, 1 << 70]))
        dtype = [np.int64, np.uint64, np.int32][trial % 3]
        rows = [np.asarray(a, dtype) for a in (st, lo, np.maximum(hi, 0))]
        clo, chi = np.minimum(lo, n), np.minimum(np.maximum(hi, 0), n)
        ok = (clo < chi).all() and len({1 + (int(b) - int(a) - 1) // binsize for a, b in zip(clo, chi)}) == 1
        if not ok:
            with pytest.raises(ValueError):
                SeriesLayout(n, *offsets_table(rows, n), binsize)
            continue
        lay = SeriesLayout(n, *offsets_table(rows, n), binsize)
        assert lay.starts.dtype == np.uintp and lay.low.dtype == np.uintp and lay.high.dtype == np.uintp
        np.testing.assert_array_equal(lay.starts, st)
        np.testing.assert_array_equal(lay.low, clo)
        np.testing.assert_array_equal(lay.high, chi)
        assert lay.nbins == 1 + (int(chi[0]) - int(clo[0]) - 1) // binsize
    with pytest.raises(ValueError, match="one-dimensional"):
        offsets_table([np.zeros((1, 1), np.int64), np.zeros(1, np.int64), np.ones(1, np.int64)], n)
    with pytest.raises(ValueError, match="same length"):
        offsets_table([np.zeros(2, np.int64), np.zeros(1, np.int64), np.ones(1, np.int64)], n)
    with pytest.raises(ValueError, match="at least one block"):
        offsets_table([np.zeros(0, np.int64)] * 3, n)
    big = np.array([np.iinfo(np.uintp).max], np.uintp)
    for rows, what in (([np.array([-1]), np.zeros(1, np.int64), np.ones(1, np.int64)], "starts"),
                       ([big, np.zeros(1, np.int64), np.ones(1, np.int64)], "starts"),
                       ([np.array([np.iinfo(np.intp).max - n + 1]), np.zeros(1, np.int64), np.ones(1, np.int64)],
                        "starts"),
                       ([np.zeros(1, np.int64), np.array([-3]), np.ones(1, np.int64)], "window"),
                       ([np.zeros(1, np.int64), np.zeros(1, np.int64), big], "window")):
        with pytest.raises(ValueError, match=what):
            offsets_table(rows, n)


def test_result_pool_reuses_and_caps_free_buffers(monkeypatch):
    """The page-locked result pool's bookkeeping, against a stub library:
    a returned buffer serves the next request of its size, free buffers
    beyond the cap go back to the library, and a failed allocation gives
    None (the caller then uses plain NumPy arrays)."""
    from apogee_amd import _hip
    freed, addrs = [], iter(range(1 << 20, 1 << 30, 1 << 16))

    class Stub:
        fail = False

        def mfh_host_alloc(self, nbytes):
            return None if Stub.fail else next(addrs)

        def mfh_host_free(self, p):
            freed.append(p.value)
            return 0

    monkeypatch.setattr(_hip, "lib", lambda: Stub())
    pool = _hip._HostPool(keep_bytes=100)
    a, b = pool.take(64), pool.take(64)
    assert a != b
    pool.give(a, 64)
    assert pool.take(64) == a and pool.free_bytes == 0
    pool.give(a, 64)
    pool.give(b, 64)                      # 128 free bytes would pass the cap
    assert freed == [b] and pool.free_bytes == 64
    assert pool.take(32) not in (a, b)    # another size: a fresh buffer
    Stub.fail = True
    assert pool.take(48) is None
