"""RefineQueue's host logic on a host double of the device queue: staging,
plans over a shared pool, per-request result blocks and validation.  Every
request must read exactly like MatchedFilter(n, 1, c).run on its spectra."""
import numpy as np
import pytest

import apogee_amd as mf
from apogee_amd import device as mfdev
from apogee_amd import refine as mfrefine
from oracle import mf_oracle as O
from _fake import FakeQueue


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(mfrefine.RefineQueue, "_queue_factory", FakeQueue)
    monkeypatch.setattr(mfdev, "devices", lambda: [mfdev.Device("gpu", 0, "fake gfx950")])
    return FakeQueue


def cnoise(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_staged_requests_match_the_flat_oracle(fake):
    rng = np.random.default_rng(1)
    q = mf.RefineQueue(device="gpu")
    reqs = [(64, 3, 16, 0.0, None), (256, 1, 100, 4.0, (10, 250)), (128, 2, None, 0.0, (0, 128)),
            (512, 0, 64, 0.0, None), (64, 5, 7, 2.5, (3, 61))]
    spectra = []
    for r, (n, c, bs, thr, win) in enumerate(reqs):
        D, T = cnoise(rng, n), cnoise(rng, (c, n))
        spectra.append((D, T))
        assert q.add(D, T, binsize=bs, threshold=thr, window=win) == r
    assert len(q) == len(reqs)
    out = q.run()
    assert len(out) == len(reqs)
    for (n, c, bs, thr, win), (D, T), got in zip(reqs, spectra, out):
        b = n if bs is None else bs
        w = (0, n) if win is None else win
        nb = O.nbins(n, b, *w)
        assert got.dtype == mf.PEAK_DTYPE and got.shape == (c, nb)
        if c:
            i, v, _ = O.flat_peaks(D[None], T, b, thr, w)
            np.testing.assert_array_equal(got["index"], i[0])
            np.testing.assert_array_equal(got["value"], v[0].astype(np.complex64))
    raw = q.run(raw=True)
    np.testing.assert_array_equal(raw[1][0], out[1]["index"])


def test_plan_over_a_shared_pool_with_broadcast_arguments(fake):
    rng = np.random.default_rng(2)
    pool = cnoise(rng, 4096)
    q = mf.RefineQueue()
    # three requests share the data spectrum at 0; templates overlap freely
    q.plan(pool, n=[256, 256, 1024], data_offset=[0, 0, 1024], template_offset=[256, 512, 2048],
           candidates=[2, 1, 2], binsize=64, threshold=0.0, window=(0, 256))
    assert q.total == 2 * 4 + 1 * 4 + 2 * 4
    np.testing.assert_array_equal(q.offsets, [0, 8, 12])
    out = q.run()
    i, v, _ = O.flat_peaks(pool[None, :256], pool[256:768].reshape(2, 256), 64, 0.0, (0, 256))
    np.testing.assert_array_equal(out[0]["index"], i[0])
    i, v, _ = O.flat_peaks(pool[None, 1024:2048], pool[2048:4096].reshape(2, 1024), 64, 0.0, (0, 256))
    np.testing.assert_array_equal(out[2]["index"], i[0])


def test_validation(fake):
    rng = np.random.default_rng(3)
    q = mf.RefineQueue()
    with pytest.raises(ValueError):
        q.run()                                   # nothing staged
    with pytest.raises(ValueError):
        q.add(cnoise(rng, 100), cnoise(rng, (1, 100)))   # not a power of two
    with pytest.raises(ValueError):
        q.add(cnoise(rng, 64), cnoise(rng, (1, 128)))    # length mismatch
    with pytest.raises(ValueError):
        q.add(cnoise(rng, 64), cnoise(rng, (1, 64)), binsize=0)
    with pytest.raises(ValueError):
        q.add(cnoise(rng, 64), cnoise(rng, (1, 64)), window=(40, 40))
    with pytest.raises(ValueError):
        q.add(cnoise(rng, (2, 64)), cnoise(rng, (1, 64)))
    with pytest.raises(ValueError):
        q.add(cnoise(rng, 1 << 21), cnoise(rng, (1, 1 << 21)))
    pool = cnoise(rng, 1024)
    with pytest.raises(ValueError):
        q.plan(pool, n=[256], data_offset=[900], template_offset=[0], candidates=[1])
    with pytest.raises(ValueError):
        q.plan(pool, n=[256], data_offset=[0], template_offset=[512], candidates=[3])
    with pytest.raises(ValueError):
        q.plan(pool, n=[256, 256], data_offset=[0, 0], template_offset=[0], candidates=1)
    with pytest.raises(ValueError):
        q.plan(pool, n=[256], data_offset=[0], template_offset=[0], candidates=[-1])
    with pytest.raises(ValueError):
        q.plan(pool.reshape(4, 256), n=[256], data_offset=[0], template_offset=[0], candidates=[1])
    # the window clamps like MatchedFilter.run's
    q.plan(pool, n=[256], data_offset=[0], template_offset=[256], candidates=[1], window=(-5, 9999),
           binsize=128)
    assert q.total == 2


def test_add_after_run_replans(fake):
    rng = np.random.default_rng(4)
    q = mf.RefineQueue()
    q.add(cnoise(rng, 64), cnoise(rng, (2, 64)))
    assert len(q.run()) == 1
    q.add(cnoise(rng, 128), cnoise(rng, (1, 128)), binsize=32)
    out = q.run()
    assert [o.shape for o in out] == [(2, 1), (1, 4)]
    q.clear()
    assert len(q) == 0


def test_uneven_workload_recipe():
    """The benchmark's uneven-refine plan (autotune.uneven_workload) on the
    host: lengths from the four classes, binsize n/8, offsets that tile the
    pool exactly."""
    from apogee_amd.autotune import uneven_workload
    pool, plan = uneven_workload(200, 5, device="cpu")
    n, cand = plan["n"], plan["candidates"]
    assert set(np.unique(n)) <= {1024, 4096, 16384, 65536} and (cand >= 1).all()
    assert (plan["binsize"] == n // 8).all()
    assert (plan["template_offset"] == plan["data_offset"] + n).all()
    assert int(pool.numel()) == int((n * (1 + cand)).sum())
    assert (plan["data_offset"][1:] == plan["template_offset"][:-1] + cand[:-1] * n[:-1]).all()
