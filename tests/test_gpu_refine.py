"""The persistent mixed-length refine queue on the MI355X.

Contract: every (request, candidate) pair reports exactly what the
per-request flat filter (mfh_flat_peaks through a Context) reports on the same
spectra -- identical indices and identical complex64 values -- at every
length 64 ... 65536, for aligned and ragged bins, windows, thresholds, window
pieces of the four-step lengths and zero-candidate requests; graph replays
repeat the eager result; and small requests agree with the FP64 oracle.
"""
import numpy as np
import pytest

import apogee_amd as mf
from apogee_amd import _hip
from _parity import assert_peaks_match

pytestmark = pytest.mark.gpu

LENGTHS = [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536]


def cnoise(rng, shape):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)).astype(np.complex64)


class PerRequest:
    """The per-request flat path: one Context per length, rows reloaded per request."""

    def __init__(self, device):
        self.device = device
        self.ctx = {}

    def peaks(self, D, T, binsize, threshold, ws, we):
        n, c = D.size, T.shape[0]
        ctx = self.ctx.get(n)
        if ctx is None or ctx.ntemplates < c:
            ctx = _hip.Context(self.device, n, 1, max(c, 16))
            self.ctx[n] = ctx
        ctx.set_rows("data", 0, D[None])
        ctx.set_rows("template", 0, T)
        idx, val, _ = ctx.peaks(0, 1, 0, c, binsize, threshold, ws, we)
        return idx[0], val[0]


def check_against_per_request(reqs, out, device, what):
    ref = PerRequest(device)
    for r, ((D, T, bs, thr, w), got) in enumerate(zip(reqs, out)):
        n = D.size
        ws, we = (0, n) if w is None else w
        b = n if bs is None else bs
        if T.shape[0] == 0:
            assert got.shape[0] == 0
            continue
        i, v = ref.peaks(D, T, b, thr, ws, we)
        ctx = "%s request %d: n=%d c=%d binsize=%d thr=%g window=(%d,%d)" % (
            what, r, n, T.shape[0], b, thr, ws, we)
        np.testing.assert_array_equal(got["index"], i, err_msg=ctx)
        np.testing.assert_array_equal(got["value"], v, err_msg=ctx)


def stage(q, reqs):
    for D, T, bs, thr, w in reqs:
        q.add(D, T, binsize=bs, threshold=thr, window=w)


def test_every_length_aligned_bins(gpu_device):
    rng = np.random.default_rng(20261206)
    reqs = []
    for n in LENGTHS:
        for c in (1, 3):
            D, T = cnoise(rng, n), cnoise(rng, (c, n))
            # a planted lag per template so bins hold real peaks and noise
            for j in range(c):
                lag = int(rng.integers(n))
                D = D + (4.0 * T[j] * np.exp(2j * np.pi * np.arange(n) * lag / n)).astype(np.complex64)
            reqs.append((D.astype(np.complex64), T, n // 8, 0.0, None))
    q = mf.RefineQueue(device="gpu")
    stage(q, reqs)
    check_against_per_request(reqs, q.run(), gpu_device, "aligned")


def test_ragged_bins_windows_thresholds_and_pieces(gpu_device):
    rng = np.random.default_rng(20261207)
    cases = [
        (64, 2, 7, 0.0, (3, 61)), (128, 1, 1, 0.0, None), (256, 4, 100, 1.5, (10, 250)),
        (512, 2, 33, 0.0, (0, 500)), (1024, 3, 96, 2.0, None), (2048, 1, 3, 0.0, (1000, 1100)),
        (4096, 2, 4096, 0.0, None), (4096, 2, 1000, 3.0, (37, 4000)), (8192, 1, 257, 0.0, None),
        (16384, 2, 64, 0.0, (100, 16300)), (32768, 1, 1024, 0.0, None), (32768, 2, 5, 2.0, (7, 9000)),
        (65536, 1, 16, 0.0, None),          # 4096 bins: window pieces of the four-step
        (65536, 2, 1000, 4.0, (123, 65000)), (65536, 1, 65536, 0.0, (0, 65536)),
        (262144, 2, 32768, 0.0, None), (1048576, 1, 4096, 0.0, (100, 1048000)),
        (1024, 0, 64, 0.0, None),           # no candidates
        (4096, 1, 64, 1000.0, None),        # nothing crosses (|z| ~ 64 on this noise)
    ]
    reqs = [(cnoise(rng, n), cnoise(rng, (c, n)), bs, thr, w) for n, c, bs, thr, w in cases]
    q = mf.RefineQueue()
    stage(q, reqs)
    out = q.run()
    check_against_per_request(reqs, out, gpu_device, "ragged")
    assert (out[-1]["index"] == -1).all() and (out[-1]["value"] == 0).all()


def config4_requests(nreq, seed):
    """The uneven-refine workload: n in {1024, 4096, 16384, 65536} with
    weights .4/.3/.2/.1, binsize n/8, candidates ~ geometric(1/3)."""
    rng = np.random.default_rng(seed)
    ns = rng.choice([1024, 4096, 16384, 65536], size=nreq, p=[0.4, 0.3, 0.2, 0.1])
    cand = rng.geometric(1.0 / 3.0, size=nreq)
    reqs = []
    for n, c in zip(ns, cand):
        reqs.append((cnoise(rng, int(n)), cnoise(rng, (int(c), int(n))), int(n) // 8, 0.0, None))
    return reqs


def test_uneven_refine_workload_sample(gpu_device):
    reqs = config4_requests(400, 20261204)
    q = mf.RefineQueue()
    stage(q, reqs)
    out = q.run()
    print("uneven refine: %d requests, %d pairs" % (len(reqs), sum(r[1].shape[0] for r in reqs)))
    check_against_per_request(reqs, out, gpu_device, "uneven")


def test_graph_replays_repeat_the_eager_result(gpu_device):
    import torch
    reqs = config4_requests(60, 5)
    q = mf.RefineQueue()
    stage(q, reqs)
    first = [o.copy() for o in q.run()]            # eager
    x = torch.randn(1 << 20, device="cuda")
    for rep in range(4):                              # capture, then replays
        _ = (x * 2).sum().item()                      # the program's own work between runs
        again = q.run()
        for a, b in zip(first, again):
            np.testing.assert_array_equal(a["index"], b["index"], err_msg="replay %d" % rep)
            np.testing.assert_array_equal(a["value"], b["value"], err_msg="replay %d" % rep)
    q.set_tuning("graphs", 0)
    plain = q.run()
    for a, b in zip(first, plain):
        np.testing.assert_array_equal(a["index"], b["index"])


def test_device_pool_async_outputs_and_in_place_updates(gpu_device):
    import torch
    rng = np.random.default_rng(9)
    n = np.array([1024, 65536, 4096, 1024, 16384], dtype=np.int64)
    cand = np.array([3, 1, 2, 5, 1], dtype=np.int64)
    sizes = n * (1 + cand)
    doff = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    toff = doff + n
    pool_h = cnoise(rng, int(sizes.sum()))
    pool = torch.from_numpy(pool_h).to("cuda")
    q = mf.RefineQueue()
    q.plan(pool, n, doff, toff, cand, binsize=n // 8)
    ref = q.run()
    idx = torch.empty(q.total, dtype=torch.int64, device="cuda")
    val = torch.empty(q.total, dtype=torch.complex64, device="cuda")
    for _ in range(3):
        q.run_async(idx, val)
    q.sync()
    got_i, got_v = idx.cpu().numpy(), val.cpu().numpy()
    for r, o in enumerate(q.offsets):
        k = ref[r].size
        np.testing.assert_array_equal(got_i[o:o + k], ref[r]["index"].reshape(-1))
        np.testing.assert_array_equal(got_v[o:o + k], ref[r]["value"].reshape(-1))
    # the plan reads the pool in place: new contents, same plan, same graph
    # (negating the data spectra negates every peak value, same lags)
    for d0, m in zip(doff, n):
        pool[int(d0):int(d0 + m)].mul_(-1.0)
    torch.cuda.synchronize()
    flipped = q.run()
    for a, b in zip(ref, flipped):
        np.testing.assert_array_equal(a["index"], b["index"])
        np.testing.assert_array_equal(a["value"], -b["value"])


def test_small_requests_against_the_fp64_oracle(gpu_device):
    rng = np.random.default_rng(12)
    q = mf.RefineQueue()
    reqs = []
    for n, c, bs, w in [(64, 4, 8, None), (256, 3, 50, (5, 250)), (1024, 2, 128, None),
                        (4096, 2, 512, (100, 4000))]:
        D, T = cnoise(rng, n), cnoise(rng, (c, n))
        reqs.append((D, T, bs, w))
        q.add(D, T, binsize=bs, window=w)
    for (D, T, bs, w), got in zip(reqs, q.run()):
        assert_peaks_match(D[None], T, got["index"][None], got["value"][None], bs, 0.0, w,
                           context="n=%d" % D.size)


def test_plan_errors_leave_the_queue_usable(gpu_device):
    rng = np.random.default_rng(3)
    q = mf.RefineQueue()
    q.add(cnoise(rng, 1024), cnoise(rng, (2, 1024)))
    good = q.run()
    pool = cnoise(rng, 4096)
    with pytest.raises(ValueError):
        q.plan(pool, n=[4096], data_offset=[0], template_offset=[0], candidates=[2])
    q.plan(pool, n=[1024], data_offset=[0], template_offset=[1024], candidates=[3])
    out = q.run()
    assert out[0].shape == (3, 1) and good[0].shape == (2, 1)
