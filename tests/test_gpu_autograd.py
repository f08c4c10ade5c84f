"""Differentiable peak values on the MI355X.

* The backward kernel against the FP64 oracle (oracle.mf_oracle.peak_grads,
  itself pinned by finite differences and PyTorch's complex autograd) for
  arbitrary incoming gradients, empty bins included.
* The differentiable peak fit: parametric chirp templates theta = (A, t0, a,
  b), injections of perturbed theta at SNR 10 in unit noise, loss
  -sum |z_peak|^2 / sigma^2 over 512-lag bins.  d loss / d theta against FP64
  central finite differences (step 1e-4, one Richardson step) and against a torch.fft autograd
  reference in complex128, both at the stored lags, to 1e-3 relative.
* The forward equals MatchedFilter.run; recomputed values equal the forward.
"""
import math

import numpy as np
import pytest

import apogee_amd as mf
from oracle import mf_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

F_LO, F_HI = 20, 1500


def cnoise(rng, shape):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2)).astype(np.complex64)


def test_backward_matches_the_fp64_oracle(gpu_device):
    from apogee_amd.autograd import PeakFilter
    rng = np.random.default_rng(31)
    n, nd, nt, bs = 1024, 5, 7, 128
    D = cnoise(rng, (nd, n))
    T = cnoise(rng, (nt, n))
    pf = PeakFilter(n, nd, nt, device="gpu")
    Dg = torch.from_numpy(D).cuda().requires_grad_(True)
    Tg = torch.from_numpy(T).cuda().requires_grad_(True)
    val, idx = pf(Dg, Tg, binsize=bs, threshold=73.0)     # about half the bins stay empty
    idx_h = idx.cpu().numpy()
    assert (idx_h < 0).any() and (idx_h >= 0).any()
    g = cnoise(rng, val.shape)
    val.backward(torch.from_numpy(g).cuda())
    gD, gT = O.peak_grads(D, T, idx_h, g)
    for got, ref, what in ((Dg.grad, gD, "data"), (Tg.grad, gT, "templates")):
        got = got.cpu().numpy().astype(np.complex128)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-5, (what, err)


def test_forward_is_the_flat_filter_and_recompute_agrees(gpu_device):
    from apogee_amd.autograd import PeakFilter
    rng = np.random.default_rng(32)
    n, nd, nt = 4096, 4, 6
    D, T = cnoise(rng, (nd, n)), cnoise(rng, (nt, n))
    pf = PeakFilter(n, nd, nt)
    Dg, Tg = torch.from_numpy(D).cuda(), torch.from_numpy(T).cuda()
    val, idx = pf(Dg, Tg, binsize=512, window=(17, 4000))
    f = mf.MatchedFilter(n, nd, nt)
    f.set_data(D)
    f.set_templates(T)
    ref = f.run(binsize=512, window=(17, 4000))
    np.testing.assert_array_equal(idx.cpu().numpy(), ref["index"])
    np.testing.assert_array_equal(val.cpu().numpy(), ref["value"])
    again = pf.recompute(Dg, Tg, idx).cpu().numpy()
    rel = np.abs(again - ref["value"]) / np.abs(ref["value"]).max()
    assert rel.max() < 2e-6 * math.log2(n) * 4


# ------------------------------------------------------ the peak fit workload

def chirp(theta, n, xp):
    """A * f^(-7/6) * exp(i (2 pi f t0 / n + a f^(-5/3) + b f)) on bins
    [F_LO, F_HI], zero elsewhere; theta (nt, 4) -> (nt, n), in theta's float
    type promoted to FP64 (the generator is caller code, not the filter)."""
    if xp is np:
        f = np.arange(n, dtype=np.float64)
        band = (f >= F_LO) & (f <= F_HI)
        fs = np.where(band, f, 1.0)
        A, t0, a, b = (theta[:, i:i + 1] for i in range(4))
        amp = A * fs ** (-7.0 / 6.0) * band
        ph = 2 * np.pi * f * t0 / n + a * fs ** (-5.0 / 3.0) + b * f
        return amp * np.exp(1j * ph)
    th = theta.to(torch.float64)
    f = torch.arange(n, dtype=torch.float64, device=theta.device)
    band = (f >= F_LO) & (f <= F_HI)
    fs = torch.where(band, f, torch.ones_like(f))
    A, t0, a, b = (th[:, i:i + 1] for i in range(4))
    amp = A * fs ** (-7.0 / 6.0) * band.to(torch.float64)
    ph = 2 * math.pi * f * t0 / n + a * fs ** (-5.0 / 3.0) + b * f
    return torch.polar(amp, ph)


def fit_case(seed=20261205, n=4096, nd=64, nt=32, snr=10.0):
    rng = np.random.default_rng(seed)
    theta = np.stack([rng.uniform(0.5, 2.0, nt), rng.uniform(0, n, nt),
                      rng.uniform(1e3, 1e4, nt), rng.uniform(-1, 1, nt)], axis=1).astype(np.float32)
    pert = theta.copy()
    src = np.arange(nd) % nt
    inj = pert[src].astype(np.float64)
    inj[:, 1] += rng.uniform(-3, 3, nd)
    inj[:, 2] *= 1 + rng.uniform(-2e-3, 2e-3, nd)
    inj[:, 3] += rng.uniform(-0.02, 0.02, nd)
    H = chirp(inj, n, np)
    sig2 = n * (np.abs(H) ** 2).sum(axis=1)
    noise = np.fft.fft(rng.standard_normal((nd, n)), axis=1)
    D = (noise + (snr * n / np.sqrt(sig2))[:, None] * H).astype(np.complex64)
    return theta, D


def loss_at_lags_np(theta64, D, idx, n):
    """-sum |z|^2 / sigma_t^2 at fixed lags, FP64 (z = n ifft(D conj T))."""
    T = chirp(theta64, n, np)
    sig2 = n * (np.abs(T) ** 2).sum(axis=1)
    z = n * np.fft.ifft(D.astype(np.complex128)[:, None, :] * np.conj(T)[None], axis=-1)
    k = np.where(idx >= 0, idx, 0)
    zk = np.take_along_axis(z, k, axis=-1)
    return -np.sum(np.where(idx >= 0, np.abs(zk) ** 2, 0.0) / sig2[None, :, None])


def test_peak_fit_gradients(gpu_device):
    from apogee_amd.autograd import PeakFilter
    n, nd, nt, bs = 4096, 64, 32, 512
    theta, D = fit_case(n=n, nd=nd, nt=nt)
    pf = PeakFilter(n, nd, nt)
    th = torch.from_numpy(theta).cuda().requires_grad_(True)
    T64 = chirp(th, n, torch)
    sig2 = n * (T64.abs() ** 2).sum(dim=1)
    val, idx = pf(torch.from_numpy(D).cuda(), T64.to(torch.complex64), binsize=bs)
    loss = -((val.to(torch.complex128).abs() ** 2) / sig2[None, :, None]).sum()
    loss.backward()
    g = th.grad.cpu().numpy().astype(np.float64)
    idx_h = idx.cpu().numpy()
    assert (idx_h >= 0).all()

    # torch.fft autograd reference, complex128 on the host, same lags
    th_r = torch.from_numpy(theta.astype(np.float64)).requires_grad_(True)
    T_r = chirp(th_r, n, torch)
    sig2_r = n * (T_r.abs() ** 2).sum(dim=1)
    z_r = n * torch.fft.ifft(torch.from_numpy(D.astype(np.complex128))[:, None, :] * torch.conj(T_r)[None], dim=-1)
    zk = torch.gather(z_r, 2, torch.from_numpy(idx_h))
    loss_r = -((zk.abs() ** 2) / sig2_r[None, :, None]).sum()
    loss_r.backward()
    g_r = th_r.grad.numpy()
    assert abs(loss.item() - loss_r.item()) <= 1e-5 * abs(loss_r.item())

    # A only rescales the template, and sigma with it: its exact gradient is 0
    scale_A = np.abs(loss_r.item()) / nt / np.abs(theta[:, 0]).min()
    assert np.abs(g[:, 0]).max() <= 1e-4 * scale_A
    assert np.abs(g_r[:, 0]).max() <= 1e-9 * scale_A
    for p, name in ((1, "t0"), (2, "a"), (3, "b")):
        rel = np.linalg.norm(g[:, p] - g_r[:, p]) / np.linalg.norm(g_r[:, p])
        assert rel <= 1e-3, (name, rel)

    # FP64 central differences at the stored lags, base step 1e-4 with one
    # Richardson step (h and h/2): b multiplies f up to 1500, so a plain
    # central difference at 1e-4 keeps an O((1500 h)^2) = 0.4% bias
    h = 1e-4
    base = theta.astype(np.float64)

    def central(t, p, step):
        sl = slice(t, t + 1)
        tp, tm = base[sl].copy(), base[sl].copy()
        tp[0, p] += step
        tm[0, p] -= step
        return (loss_at_lags_np(tp, D, idx_h[:, sl], n) - loss_at_lags_np(tm, D, idx_h[:, sl], n)) / (2 * step)

    for t in (0, 7, 19):
        for p in (1, 2, 3):
            fd = (4 * central(t, p, h / 2) - central(t, p, h)) / 3
            scale = np.abs(g_r[:, p]).max()
            assert abs(g[t, p] - fd) <= 1e-3 * abs(fd) + 1e-5 * scale, (t, p, g[t, p], fd)
            assert abs(g_r[t, p] - fd) <= 1e-5 * abs(fd) + 1e-7 * scale, (t, p, g_r[t, p], fd)
