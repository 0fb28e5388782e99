"""GPU parity of the SOCS forward / adjoint / CTM grad-eval against the FP64
oracle, through the C ABI (include/litho_amd.h).

Tolerances (stated per result): aerial image relative L-inf <= 1e-5
(max |I_gpu - I_ref| / max |I_ref|), wafer image absolute L-inf <= 1e-4,
gradient relative L2 <= 1e-4, PE relative <= 1e-5.
"""
import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import ctm as oc
from oracle import imaging as im
from oracle import optics as op

pytestmark = pytest.mark.gpu

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)
_KCACHE = {}


def socs(n, k):
    key = (n, k)
    if key not in _KCACHE:
        _KCACHE[key] = op.socs_kernels(n, **OPT, num_kernels=k)
    return _KCACHE[key]


class Synth:
    """The reference tests' closed-form 2-kernel cache as a full-grid window."""

    def __init__(self, n):
        c = im.Cache.synthetic(n)
        self.cache = c
        self.n = n
        self.kernels = np.stack([np.fft.fftshift(h) for h in c.kernels])
        self.weights = c.weights
        self.weight_sum = c.weight_sum
        self.win_y0 = self.win_x0 = 0


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def empty(n, b=1):
    import torch
    return torch.zeros((b, n, n), dtype=torch.float32, device="cuda")


def host(t):
    return t.detach().cpu().double().numpy()


def rel_linf(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


def rel_l2(a, ref):
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


def window_kernels(n, k):
    """Random complex kernels on the true SOCS window of the optics at size n:
    the forward / adjoint identities hold for any kernels, so these pin the
    pruned transforms at tile sizes whose SVD would be slow here."""
    r, c, npts, wsum, y0, x0 = op.source_summary(n, **OPT)
    rng = np.random.default_rng(n + k)
    kern = rng.standard_normal((k, r, c)) + 1j * rng.standard_normal((k, r, c))
    w = np.sort(rng.random(k) + 0.05)[::-1] * wsum / k
    # scaled so that a clear mask images to I = sum_k w_k |H_k(DC)|^2 / sum w = 1, as the optics do
    dc = kern[:, n // 2 - y0, n // 2 - x0]
    kern = kern / np.sqrt(np.sum(w * np.abs(dc) ** 2) / wsum)
    return op.SocsKernels(n=n, win_y0=y0, win_x0=x0, kernels=kern, weights=w, weight_sum=wsum)


def wide_window_kernels(n, k, w):
    """Random complex kernels on a centred w x w window: w = 300 at n = 640
    needs correlations of length L = 1024 >= 2 w - 1 in the band-limited
    column path (the true optics windows of the supported tiles stay <= 512)."""
    rng = np.random.default_rng(n + k + w)
    y0 = x0 = n // 2 - w // 2
    kern = rng.standard_normal((k, w, w)) + 1j * rng.standard_normal((k, w, w))
    wts = np.sort(rng.random(k) + 0.05)[::-1]
    wsum = float(np.sum(wts))
    dc = kern[:, n // 2 - y0, n // 2 - x0]
    kern = kern / np.sqrt(np.sum(wts * np.abs(dc) ** 2) / wsum)
    return op.SocsKernels(n=n, win_y0=y0, win_x0=x0, kernels=kern, weights=wts, weight_sum=wsum)


def make_case(kind, n, k=24, seed=0):
    if kind == "wide":
        s = wide_window_kernels(n, k, 300)
        cache = im.Cache.from_socs(s)
    elif kind == "synthetic":
        s = Synth(n)
        cache = s.cache
    elif kind == "window":
        s = window_kernels(n, k)
        cache = im.Cache.from_socs(s)
    else:
        s = socs(n, k)
        cache = im.Cache.from_socs(s)
    rng = np.random.default_rng(seed)
    tgt = targets().manhattan_target(n, seed) if n >= 64 else (rng.random((n, n)) > 0.6).astype(np.float32)
    return s, cache, tgt


CASES = [("synthetic", 20, 2), ("synthetic", 24, 2), ("socs", 64, 8), ("socs", 128, 24), ("socs", 512, 24),
         ("window", 384, 6), ("window", 640, 6), ("window", 896, 6), ("window", 1024, 6), ("window", 1152, 6),
         ("window", 2176, 4),
         ("wide", 640, 4)]


@pytest.mark.parametrize("kind,n,k", CASES)
def test_forward_parity(gpu_ctx, kind, n, k):
    print(f"case forward {kind} n={n} k={k}", flush=True)
    la = litho_amd()
    s, cache, tgt = make_case(kind, n, k)
    plan = la.Plan.from_socs(gpu_ctx, s)
    rng = np.random.default_rng(7)
    mask = np.clip(tgt + 0.3 * rng.standard_normal((n, n)), 0, 1).astype(np.float32)
    a, w = empty(n), empty(n)
    plan.forward(dev(mask), a, w, threshold=0.225, alpha=50.0)
    ra, rw, _ = im.forward(cache, mask.astype(np.float64), 0.225, 50.0)
    assert rel_linf(host(a)[0], ra) <= 1e-5
    assert np.max(np.abs(host(w)[0] - rw)) <= 1e-4
    plan.close()


@pytest.mark.parametrize("kind,n,k", CASES)
def test_gradient_parity(gpu_ctx, kind, n, k):
    print(f"case gradient {kind} n={n} k={k}", flush=True)
    la = litho_amd()
    s, cache, tgt = make_case(kind, n, k)
    plan = la.Plan.from_socs(gpu_ctx, s)
    theta = oc.initialize_theta(tgt)
    mask = im.mask_from_theta(theta).astype(np.float32)
    alpha, thr = (8.0, 0.25) if kind == "synthetic" else (50.0, 0.225)
    g = empty(n)
    pe, bg = plan.pe_gradient(dev(mask), dev(tgt), grad_mask=g, threshold=thr, alpha=alpha)
    _, rw, fields = im.forward(cache, mask.astype(np.float64), thr, alpha)
    rg = im.pe_gradient(cache, rw, tgt.astype(np.float64), fields, alpha)
    rpe = im.pe_loss(rw, tgt)
    assert rel_l2(host(g)[0], rg) <= 1e-4
    assert abs(pe[0] - rpe) <= 1e-5 * max(rpe, 1.0)
    plan.close()


def test_windowed_and_penalty_gradient(gpu_ctx):
    print("case windowed+penalty n=64", flush=True)
    la = litho_amd()
    n = 64
    s, cache, tgt = make_case("socs", n, 8)
    plan = la.Plan.from_socs(gpu_ctx, s)
    rng = np.random.default_rng(11)
    mask = (0.35 + 0.3 * rng.random((n, n))).astype(np.float32)
    pw = np.zeros((n, n), np.float32)
    pw[10:50, 12:52] = (tgt[10:50, 12:52] > 0.5)
    bw = np.zeros((n, n), np.float32)
    bw[10:50, 12:52] = (tgt[10:50, 12:52] < 0.5)
    g = empty(n)
    pe, bg = plan.pe_gradient(dev(mask), dev(tgt), grad_mask=g, pe_window=dev(pw), penalty_window=dev(bw),
                              pe_weight=0.6, penalty=2.0, penalty_threshold=0.01, threshold=0.25, alpha=20.0)
    _, rw, fields = im.forward(cache, mask.astype(np.float64), 0.25, 20.0)
    t64 = tgt.astype(np.float64)
    rg = im.pe_gradient_with_penalty(cache, rw, t64, fields, 20.0, 2.0, 0.01, pw.astype(float), 0.6,
                                     bw.astype(float))
    assert rel_l2(host(g)[0], rg) <= 1e-4
    assert abs(pe[0] - im.pe_loss(rw, t64, pw)) <= 1e-5 * max(1.0, pe[0])
    assert abs(bg[0] - im.background_penalty_loss(rw, t64, 0.01, bw)) <= 1e-5 * max(1.0, bg[0])
    plan.close()


@pytest.mark.parametrize("n", [64, 512])
def test_ctm_eval_parity(gpu_ctx, n):
    print(f"case ctm_eval n={n}", flush=True)
    la = litho_amd()
    s, cache, tgt = make_case("socs", n, 24 if n > 64 else 8)
    plan = la.Plan.from_socs(gpu_ctx, s)
    theta = oc.initialize_theta(tgt) + 0.2 * np.random.default_rng(1).standard_normal((n, n))
    m, gt = empty(n), empty(n)
    pe, gn2 = plan.ctm_eval(dev(theta), dev(tgt), mask=m, grad_theta=gt, threshold=0.225, alpha=50.0)
    th32 = dev(theta).double().cpu().numpy()
    rpe, rgt, rmask, *_ = im.ctm_evaluate(cache, th32, tgt.astype(np.float64), 0.225, 50.0)
    assert rel_l2(host(gt)[0], rgt) <= 1e-4
    assert abs(pe[0] - rpe) <= 1e-5 * rpe
    assert abs(np.sqrt(gn2[0]) - np.linalg.norm(rgt)) <= 1e-4 * np.linalg.norm(rgt)
    assert np.max(np.abs(host(m)[0] - rmask)) <= 1e-6
    plan.close()


def test_batched_equals_single(gpu_ctx):
    print("case batch n=128 b=3", flush=True)
    import torch
    la = litho_amd()
    n = 128
    s, cache, _ = make_case("socs", n, 24)
    plan1 = la.Plan.from_socs(gpu_ctx, s, capacity=1)
    plan3 = la.Plan.from_socs(gpu_ctx, s, capacity=3)
    tg = np.stack([targets().manhattan_target(n, 100 + b) for b in range(3)])
    th = oc.initialize_theta(tg).astype(np.float32)
    gt3 = empty(n, 3)
    pe3, gn3 = plan3.ctm_eval(dev(th), dev(tg), grad_theta=gt3, batch=3, threshold=0.225, alpha=50.0)
    for b in range(3):
        g1 = empty(n)
        pe1, gn1 = plan1.ctm_eval(dev(th[b]), dev(tg[b]), grad_theta=g1, threshold=0.225, alpha=50.0)
        assert pe1[0] == pe3[b] and gn1[0] == gn3[b]
        assert torch.equal(g1[0], gt3[b])
    plan1.close()
    plan3.close()


def test_argument_validation(gpu_ctx):
    la = litho_amd()
    s = socs(64, 8)
    with pytest.raises(la.LithoError) as e:
        la.Plan(gpu_ctx, 64, s.kernels, -s.weights, s.weight_sum, s.win_y0, s.win_x0)
    assert e.value.code == la.EINVAL
    with pytest.raises(la.LithoError) as e:
        la.Plan(gpu_ctx, 64, s.kernels, s.weights, 0.0, s.win_y0, s.win_x0)
    assert e.value.code == la.EINVAL
    with pytest.raises(la.LithoError) as e:   # no compiled FFT and beyond the direct-DFT path's 4096
        la.Plan(gpu_ctx, 5000, s.kernels, s.weights, s.weight_sum, s.win_y0, s.win_x0)
    assert e.value.code == la.EUNSUPPORTED
    plan = la.Plan.from_socs(gpu_ctx, s, capacity=1)
    with pytest.raises(la.LithoError):
        plan.forward(dev(np.zeros((2, 64, 64))), empty(64, 2), batch=2)
    plan.close()
