"""Tile sizes without compiled transforms: the direct-DFT pipeline
(generic_ops.hip) on the GPU.

The reference images any grid size (its finite-difference test runs N = 65,
test_ctm_optimizer.cpp:164; a 3 x 3 partition of a 1000^2 chip gives 462 px
tiles).  Plans of sizes outside the compiled list run the same stages with
direct window-restricted DFTs.
* Forward and gradient parity against the FP64 oracle with the optics' own
  SOCS kernels at N = 65 (odd), 100 and 462, and random window kernels at
  N = 1000: aerial relative L-inf <= 1e-5, gradient relative L2 <= 1e-4,
  PE relative <= 1e-5.
* At a compiled size (640) the direct-DFT pipeline (LITHO_SOCS_PATH=generic)
  agrees with the FFT pipeline to FP32 rounding: aerial <= 1e-5 relative
  L-inf, gradient <= 1e-5 relative L2.
* CTM gradient descent at N = 65, 10 iterations, lr 0.9: PE history within
  1e-4 of the oracle per iteration.
* A batch of three tiles gives each tile its single-tile result bit for bit.
"""
import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import ctm as oc
from oracle import imaging as im
from oracle import optics as op
from test_gpu_socs import dev, empty, host, rel_l2, rel_linf, window_kernels

pytestmark = pytest.mark.gpu

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)


def case(n, k):
    s = op.socs_kernels(n, **OPT, num_kernels=k) if n < 1000 else window_kernels(n, k)
    tgt = targets().manhattan_target(n, 5)
    rng = np.random.default_rng(n)
    mask = np.clip(im.mask_from_theta(oc.initialize_theta(tgt)) + 0.2 * rng.standard_normal((n, n)), 0, 1)
    return s, tgt, mask.astype(np.float32)


def run(plan, n, mask, tgt, batch=1):
    a, w, g = empty(n, batch), empty(n, batch), empty(n, batch)
    m = dev(np.broadcast_to(mask, (batch, n, n)))
    t = dev(np.broadcast_to(tgt, (batch, n, n)))
    plan.forward(m, a, w, threshold=0.225, alpha=50.0, batch=batch)
    pe, _ = plan.pe_gradient(m, t, grad_mask=g, threshold=0.225, alpha=50.0, batch=batch)
    return host(a), host(w), host(g), list(pe)


@pytest.mark.parametrize("n,k", [(65, 8), (100, 8), (462, 24), (1000, 6)])
def test_direct_dft_parity(gpu_ctx, n, k):
    la = litho_amd()
    assert not la.supported_size(n)
    s, tgt, mask = case(n, k)
    print(f"case direct-dft n={n} k={k} window {s.kernels.shape[1]}x{s.kernels.shape[2]}", flush=True)
    plan = la.Plan.from_socs(gpu_ctx, s)
    path = plan.column_path()
    a, w, g, pe = run(plan, n, mask, tgt)
    plan.close()
    cache = im.Cache.from_socs(s)
    ra, rw, fields = im.forward(cache, mask.astype(np.float64), 0.225, 50.0)
    rg = im.pe_gradient(cache, rw, tgt.astype(np.float64), fields, 50.0)
    rpe = im.pe_loss(rw, tgt)
    ea, eg = rel_linf(a[0], ra), rel_l2(g[0], rg)
    print(f"n={n} {path}: aerial rel Linf {ea:.2e}, wafer abs {np.max(np.abs(w[0] - rw)):.2e}, gradient rel L2 "
          f"{eg:.2e}, PE {pe[0]:.6f} vs {rpe:.6f}")
    assert path["transforms"] == "direct_dft" and path["corr_len"] > 0
    assert ea <= 1e-5 and np.max(np.abs(w[0] - rw)) <= 1e-4
    assert eg <= 1e-4
    assert abs(pe[0] - rpe) <= 1e-5 * max(rpe, 1.0)


def test_direct_dft_equals_fft_pipeline(gpu_ctx, monkeypatch):
    la = litho_amd()
    n, k = 640, 6
    s = window_kernels(n, k)
    tgt = targets().manhattan_target(n, 9)
    mask = im.mask_from_theta(oc.initialize_theta(tgt) + 0.3 * np.random.default_rng(1).standard_normal((n, n)))
    mask = mask.astype(np.float32)
    print(f"case direct-dft vs fft n={n} k={k}", flush=True)
    plan = la.Plan.from_socs(gpu_ctx, s)
    a0, _, g0, pe0 = run(plan, n, mask, tgt)
    assert plan.column_path()["transforms"] == "fft"
    plan.close()
    monkeypatch.setenv("LITHO_SOCS_PATH", "generic")
    plan = la.Plan.from_socs(gpu_ctx, s)
    monkeypatch.delenv("LITHO_SOCS_PATH", raising=False)
    assert plan.column_path()["transforms"] == "direct_dft"
    a1, _, g1, pe1 = run(plan, n, mask, tgt)
    plan.close()
    ea, eg = rel_linf(a1[0], a0[0]), rel_l2(g1[0], g0[0])
    print(f"n={n}: direct-dft vs fft aerial {ea:.2e}, gradient {eg:.2e}, PE {pe1[0]:.6f} vs {pe0[0]:.6f}")
    assert ea <= 1e-5 and eg <= 1e-5
    assert abs(pe1[0] - pe0[0]) <= 1e-6 * max(pe0[0], 1.0)


def test_direct_dft_gd_trajectory_65(gpu_ctx):
    la = litho_amd()
    n, k = 65, 8
    s, tgt, _ = case(n, k)
    print(f"case direct-dft GD n={n}", flush=True)
    plan = la.Plan.from_socs(gpu_ctx, s)
    out = empty(n)
    r = plan.ctm_optimize(dev(tgt), out[0], method="gradient_descent", iterations=10, learning_rate=0.9, tol=0.0,
                          threshold=0.225, alpha=50.0)
    plan.close()
    th0 = oc.initialize_theta(tgt).astype(np.float32).astype(np.float64)
    best, hist, _ = oc.gradient_descent(im.Cache.from_socs(s), tgt.astype(np.float64), 0.225, 50.0, 10, 0.9, 0.0,
                                        theta0=th0)
    h = np.asarray(r["history"])
    rel = np.abs(h - np.asarray(hist)) / np.asarray(hist)
    print(f"GD n={n}: PE {h[0]:.4f} -> {h[-1]:.4f}, max rel dev {rel.max():.2e}")
    assert r["evaluations"] == 10
    assert rel.max() <= 1e-4


def test_direct_dft_batch_equals_single(gpu_ctx):
    la = litho_amd()
    n, k = 100, 8
    s, tgt, mask = case(n, k)
    print(f"case direct-dft batch n={n}", flush=True)
    plan = la.Plan.from_socs(gpu_ctx, s, capacity=3)
    a1, _, g1, pe1 = run(plan, n, mask, tgt, batch=1)
    a3, _, g3, pe3 = run(plan, n, mask, tgt, batch=3)
    plan.close()
    for b in range(3):
        assert np.array_equal(a3[b], a1[0]) and np.array_equal(g3[b], g1[0]) and pe3[b] == pe1[0]


@pytest.mark.parametrize("n", [65, 462])
def test_direct_dft_lbfgs(gpu_ctx, n):
    """L-BFGS (m = 10, 10 iterations) at an odd size (its N^2 elements are not a
    multiple of 4: the vector kernels run per float) and at 462: the oracle's
    evaluation count and per-evaluation PE history, as in test_gpu_ctm.py."""
    from test_gpu_ctm import assert_lbfgs_trajectory
    la = litho_amd()
    s, tgt, _ = case(n, 8)
    print(f"case direct-dft lbfgs n={n}", flush=True)
    plan = la.Plan.from_socs(gpu_ctx, s)
    out = empty(n)
    r = plan.ctm_optimize(dev(tgt), out[0], method="lbfgs", iterations=10, history=10, tol=1e-2, threshold=0.225,
                          alpha=50.0)
    plan.close()
    cache = im.Cache.from_socs(s)
    mask, hist, theta, info = oc.lbfgs(cache, tgt.astype(np.float64), 0.225, 50.0, 10, 10, 1e-2)
    assert_lbfgs_trajectory(r, hist, info, f"direct-dft lbfgs n={n}")
