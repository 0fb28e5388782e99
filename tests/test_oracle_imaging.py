"""Pins of the oracle's forward imaging, losses and adjoint (no GPU)."""
import numpy as np
import pytest

from oracle import imaging as im
from oracle import optics as op

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)


def brute_forward(cache, mask):
    """Direct O(N^4) DFT evaluation of I = sum_k w_k |E_k|^2 / sum w with the
    reference's shifts written out literally (tiny N only)."""
    n = cache.n
    idx = np.arange(n)
    f = np.exp(-2j * np.pi * np.outer(idx, idx) / n)          # forward DFT matrix
    finv = np.conj(f) / n                                      # normalised inverse
    m = np.fft.ifftshift(mask)
    mf = f @ m @ f.T
    out = np.zeros((n, n))
    for k, h in enumerate(cache.kernels):
        e = np.fft.fftshift(finv @ (mf * h) @ finv.T)
        out += cache.weights[k] * np.abs(e) ** 2
    return out / cache.weight_sum


@pytest.mark.parametrize("n", [8, 12, 16])
def test_forward_matches_brute_force(n):
    cache = im.Cache.synthetic(n)
    mask = np.random.default_rng(n).random((n, n))
    aerial, wafer, _ = im.forward(cache, mask, 0.25, 50.0)
    assert np.max(np.abs(aerial - brute_forward(cache, mask))) < 1e-14
    assert np.allclose(wafer, 1 / (1 + np.exp(-50.0 * (aerial - 0.25))), rtol=0, atol=1e-15)


def test_known_answers_real_optics():
    sk = op.socs_kernels(64, **OPT, num_kernels=0)
    cache = im.Cache.from_socs(sk)
    a0, w0, _ = im.forward(cache, np.zeros((64, 64)), 0.225, 50.0)
    assert np.all(a0 == 0.0)
    assert np.allclose(w0, 1.0 / (1.0 + np.exp(50.0 * 0.225)), rtol=0, atol=1e-16)
    # M == 1 with every kernel kept: I == sum_i w_i |P(f_i)|^2 / sum w, P bilinear at the DC bin
    a1, _, _ = im.forward(cache, np.ones((64, 64)), 0.225, 50.0)
    _, f = op.grid_axes(64, 4.0, 1.35, 193.0)
    coords, wmap = op.source_map(f, 0.6, 0.9)
    fs, gs, ws = op.source_points(coords, wmap)
    d = f[1] - f[0]
    pdc = []
    for fx, fy in zip(fs, gs):   # bilinear P((0,0) + f_i) of the 0/1 pupil on the centred grid
        c0, r0 = int(np.floor(fx / d)), int(np.floor(fy / d))
        tc, tr = fx / d - c0, fy / d - r0
        val = 0.0
        for dr, wr_ in ((0, 1 - tr), (1, tr)):
            for dc, wc_ in ((0, 1 - tc), (1, tc)):
                yy, xx = 32 + r0 + dr, 32 + c0 + dc
                val += wr_ * wc_ * float(f[xx] ** 2 + f[yy] ** 2 <= 1.0)
        pdc.append(val)
    expect1 = float(np.sum(ws * np.square(pdc)) / ws.sum())
    assert np.max(np.abs(a1 - expect1)) < 1e-12
    # fine sampling (N = 512): the annulus lies well inside the pupil, I(M = 1) -> 1
    sk512 = op.socs_kernels(512, **OPT, num_kernels=0)
    a512, _, _ = im.forward(im.Cache.from_socs(sk512), np.ones((512, 512)), 0.225, 50.0)
    assert np.max(np.abs(a512 - 1.0)) < 1e-12
    # M == 1 with K kernels: I == sum_k lambda_k |H_k[0,0]|^2 / sum w
    sk24 = op.socs_kernels(64, **OPT, num_kernels=8)
    c24 = im.Cache.from_socs(sk24)
    a24, _, _ = im.forward(c24, np.ones((64, 64)), 0.225, 50.0)
    expect = sum(c24.weights[k] * abs(c24.kernels[k][0, 0]) ** 2 for k in range(8)) / c24.weight_sum
    assert np.max(np.abs(a24 - expect)) < 1e-13


def test_single_pixel_psf_and_shift_equivariance():
    sk = op.socs_kernels(64, **OPT, num_kernels=8)
    cache = im.Cache.from_socs(sk)
    mask = np.zeros((64, 64))
    mask[20, 33] = 1.0
    a, _, _ = im.forward(cache, mask, 0.25, 50.0)
    # single pixel: I(x) = sum_k w_k |h_k(x - x0)|^2 / sum w, total energy = sum_k w_k ||H_k||^2 / (N^2 sum w)
    energy = sum(cache.weights[k] * np.sum(np.abs(cache.kernels[k]) ** 2) for k in range(8)) / (64 * 64)
    assert abs(a.sum() - energy / cache.weight_sum) < 1e-13
    rolled, _, _ = im.forward(cache, np.roll(mask, (5, -7), axis=(0, 1)), 0.25, 50.0)
    assert np.max(np.abs(rolled - np.roll(a, (5, -7), axis=(0, 1)))) < 1e-15


def test_adjoint_dot_product():
    # <J dM, g> = <dM, J^T g> with J = dI/dM (JVP of imaging.cpp:173-241)
    n = 32
    cache = im.Cache.from_socs(op.socs_kernels(n, **OPT, num_kernels=6))
    rng = np.random.default_rng(3)
    mask = rng.random((n, n))
    _, _, fields = im.forward(cache, mask, 0.25, 50.0)
    dm = rng.standard_normal((n, n))
    g = rng.standard_normal((n, n))
    jv = im.aerial_directional_derivative(cache, dm, fields)
    # _backpropagate(g, scale 2) is J^T g (gradient.cpp:34-109 with output_scale 2)
    jt = im._backpropagate(cache, g, fields, 2.0)
    lhs, rhs = float(np.sum(jv * g)), float(np.sum(dm * jt))
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)


def test_jvp_matches_finite_difference():
    n = 24
    cache = im.Cache.synthetic(n)
    rng = np.random.default_rng(4)
    mask = 0.4 + 0.1 * rng.random((n, n))
    d = rng.standard_normal((n, n))
    _, _, fields = im.forward(cache, mask, 0.25, 8.0)
    jv = im.aerial_directional_derivative(cache, d, fields)
    h = 1e-6
    ap, _, _ = im.forward(cache, mask + h * d, 0.25, 8.0)
    am, _, _ = im.forward(cache, mask - h * d, 0.25, 8.0)
    assert np.max(np.abs((ap - am) / (2 * h) - jv)) < 1e-8


@pytest.mark.parametrize("n", [64, 65])
@pytest.mark.parametrize("penalty", [False, True])
def test_pe_gradient_finite_difference(n, penalty):
    # test_ctm_optimizer.cpp:61-96: directional FD of PE (and penalty PE) w.r.t. theta, step 1e-3, rel 1e-3
    target = np.zeros((n, n))
    o = (n - 64) // 2
    target[o + 23:o + 42, o + 23:o + 42] = 1.0
    cache = im.Cache.from_socs(op.socs_kernels(n, **OPT, num_kernels=36))
    thr, alpha = 0.25, (10.0 if penalty else 85.0)
    theta = np.where(target > 0, 0.3 * np.pi, 0.7 * np.pi)
    mask = im.mask_from_theta(theta)
    _, wafer, fields = im.forward(cache, mask, thr, alpha)
    if penalty:
        assert im.pe_loss_with_penalty(wafer, target, 2.0, 0.01) > im.pe_loss(wafer, target)
        gm = im.pe_gradient_with_penalty(cache, wafer, target, fields, alpha, 2.0, 0.01)
    else:
        gm = im.pe_gradient(cache, wafer, target, fields, alpha)
    gt = -0.5 * gm * np.sin(theta)
    assert np.linalg.norm(gt) > 1e-8
    d = -gt / np.linalg.norm(gt)

    def f(th):
        _, w, _ = im.forward(cache, im.mask_from_theta(th), thr, alpha)
        return im.pe_loss_with_penalty(w, target, 2.0, 0.01) if penalty else im.pe_loss(w, target)

    step = 1e-3
    num = (f(theta + step * d) - f(theta - step * d)) / (2 * step)
    ana = float(np.sum(gt * d))
    assert abs(num - ana) <= 1e-3 * max(1.0, abs(num))


def test_windowed_gradient_finite_difference():
    # test_micro_gradient_exchange.cpp:200-227: windowed PE FD incl. a pixel outside the window
    n = 20
    cache = im.Cache.synthetic(n)
    rng = np.random.default_rng(5)
    mask = 0.45 + 0.15 * np.sin(0.31 * np.arange(n)[None, :] + 0.17 * np.arange(n)[:, None])
    target = (rng.random((n, n)) > 0.6).astype(float)
    win = np.zeros((n, n))
    win[4:16, 4:16] = 1.0
    _, wafer, fields = im.forward(cache, mask, 0.25, 8.0)
    g = im.pe_gradient(cache, wafer, target, fields, 8.0, win)

    def loss(m):
        _, w, _ = im.forward(cache, m, 0.25, 8.0)
        return im.pe_loss(w, target, win)

    d = g / np.linalg.norm(g)
    h = 1e-5
    num = (loss(mask + h * d) - loss(mask - h * d)) / (2 * h)
    assert abs(num - np.linalg.norm(g)) < 1e-6 * max(1.0, np.linalg.norm(g))
    p = np.zeros((n, n))
    p[3, 5] = 1.0
    num_p = (loss(mask + h * p) - loss(mask - h * p)) / (2 * h)
    assert abs(g[3, 5]) > 1e-8 and abs(num_p - g[3, 5]) < 1e-6 * max(1.0, abs(g[3, 5]))
    # all-ones window equals the unwindowed gradient exactly
    assert np.array_equal(im.pe_gradient(cache, wafer, target, fields, 8.0, np.ones((n, n))),
                          im.pe_gradient(cache, wafer, target, fields, 8.0))


def test_validation_errors():
    cache = im.Cache.synthetic(8)
    with pytest.raises(ValueError):
        im.forward(cache, np.zeros((8, 9)))
    with pytest.raises(ValueError):
        im.forward(cache, np.full((8, 8), np.nan))
    bad = im.Cache(8, cache.kernels, [0.5, -0.1], 1.0)
    with pytest.raises(ValueError):
        im.forward(bad, np.zeros((8, 8)))
    zero = im.Cache(8, cache.kernels, [0.5, 0.5], 0.0)
    with pytest.raises(ValueError):
        im.forward(zero, np.zeros((8, 8)))
