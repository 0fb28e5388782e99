"""Pins of the EPE oracle (no GPU): bilinear stencils by hand, the
one-sided / central difference operator on polynomials it differentiates
exactly, and the EPE gradient against finite differences of the loss with the
gradient magnitude frozen (how gradient.cpp:215-216 derives it)."""
import numpy as np
import pytest

from oracle import epe as oe
from oracle import imaging as im
from oracle import optics as op


def test_bilinear_stencil_by_hand():
    st = oe.bilinear_stencil(10, 10, 2.25, 3.5)
    got = {(y, x): w for y, x, w in st}
    assert got == {(2, 3): 0.75 * 0.5, (2, 4): 0.75 * 0.5, (3, 3): 0.25 * 0.5, (3, 4): 0.25 * 0.5}
    # on the last row / column the duplicated pixels merge; outside points clamp
    assert oe.bilinear_stencil(10, 10, 9.0, 9.0) == [[9, 9, 1.0]]
    assert oe.bilinear_stencil(10, 10, -3.0, 4.0) == [[0, 4, 1.0]]
    assert sum(w for _, _, w in oe.bilinear_stencil(7, 5, 6.7, 1.2)) == pytest.approx(1.0, abs=1e-15)


def test_gradient_1d_exact_on_quadratics():
    y, x = np.mgrid[0:9, 0:11].astype(float)
    f = 3 * x * x - 2 * x * y + 5 * y * y + x - 7
    assert np.allclose(oe.gradient_1d(f, 1.0, 1), 6 * x - 2 * y + 1, atol=1e-10)
    assert np.allclose(oe.gradient_1d(f, 1.0, 0), -2 * x + 10 * y, atol=1e-10)
    assert np.allclose(oe.gradient_1d(f, 0.5, 1), 2 * (6 * x - 2 * y + 1), atol=1e-10)


def test_epe_gradient_finite_differences():
    n = 32
    cache = im.Cache.from_socs(op.socs_kernels(n, 4.0, 193.0, 1.35, 0.6, 0.9, num_kernels=8))
    rng = np.random.default_rng(1)
    mask = np.zeros((n, n))
    mask[10:22, 8:24] = 1.0
    mask = np.clip(mask + 0.1 * rng.standard_normal((n, n)), 0, 1)
    pts = np.array([[10.0, 15.5], [21.3, 12.0], [16.0, 8.2], [15.5, 23.7], [3.0, 3.0]])
    wts = np.array([1.0, 0.5, 2.0, 1.0, 0.0])
    thr = 0.225
    g, loss, absnm, act, aerial, _, _ = oe.epe_gradient(cache, mask, pts, wts, thr, 50.0)
    assert act == 4
    gm = np.maximum(np.sqrt(oe.gradient_1d(aerial, 1.0, 1) ** 2 + oe.gradient_1d(aerial, 1.0, 0) ** 2), 1e-6)

    def frozen_loss(m):
        a, _, _ = im.forward(cache, m, thr, 50.0)
        tot = 0.0
        for p, w in zip(pts, wts):
            if w <= 0:
                continue
            st = oe.bilinear_stencil(n, n, p[0], p[1])
            iv = sum(sw * a[py, px] for py, px, sw in st)
            mag = max(sum(sw * gm[py, px] for py, px, sw in st), 1e-6)
            tot += w * ((iv - thr) / mag) ** 2
        return tot

    assert frozen_loss(mask) == pytest.approx(loss, rel=1e-12)
    eps = 1e-6
    for (y, x) in [(10, 15), (16, 9), (21, 12), (5, 5)]:
        mp, mm = mask.copy(), mask.copy()
        mp[y, x] += eps
        mm[y, x] -= eps
        fd = (frozen_loss(mp) - frozen_loss(mm)) / (2 * eps)
        assert g[y, x] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(g).max())
