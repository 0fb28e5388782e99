"""Pins of the oracle's optics / SOCS precompute (no GPU)."""
import numpy as np
import pytest

from oracle import imaging as im
from oracle import optics as op

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)


@pytest.mark.parametrize("n", [1, 4, 5, 62, 63, 64, 65, 256, 257])
def test_frequency_axis(n):
    # test_ctm_optimizer.cpp:42-59 and test_full_chip_imaging_compare.cpp:46-81
    if n < 2:
        ax = op.fft_freq_axis(n, 4.0)
        assert ax[0] == 0.0
        return
    _, f = op.grid_axes(n, 4.0, 1.35, 193.0)
    step = (1.0 / (n * 4.0)) * 193.0 / 1.35
    assert f[n // 2] == 0.0
    assert abs(f[0] + (n // 2) * step) < 1e-12
    assert abs(f[-1] - (n - 1 - n // 2) * step) < 1e-12
    assert np.all(np.abs(np.diff(f) - step) < 1e-12)
    phys = op.fft_freq_axis(n, 4.0)
    assert np.allclose(phys, f * 1.35 / 193.0, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [8, 9, 64, 65])
def test_shift_inverse(n):
    x = np.random.default_rng(n).standard_normal((n, n))
    assert np.array_equal(np.fft.fftshift(np.fft.ifftshift(x)), x)
    # numpy semantics: ifftshift moves the centred DC (index n//2) to 0
    e = np.zeros((n, n))
    e[n // 2, n // 2] = 1.0
    assert np.fft.ifftshift(e)[0, 0] == 1.0


@pytest.mark.parametrize("n,w,ns,wsum", [(384, 41, 228, 170.36), (512, 55, 380, 299.64), (640, 67, 552, 465.12),
                                         (896, 97, 1044, 904.56), (1152, 123, 1680, 1489.24),
                                         (2048, 219, 5008, 4679.04), (2176, 231, 5628, 5279.12)])
def test_window_and_source_points(n, w, ns, wsum):
    # configs: 55x55 window at N=512 and 219x219 at N=2048 (annular 0.6-0.9, NA 1.35, 193 nm, 4 nm/px)
    rows, cols, npts, s, y0, x0 = op.source_summary(n, **OPT)
    assert (rows, cols, npts) == (w, w, ns)
    assert abs(s - wsum) < 1e-9
    # the window is centred on DC: symmetric around n//2
    assert y0 + rows // 2 == n // 2 and x0 + cols // 2 == n // 2


def test_source_weights_are_pool_fractions():
    _, f = op.grid_axes(128, **{k: OPT[k] for k in ("pixel_nm",)}, na=1.35, wavelength_nm=193.0)
    coords, wmap = op.source_map(f, 0.6, 0.9)
    assert np.all((wmap * 100.0) - np.round(wmap * 100.0) < 1e-12)   # k/100 block means
    assert 0.0 < wmap.max() <= 1.0 and wmap.min() == 0.0


def test_svd_energy_identity():
    sk = op.socs_kernels(64, **OPT, num_kernels=0)
    _, f = op.grid_axes(64, 4.0, 1.35, 193.0)
    p = op.pupil(f)
    coords, wmap = op.source_map(f, 0.6, 0.9)
    fs, gs, ws = op.source_points(coords, wmap)
    d = f[1] - f[0]
    cols = op.shifted_pupils(p, fs, gs, ws, d, d, sk.win_y0, sk.win_x0, sk.win_rows, sk.win_cols)
    fro2 = float(np.sum(np.abs(cols) ** 2))
    assert abs(np.sum(sk.singular_values ** 2) - fro2) < 1e-10 * fro2
    # kernels are orthonormal
    u = sk.kernels.reshape(sk.k, -1)
    assert np.allclose(u.conj() @ u.T, np.eye(sk.k), atol=1e-10)


def test_compact_equals_dense():
    # test_full_chip_imaging_compare.cpp:83-103: eigenvalues 1e-9, AI/WI 1e-10
    n = 48
    a = op.socs_kernels(n, **OPT, num_kernels=12, compact=True)
    b = op.socs_kernels(n, **OPT, num_kernels=12, compact=False)
    assert np.allclose(a.weights, b.weights, rtol=1e-9, atol=0)
    mask = (np.random.default_rng(1).random((n, n)) > 0.6).astype(float)
    ia, wa, _ = im.forward(im.Cache.from_socs(a), mask, 0.25, 50.0)
    ib, wb, _ = im.forward(im.Cache.from_socs(b), mask, 0.25, 50.0)
    assert np.max(np.abs(ia - ib)) < 1e-10 and np.max(np.abs(wa - wb)) < 1e-10


def test_socs_converges_to_abbe():
    # with every singular vector kept, SOCS equals the Abbe per-point sum
    n = 32
    s = op.socs_kernels(n, **OPT, num_kernels=0)
    a = op.abbe_kernels(n, **OPT)
    mask = (np.random.default_rng(2).random((n, n)) > 0.5).astype(float)
    i_s, _, _ = im.forward(im.Cache.from_socs(s), mask, 0.25, 50.0)
    i_a, _, _ = im.forward(im.Cache.from_socs(a), mask, 0.25, 50.0)
    assert np.max(np.abs(i_s - i_a)) < 1e-11 * max(1.0, np.max(np.abs(i_a)))


def test_eigengap_at_k24():
    # K = 24 must not split a degenerate pair: at 512 the 24th and 25th eigenvalues differ by (l24-l25)/l24 = 0.164
    s = op.socs_kernels(512, **OPT, num_kernels=25)
    gap = (s.weights[23] - s.weights[24]) / s.weights[23]
    assert abs(gap - 0.164) < 0.01


def test_fringe_zernike_orthonormal_on_unit_disk():
    """Z1..Z11 of zernike.cpp:11-43 are orthonormal over the unit disk under the
    disk average (a wrong factor or sign in the table breaks this); pinned by a
    midpoint rule on a fine polar grid."""
    nr, nt = 400, 720
    r = (np.arange(nr) + 0.5) / nr
    t = (np.arange(nt) + 0.5) * 2.0 * np.pi / nt
    rr, tt = np.meshgrid(r, t, indexing="ij")
    f, g = rr * np.cos(tt), rr * np.sin(tt)
    w = rr / (nr * nt) * 2.0          # dA / (pi) with dA = r dr dtheta
    z = [op.zernike(j, f, g) for j in range(1, 12)]
    gram = np.array([[np.sum(a * b * w) for b in z] for a in z])
    assert np.allclose(gram, np.eye(11), atol=2e-4)
    assert np.all(op.zernike(12, f, g) == 0.0)


def test_pupil_with_aberrations_is_unit_modulus_inside():
    n = 64
    _, f_norm = op.grid_axes(n, 4.0, 1.35, 193.0)
    p = op.pupil(f_norm, {4: 0.05, 9: -0.02})
    p0 = op.pupil(f_norm)
    inside = p0 != 0
    assert np.allclose(np.abs(p[inside]), 1.0) and np.all(p[~inside] == 0)
    assert np.max(np.abs(np.angle(p[inside]))) > 0.1
