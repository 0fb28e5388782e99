"""Pins of the oracle's level-set numerics (no GPU).

Ports tests/test_level_set_godunov.cpp (plane, extremum, circle symmetry,
size mismatch) and checks the chamfer distance transform against a
brute-force shortest path over the 5x5 chamfer mask.
"""
import heapq

import numpy as np
import pytest

from oracle import levelset as ls


def x_plane(rows, cols, slope):
    c = 0.5 * (cols - 1)
    return np.tile(slope * (np.arange(cols) - c), (rows, 1))


def test_planar_signed_distance():
    rows, cols = 25, 29
    pos = np.full((rows, cols), 2.0)
    inc = ls.evolve_normal_weno_godunov(x_plane(rows, cols, 1.0), pos, 1.0, 1.0)
    dec = ls.evolve_normal_weno_godunov(x_plane(rows, cols, -1.0), pos, 1.0, 1.0)
    rev = ls.evolve_normal_weno_godunov(x_plane(rows, cols, 1.0), -pos, 1.0, 1.0)
    r, c = rows // 2, cols // 2
    assert abs(inc[0][r, c] - 2.0) < 1e-12
    assert abs(dec[0][r, c] - 2.0) < 1e-12
    assert abs(rev[0][r, c] + 2.0) < 1e-12
    assert abs(inc[1][r, c] - 2.0) < 1e-12 and abs(inc[2][r, c] - 2.0) < 1e-12
    assert inc[0].shape == (rows, cols)


def test_godunov_at_extremum():
    rows, cols = 17, 25
    center = cols // 2
    mn = np.tile(np.abs(np.arange(cols) - center).astype(float), (rows, 1))
    mx = -mn
    one = np.ones((rows, cols))
    assert abs(ls.evolve_normal_weno_godunov(mn, one, 1, 1)[0][rows // 2, center]) < 1e-12
    assert abs(ls.evolve_normal_weno_godunov(mn, -one, 1, 1)[0][rows // 2, center] + 1.0) < 1e-6
    assert abs(ls.evolve_normal_weno_godunov(mx, one, 1, 1)[0][rows // 2, center] - 1.0) < 1e-6
    assert abs(ls.evolve_normal_weno_godunov(mx, -one, 1, 1)[0][rows // 2, center]) < 1e-12


def test_size_mismatch_rejected():
    with pytest.raises(ValueError):
        ls.evolve_normal_weno_godunov(np.zeros((5, 7)), np.zeros((5, 6)), 1.0, 1.0)


def test_circle_mirror_symmetry():
    size, radius = 33, 10.0
    c = 0.5 * (size - 1)
    y, x = np.mgrid[0:size, 0:size]
    phi = radius - np.hypot(x - c, y - c)
    delta = ls.evolve_normal_weno_godunov(phi, np.ones((size, size)), 1.0, 1.0)[0]
    m = 4
    band = (np.abs(phi) <= 3.0)
    band[:m, :] = band[-m:, :] = band[:, :m] = band[:, -m:] = False
    assert np.max(np.abs(delta - delta[:, ::-1])[band]) < 1e-12
    assert np.max(np.abs(delta - delta[::-1, :])[band]) < 1e-12


def test_weno5_exact_for_smooth_quadratic():
    # WENO5-JS reproduces the exact derivative of a quadratic where all stencils agree
    x = np.arange(40, dtype=float)
    v = np.tile(0.01 * x * x, (3, 1))
    d = ls.der_weno5(v, 1.0, "minus", 1)
    exact = 0.02 * x
    assert np.max(np.abs(d[:, 6:-6] - exact[6:-6])) < 1e-6


def test_kappa_and_cfl():
    # curvature of a circle phi = R - r is -1/r (b * kappa |grad phi| with |grad| = 1)
    size, r0 = 65, 20.0
    c = 32.0
    y, x = np.mgrid[0:size, 0:size]
    phi = r0 - np.hypot(x - c, y - c)
    k = ls.evolve_kappa(phi, 1.0, 1.0, 1.0)
    assert abs(k[32, 32 + 20] + 1.0 / 20.0) < 2e-3
    dt = ls.get_dt_normal_kappa(0.5, 1.0, 1.0, np.full((3, 3), 2.0), np.full((3, 3), 2.0), 0.01)
    assert abs(dt - 0.5 / (4.0 + 0.04)) < 1e-15
    assert ls.get_dt_normal_kappa(0.5, 1.0, 1.0, np.zeros((2, 2)), np.zeros((2, 2)), 0.0) == 1e-3


def chamfer_dijkstra(nonzero):
    """Brute-force shortest path over the 5x5 chamfer mask (fixed-point weights)."""
    hv, dg, lg, init, dmax = ls.chamfer_weights()
    h, w = nonzero.shape
    dist = np.full((h, w), np.inf)
    pq = []
    for y, x in zip(*np.nonzero(~nonzero)):
        dist[y, x] = 0
        heapq.heappush(pq, (0, y, x))
    moves = [(0, 1, hv), (1, 0, hv), (1, 1, dg), (1, 2, lg), (2, 1, lg)]
    steps = []
    for dy, dx, wgt in moves:
        for sy in (-1, 1):
            for sx in (-1, 1):
                steps.append((sy * dy, sx * dx, wgt))
    while pq:
        d, y, x = heapq.heappop(pq)
        if d > dist[y, x]:
            continue
        for dy, dx, wgt in steps:
            yy, xx = y + dy, x + dx
            if 0 <= yy < h and 0 <= xx < w and d + wgt < dist[yy, xx]:
                dist[yy, xx] = d + wgt
                heapq.heappush(pq, (d + wgt, yy, xx))
    return (np.minimum(dist, dmax) / 65536.0).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_chamfer_equals_shortest_path(seed):
    rng = np.random.default_rng(seed)
    img = np.ones((23, 31), dtype=bool)
    for _ in range(4):
        y, x = rng.integers(0, 23), rng.integers(0, 31)
        img[y:y + rng.integers(1, 5), x:x + rng.integers(1, 5)] = False
    got = ls.distance_transform_5x5(img)
    assert np.array_equal(got, chamfer_dijkstra(img))


def test_chamfer_weights_fixed_point():
    hv, dg, lg, init, dmax = ls.chamfer_weights()
    assert (hv, dg, lg) == (65536, 91750, 143976)


def test_reinit_sdf_signs_and_empty_tile():
    phi = -np.ones((16, 16))
    phi[5:11, 4:12] = 1.0
    sdf = ls.reinit_sdf(phi, 1.0, 1.0)
    assert np.all(sdf[phi > 0] > 0) and np.all(sdf[phi <= 0] < 0)
    assert sdf[7, 7] == pytest.approx(3.0, abs=1e-6)   # 3 pixels from the nearest outside pixel
    empty = ls.reinit_sdf(-np.ones((8, 8)), 1.0, 1.0)
    assert np.all(empty == -8192.0)                     # no inside pixel: OpenCV's DIST_MAX clamp


# ---------------------------------------------------------------- renderer
# render_levelset_mask (levelset_mask_renderer.cpp:7-33).  OpenCV is absent,
# so the Gray restatement of cv::resize INTER_LINEAR is pinned by closed
# forms: scale 1 is the Binary render, a constant field is all in or all
# out, and a field linear in x is interpolated exactly, so each pixel covers
# the fraction of its subpixel centres u = (x s + k + 0.5) / s - 0.5 (clamped
# to the tile) with u >= c for the zero at x = c.

def test_render_gray_scale_one_is_binary():
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((9, 13)).astype(np.float32)
    assert np.array_equal(ls.render_levelset_mask(phi, "Gray", 1), ls.render_binary(phi))


def test_render_gray_constant_fields():
    assert np.all(ls.render_gray(np.full((5, 7), 0.25), 4) == 1.0)
    assert np.all(ls.render_gray(np.full((5, 7), -0.25), 3) == 0.0)
    assert np.all(ls.render_gray(np.zeros((4, 4)), 2) == 1.0)      # phi >= 0 counts


@pytest.mark.parametrize("scale", [2, 3, 4, 5])
@pytest.mark.parametrize("c", [3.3, 7.0, 7.5, 10.125, -2.0, 20.0])
def test_render_gray_linear_field(scale, c):
    n = 16
    phi = np.tile(np.arange(n, dtype=np.float64) - c, (4, 1))
    got = ls.render_gray(phi, scale)
    for x in range(n):
        u = np.clip((x * scale + np.arange(scale) + 0.5) / scale - 0.5, 0, n - 1)
        assert np.array_equal(got[:, x], np.full(4, np.mean(u - c >= 0)))


def test_render_gray_transpose_and_scale_validation():
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((11, 6))
    assert np.array_equal(ls.render_gray(phi, 3).T, ls.render_gray(phi.T, 3))
    with pytest.raises(ValueError):
        ls.render_gray(phi, 0)
