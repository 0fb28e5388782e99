"""Level-set numerics, FP64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates source/utils/level_set_utils.cpp (edge padding, WENO5-JS one-sided
derivatives, Godunov |grad phi|, curvature term, CFL step) and
source/utils/levelset_mask_renderer.cpp.

reinit_sdf follows OpenCV's distanceTransform(DIST_L2, 5) as the reference calls
it (level_set_utils.cpp:290-319): the 5x5 chamfer with weights 1 / 1.4 / 2.1969
evaluated in 16-bit fixed point by a forward and a backward raster scan, then
scaled to float32.  OpenCV is not present here, so that routine is PARITY
UNPINNED; the in-row recurrences are evaluated as exact integer prefix minima,
which equals the sequential scan.
"""
from __future__ import annotations

import numpy as np


def pad_edge(data: np.ndarray, pad: int, axis: int) -> np.ndarray:
    """LevelSetUtils::pad_edge (level_set_utils.cpp:19-43)."""
    if axis == 0:
        return np.pad(data, ((pad, pad), (0, 0)), mode="edge")
    return np.pad(data, ((0, 0), (pad, pad)), mode="edge")


def weno5_rows(v: np.ndarray, n_orig: int, dx: float, direction: str) -> np.ndarray:
    """LevelSetUtils::weno5_1d applied to every row of v (level_set_utils.cpp:45-113).

    v has n_orig + 6 columns (already edge padded by 3)."""
    d1 = (v[:, 1:] - v[:, :-1]) / dx
    i = np.arange(n_orig)
    if direction == "plus":
        v1, v2, v3, v4, v5 = (d1[:, i + 5], d1[:, i + 4], d1[:, i + 3], d1[:, i + 2], d1[:, i + 1])
    else:
        v1, v2, v3, v4, v5 = (d1[:, i], d1[:, i + 1], d1[:, i + 2], d1[:, i + 3], d1[:, i + 4])
    c13, c76, c116 = 1.0 / 3.0, 7.0 / 6.0, 11.0 / 6.0
    c16, c56 = 1.0 / 6.0, 5.0 / 6.0
    cs1, cs2 = 13.0 / 12.0, 0.25
    g1, g2, g3 = 0.1, 0.6, 0.3
    d_1 = c13 * v1 - c76 * v2 + c116 * v3
    d_2 = -c16 * v2 + c56 * v3 + c13 * v4
    d_3 = c13 * v3 + c56 * v4 - c16 * v5
    t1 = v1 - 2 * v2 + v3
    t2 = v2 - 2 * v3 + v4
    t3 = v3 - 2 * v4 + v5
    s1 = cs1 * t1 * t1 + cs2 * (v1 - 4 * v2 + 3 * v3) * (v1 - 4 * v2 + 3 * v3)
    s2 = cs1 * t2 * t2 + cs2 * (v2 - v4) * (v2 - v4)
    s3 = cs1 * t3 * t3 + cs2 * (3 * v3 - 4 * v4 + v5) * (3 * v3 - 4 * v4 + v5)
    vmax2 = np.maximum.reduce([v1 * v1, v2 * v2, v3 * v3, v4 * v4, v5 * v5])
    eps = 1e-6 * vmax2 + 1e-99
    a1 = g1 / ((s1 + eps) * (s1 + eps))
    a2 = g2 / ((s2 + eps) * (s2 + eps))
    a3 = g3 / ((s3 + eps) * (s3 + eps))
    return (a1 * d_1 + a2 * d_2 + a3 * d_3) / (a1 + a2 + a3)


def der_weno5(m: np.ndarray, dx: float, direction: str, axis: int = -1) -> np.ndarray:
    """LevelSetUtils::der_weno5 (level_set_utils.cpp:115-148)."""
    if direction not in ("plus", "minus"):
        raise ValueError("direction must be plus or minus")
    is_2d = m.shape[0] > 1 and m.shape[1] > 1
    ax = (1 if is_2d else 0) if axis == -1 else axis
    padded = pad_edge(m, 3, ax)
    if ax == 0:
        return weno5_rows(padded.T, m.shape[0], dx, direction).T
    return weno5_rows(padded, m.shape[1], dx, direction)


def evolve_normal_weno_godunov(phi: np.ndarray, vn: np.ndarray, dy: float, dx: float):
    """LevelSetUtils::evolve_normal_WENO_godunov (level_set_utils.cpp:150-225).

    Returns (delta, H1_abs, H2_abs)."""
    if phi.shape != vn.shape:
        raise ValueError("evolve_normal_WENO_godunov: phi and Vn sizes must match")
    if phi.size == 0 or dx <= 0.0 or dy <= 0.0:
        raise ValueError("evolve_normal_WENO_godunov: grid must be non-empty and dx/dy must be positive")
    data_ext = pad_edge(pad_edge(phi, 3, 0), 3, 1)
    vn_ext = pad_edge(pad_edge(vn, 3, 0), 3, 1)
    pxm = der_weno5(data_ext, dx, "minus", 1)
    pxp = der_weno5(data_ext, dx, "plus", 1)
    pym = der_weno5(data_ext.T, dy, "minus", 1).T
    pyp = der_weno5(data_ext.T, dy, "plus", 1).T
    pos = vn_ext > 0.0
    neg = vn_ext < 0.0
    gx_pos = np.maximum(np.maximum(pxm, 0.0) ** 2, np.minimum(pxp, 0.0) ** 2)
    gy_pos = np.maximum(np.maximum(pym, 0.0) ** 2, np.minimum(pyp, 0.0) ** 2)
    gx_neg = np.maximum(np.minimum(pxm, 0.0) ** 2, np.maximum(pxp, 0.0) ** 2)
    gy_neg = np.maximum(np.minimum(pym, 0.0) ** 2, np.maximum(pyp, 0.0) ** 2)
    gx = np.where(pos, gx_pos, np.where(neg, gx_neg, 0.0))
    gy = np.where(pos, gy_pos, np.where(neg, gy_neg, 0.0))
    grad = np.sqrt(gx + gy)
    delta_ext = vn_ext * grad
    h_ext = np.abs(vn_ext)
    r, c = phi.shape
    return delta_ext[3:3 + r, 3:3 + c].copy(), h_ext[3:3 + r, 3:3 + c].copy(), h_ext[3:3 + r, 3:3 + c].copy()


def gradient_1d(m: np.ndarray, d: float, axis: int) -> np.ndarray:
    """LevelSetUtils::_gradient_1d (level_set_utils.cpp:228-255): central
    differences, second-order one-sided at the two ends."""
    a = m if axis == 1 else m.T
    out = np.empty_like(a)
    out[:, 0] = (-3 * a[:, 0] + 4 * a[:, 1] - a[:, 2]) / (2 * d)
    out[:, -1] = (3 * a[:, -1] - 4 * a[:, -2] + a[:, -3]) / (2 * d)
    out[:, 1:-1] = (a[:, 2:] - a[:, :-2]) / (2 * d)
    return out if axis == 1 else out.T


def evolve_kappa(phi: np.ndarray, dx: float, dy: float, b: float) -> np.ndarray:
    """LevelSetUtils::evolve_kappa (level_set_utils.cpp:257-278)."""
    px = gradient_1d(phi, dx, 1)
    py = gradient_1d(phi, dy, 0)
    pxy = gradient_1d(px, dy, 0)
    pxx = gradient_1d(px, dx, 1)
    pyy = gradient_1d(py, dy, 0)
    g2 = px * px + py * py
    num = pxx * py * py - 2.0 * px * py * pxy + pyy * px * px
    eps = np.finfo(np.float64).eps
    out = np.zeros_like(phi)
    ok = g2 > eps
    out[ok] = num[ok] / g2[ok]
    return b * out


def get_dt_normal_kappa(cfl: float, dx: float, dy: float, h1, h2, b: float) -> float:
    """LevelSetUtils::get_dt_normal_kappa (level_set_utils.cpp:280-288)."""
    normal = float(np.max(h1)) / dx + float(np.max(h2)) / dy
    kappa = b * (1.0 / (dx * dx) + 1.0 / (dy * dy)) * 2
    den = normal + kappa
    if den < 1e-9:
        return 1e-3
    return cfl / den


# ------------------------------------------------------------------ chamfer SDF
_SHIFT = 16
_HV = int(round(1.0 * (1 << _SHIFT)))
_DIAG = int(np.round(np.float32(1.4) * np.float32(1 << _SHIFT)))
_LONG = int(np.round(np.float32(2.1969) * np.float32(1 << _SHIFT)))
_INIT = 2 ** 31 - 1
_DIST_MAX = _INIT >> 2


def chamfer_weights():
    return _HV, _DIAG, _LONG, _INIT, _DIST_MAX


def distance_transform_5x5(src_nonzero: np.ndarray) -> np.ndarray:
    """Distance of every nonzero pixel to the nearest zero pixel, 5x5 chamfer
    (OpenCV DIST_L2 mask 5 semantics).  Returns float32."""
    h, w = src_nonzero.shape
    big = np.int64(_INIT)
    tmp = np.full((h + 4, w + 4), big, dtype=np.int64)   # 2-pixel border = INIT
    cols = np.arange(w)
    for i in range(h):
        r = i + 2
        up1 = tmp[r - 1]
        up2 = tmp[r - 2]
        t0 = np.minimum.reduce([up2[cols + 1] + _LONG, up2[cols + 3] + _LONG,
                                up1[cols + 0] + _LONG, up1[cols + 1] + _DIAG,
                                up1[cols + 2] + _HV, up1[cols + 3] + _DIAG,
                                up1[cols + 4] + _LONG])
        u = np.where(src_nonzero[i], t0, 0)
        # tmp[j] = min(u[j], tmp[j-1] + HV), left border = INIT
        left = big + _HV * (cols + 1)
        run = np.minimum.accumulate(u - _HV * cols) + _HV * cols
        tmp[r, 2:2 + w] = np.minimum(run, left)
    out = np.empty((h, w), dtype=np.float32)
    scale = np.float32(1.0 / (1 << _SHIFT))
    for i in range(h - 1, -1, -1):
        r = i + 2
        dn1 = tmp[r + 1]
        dn2 = tmp[r + 2]
        cur = tmp[r, 2:2 + w]
        t0 = np.minimum.reduce([cur, dn2[cols + 3] + _LONG, dn2[cols + 1] + _LONG,
                                dn1[cols + 4] + _LONG, dn1[cols + 3] + _DIAG,
                                dn1[cols + 2] + _HV, dn1[cols + 1] + _DIAG,
                                dn1[cols + 0] + _LONG])
        rev = t0[::-1]
        rc = np.arange(w)
        run = (np.minimum.accumulate(rev - _HV * rc) + _HV * rc)[::-1]
        right = big + _HV * (w - cols)
        val = np.minimum(run, right)
        tmp[r, 2:2 + w] = val
        out[i] = (np.minimum(val, _DIST_MAX).astype(np.float32) * scale)
    return out


def reinit_sdf(phi: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """LevelSetUtils::reinit_sdf (level_set_utils.cpp:290-319): inside (phi > 0)
    positive, sdf = dist_in*dx - dist_out*dx."""
    inside = phi > 0
    dist_out = distance_transform_5x5(~inside)   # nonzero outside
    dist_in = distance_transform_5x5(inside)     # nonzero inside
    return dist_in.astype(np.float64) * dx - dist_out.astype(np.float64) * dx


def render_binary(phi: np.ndarray) -> np.ndarray:
    """render_levelset_mask, Binary mode (levelset_mask_renderer.cpp:7-8)."""
    return (phi >= 0.0).astype(np.float64)


def _linear_taps(n_src: int, scale: int):
    """Source indices and float32 weights of a bilinear upscale by `scale`
    with pixel-centre alignment and replicated borders, as cv::resize with
    INTER_LINEAR defines them: f = (d + 0.5) / scale - 0.5 in FP64, rounded to
    float32; s = floor(f); f -= s; s < 0 -> (0, f = 0); s >= n - 1 -> (n - 1,
    f = 0)."""
    d = np.arange(n_src * scale, dtype=np.float64)
    f = ((d + 0.5) * (1.0 / scale) - 0.5).astype(np.float32)
    s = np.floor(f).astype(np.int64)
    f = (f - s.astype(np.float32)).astype(np.float32)
    low = s < 0
    f[low] = 0.0
    s[low] = 0
    high = s >= n_src - 1
    f[high] = 0.0
    s[high] = n_src - 1
    return s, np.minimum(s + 1, n_src - 1), f


def render_gray(phi: np.ndarray, scale: int) -> np.ndarray:
    """render_levelset_mask, Gray mode (levelset_mask_renderer.cpp:10-33):
    float32 phi upscaled by `scale` (bilinear, each output row interpolated
    along x first, then the two rows along y, float32 products and sums),
    subpixels >= 0 counted over each pixel's scale x scale block, divided by
    scale^2.  OpenCV is not present here: parity with cv::resize is unpinned
    (its float rounding depends on the build's vector code); this restatement
    is pinned by closed forms (constant and linear fields, scale 1)."""
    h, w = np.shape(phi)
    v = upscale_linear(phi, scale)
    inside = (v >= np.float32(0.0)).reshape(h, scale, w, scale).sum(axis=(1, 3))
    return inside.astype(np.float64) / float(scale * scale)


def upscale_linear(phi: np.ndarray, scale: int) -> np.ndarray:
    """The float32 bilinear upscale of render_gray, (rows * scale, cols * scale)."""
    if scale <= 0:
        raise ValueError("render_levelset_mask: scale must be positive")
    p = np.asarray(phi, dtype=np.float32)
    h, w = p.shape
    ry0, ry1, fy = _linear_taps(h, scale)
    cx0, cx1, fx = _linear_taps(w, scale)
    one = np.float32(1.0)
    a0, a1 = (one - fx).astype(np.float32), fx
    b0, b1 = (one - fy).astype(np.float32)[:, None], fy[:, None]
    r0, r1 = p[ry0], p[ry1]
    h0 = r0[:, cx0] * a0 + r0[:, cx1] * a1
    h1 = r1[:, cx0] * a0 + r1[:, cx1] * a1
    return h0 * b0 + h1 * b1


def render_levelset_mask(phi: np.ndarray, mask_type: str = "Binary", scale: int = 4) -> np.ndarray:
    """render_levelset_mask(phi, type, scale) (levelset_mask_renderer.cpp:7-33)."""
    if mask_type == "Binary":
        return render_binary(phi)
    return render_gray(phi, scale)
