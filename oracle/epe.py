"""Edge-placement-error loss and gradient at given edge points (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates make_bilinear_stencil / bilinear_sample / bilinear_scatter
(epe_utils.cpp:38-84), Loss::evaluate_sparse_epe (loss.cpp:125-163) and
Gradient::epe_gradient (gradient.cpp:198-227) with
LevelSetUtils::_gradient_1d (level_set_utils.cpp:228-255).  Pinned in
tests/test_oracle_epe.py: stencils by hand, _gradient_1d on polynomials it
differentiates exactly, and the EPE gradient by finite differences of the
loss with the gradient magnitude frozen (the reference's own derivation,
gradient.cpp:215-216).
"""
from __future__ import annotations

import math

import numpy as np

from .imaging import Cache, _backpropagate, forward


def bilinear_stencil(h: int, w: int, y: float, x: float):
    """make_bilinear_stencil: [(row, col, weight)], up to 4, duplicates merged."""
    yy = min(max(y, 0.0), float(h - 1))
    xx = min(max(x, 0.0), float(w - 1))
    y0, x0 = int(math.floor(yy)), int(math.floor(xx))
    y1, x1 = min(y0 + 1, h - 1), min(x0 + 1, w - 1)
    wy1 = 0.0 if y1 == y0 else yy - y0
    wx1 = 0.0 if x1 == x0 else xx - x0
    out = []
    for py, wy in ((y0, 1.0 - wy1), (y1, wy1)):
        for px, wx in ((x0, 1.0 - wx1), (x1, wx1)):
            wt = wy * wx
            if wt <= 0.0:
                continue
            for e in out:
                if e[0] == py and e[1] == px:
                    e[2] += wt
                    break
            else:
                out.append([py, px, wt])
    return out


def gradient_1d(m: np.ndarray, d: float, axis: int) -> np.ndarray:
    """LevelSetUtils::_gradient_1d: central differences, one-sided second order at both ends."""
    m = np.asarray(m, dtype=np.float64)
    r = np.empty_like(m)
    if axis == 0:
        r[0] = (-3 * m[0] + 4 * m[1] - m[2]) / (2 * d)
        r[-1] = (3 * m[-1] - 4 * m[-2] + m[-3]) / (2 * d)
        r[1:-1] = (m[2:] - m[:-2]) / (2 * d)
    else:
        r[:, 0] = (-3 * m[:, 0] + 4 * m[:, 1] - m[:, 2]) / (2 * d)
        r[:, -1] = (3 * m[:, -1] - 4 * m[:, -2] + m[:, -3]) / (2 * d)
        r[:, 1:-1] = (m[:, 2:] - m[:, :-2]) / (2 * d)
    return r


def _sample(field, st):
    return sum(wt * field[py, px] for py, px, wt in st)


def sparse_epe(aerial: np.ndarray, points_yx, weights, threshold: float, pixel_nm: float):
    """Loss::evaluate_sparse_epe: (weighted squared loss, abs sum nm, active
    count, per-point dL/dI_ep = 2 w (I - tau) / G^2)."""
    a = np.asarray(aerial, dtype=np.float64)
    gm = np.maximum(np.sqrt(gradient_1d(a, 1.0, 1) ** 2 + gradient_1d(a, 1.0, 0) ** 2), 1e-6)
    h, w = a.shape
    loss = absnm = 0.0
    act = 0
    sg = np.zeros(len(weights))
    for i, (pt, wt) in enumerate(zip(points_yx, weights)):
        if wt <= 0.0:
            continue
        st = bilinear_stencil(h, w, float(pt[0]), float(pt[1]))
        iv = _sample(a, st)
        mag = max(_sample(gm, st), 1e-6)
        s = (iv - threshold) / mag
        loss += wt * s * s
        absnm += abs(s) * pixel_nm
        act += 1
        sg[i] = 2.0 * wt * (iv - threshold) / (mag * mag)
    return loss, absnm, act, sg


def epe_dldi(shape, points_yx, weights, sg) -> np.ndarray:
    """bilinear_scatter of every active point's dL/dI_ep, in point order."""
    h, w = shape
    d = np.zeros(shape)
    for pt, wt, g in zip(points_yx, weights, sg):
        if wt <= 0.0:
            continue
        for py, px, sw in bilinear_stencil(h, w, float(pt[0]), float(pt[1])):
            d[py, px] += sw * g
    return d


def epe_gradient(cache: Cache, mask, points_yx, weights, threshold: float, alpha: float, pixel_nm: float = 4.0):
    """Gradient::epe_gradient and the loss of the same evaluation:
    returns (grad, loss, abs_sum_nm, active, aerial, wafer, fields)."""
    aerial, wafer, fields = forward(cache, mask, threshold, alpha)
    loss, absnm, act, sg = sparse_epe(aerial, points_yx, weights, threshold, pixel_nm)
    grad = _backpropagate(cache, epe_dldi(aerial.shape, points_yx, weights, sg), fields, 2.0)
    return grad, loss, absnm, act, aerial, wafer, fields
