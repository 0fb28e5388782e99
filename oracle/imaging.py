"""SOCS forward imaging, PE losses and adjoint gradients, FP64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Literal restatement of source/litho_model/imaging.cpp, source/utils/gradient.cpp
and source/utils/loss.cpp: full N x N complex FFTs (1 + K forward, 2K adjoint),
the ifftshift / fftshift pairs kept exactly as the reference writes them, and
the fixed k = 0..K-1 summation order.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft as sfft

from .optics import SocsKernels

_WORKERS = int(os.environ.get("LITHO_ORACLE_THREADS", "1"))


def set_threads(n: int) -> None:
    """Number of FFT worker threads (the reference's OpenMP over k)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def _fft2(x):
    return sfft.fft2(x, workers=_WORKERS)


def _ifft2_unnorm(x):
    # FFTW_BACKWARD is unnormalised; the reference multiplies by 1/N^2 itself.
    return sfft.ifft2(x, norm="forward", workers=_WORKERS)


class Cache:
    """Dense view of the imaging cache: N, kernels (DC corner), weights, sum(w)."""

    def __init__(self, n: int, kernels_dc: list, weights, weight_sum: float):
        self.n = int(n)
        self.kernels = [np.asarray(k, dtype=np.complex128) for k in kernels_dc]
        self.weights = np.asarray(weights, dtype=np.float64)
        self.weight_sum = float(weight_sum)

    @classmethod
    def from_socs(cls, sk: SocsKernels) -> "Cache":
        return cls(sk.n, [sk.dense(k) for k in range(sk.k)], sk.weights, sk.weight_sum)

    @classmethod
    def synthetic(cls, n: int, socs: bool = True, source_ws=(0.5, 0.5)) -> "Cache":
        """Closed-form 2-kernel cache of the reference tests
        (test_micro_gradient_exchange.cpp:28-46): H_k(f) = exp(-(12+8k)|f|^2)
        * exp(i k (0.35 fx - 0.2 fy)), DC at the corner, lambda = {0.7, 0.3}
        in SOCS mode, else the source weights."""
        idx = np.arange(n)
        f = np.where(idx <= n // 2, idx, idx - n).astype(np.float64) / n
        fy, fx = f[:, None], f[None, :]
        kernels = []
        for k in range(2):
            amp = np.exp(-(12.0 + 8.0 * k) * (fx * fx + fy * fy))
            kernels.append(amp * np.exp(1j * (k * (0.35 * fx - 0.2 * fy))))
        ws = np.asarray(source_ws, dtype=np.float64)
        weights = np.array([0.7, 0.3]) if socs else ws
        return cls(n, kernels, weights, float(ws.sum()))


def check_inputs(cache: Cache, mask: np.ndarray, threshold: float, alpha: float) -> None:
    """Validation of Imaging::compute (imaging.cpp:68-107)."""
    n = cache.n
    if mask.shape != (n, n):
        raise ValueError("Imaging: mask dimensions must match cache.N")
    if not np.all(np.isfinite(mask)) or not np.isfinite(threshold) or not np.isfinite(alpha):
        raise ValueError("Imaging: mask, threshold, and alpha must be finite")
    if len(cache.kernels) == 0:
        raise ValueError("Imaging: no optical kernels available")
    if len(cache.weights) != len(cache.kernels):
        raise ValueError("Imaging: SOCS kernel/value counts do not match")
    if not np.isfinite(cache.weight_sum) or cache.weight_sum <= 0.0:
        raise ValueError("Imaging: source weights must have a positive sum")
    if np.any(~np.isfinite(cache.weights)) or np.any(cache.weights < 0.0):
        raise ValueError("Imaging: SOCS values must be finite and nonnegative")


def forward(cache: Cache, mask: np.ndarray, threshold: float = 0.25, alpha: float = 50.0):
    """Imaging::compute (imaging.cpp:60-171).

    Returns (aerial, wafer, fields) where fields[k] = E_k (N x N complex128).
    """
    mask = np.asarray(mask, dtype=np.float64)
    check_inputs(cache, mask, threshold, alpha)
    n = cache.n
    inv = 1.0 / float(n * n)
    mf = _fft2(np.fft.ifftshift(mask.astype(np.complex128)))          # :114-118
    fields = []
    for k in range(len(cache.kernels)):                                # :137-149
        e = _ifft2_unnorm(mf * cache.kernels[k]) * inv
        fields.append(np.fft.fftshift(e))
    aerial = np.zeros((n, n))
    for k in range(len(cache.kernels)):                                # :156-162
        e = fields[k]
        aerial += cache.weights[k] * (e.real * e.real + e.imag * e.imag)
    aerial /= cache.weight_sum                                         # :164
    wafer = 1.0 / (1.0 + np.exp(-alpha * (aerial - threshold)))        # :165-168
    return aerial, wafer, fields


def aerial_directional_derivative(cache: Cache, direction: np.ndarray, fields) -> np.ndarray:
    """Imaging::compute_aerial_directional_derivative (imaging.cpp:173-241):
    dI = 2 sum_k w_k Re(conj(E_k) dE_k) / sum(w)."""
    n = cache.n
    inv = 1.0 / float(n * n)
    df = _fft2(np.fft.ifftshift(np.asarray(direction, dtype=np.complex128)))
    out = np.zeros((n, n))
    for k in range(len(cache.kernels)):
        de = np.fft.fftshift(_ifft2_unnorm(df * cache.kernels[k]) * inv)
        out += 2.0 * cache.weights[k] * (np.conj(fields[k]) * de).real
    return out / cache.weight_sum


def _backpropagate(cache: Cache, dldi: np.ndarray, fields, output_scale: float) -> np.ndarray:
    """Gradient::_backpropagate (gradient.cpp:34-109)."""
    n = cache.n
    inv = 1.0 / float(n * n)
    if dldi.shape != (n, n):
        raise ValueError("Gradient: dLdI dimensions do not match cache")
    if len(fields) != len(cache.kernels):
        raise ValueError("Gradient: optical kernel weights are incomplete")
    grad = np.zeros((n, n))
    for k in range(len(cache.kernels)):
        buf = np.fft.ifftshift(fields[k] * dldi)                        # :76-78
        buf = _fft2(buf) * np.conj(cache.kernels[k])                   # :80-83
        buf = np.fft.fftshift(_ifft2_unnorm(buf))                       # :85-88
        grad += output_scale * cache.weights[k] * inv * buf.real       # :90-91, 98-101
    if cache.weight_sum == 0.0:
        raise RuntimeError("Gradient: source weight sum is zero")
    return grad / cache.weight_sum                                      # :107


def pe_gradient(cache: Cache, wafer, target, fields, alpha: float, window=None) -> np.ndarray:
    """Gradient::pe_gradient with optional loss window (gradient.cpp:121-141)."""
    dldi = (wafer - target) * wafer * (1.0 - wafer)
    if window is not None:
        dldi = dldi * window
    return _backpropagate(cache, dldi, fields, 4.0 * alpha)


def pe_gradient_with_penalty(cache: Cache, wafer, target, fields, alpha: float, penalty: float,
                             penalty_threshold: float = 0.01, pe_window=None, pe_weight: float = 1.0,
                             penalty_window=None) -> np.ndarray:
    """Gradient::pe_gradient_with_penalty, all overloads (gradient.cpp:143-196)."""
    residual = wafer - target
    background = (target < 0.5).astype(np.float64)
    excess = np.maximum(wafer - penalty_threshold, 0.0)
    pw = 1.0 if pe_window is None else pe_window
    bw = pw if penalty_window is None else penalty_window
    dldi = (pe_weight * residual * pw + penalty * background * excess * bw) * wafer * (1.0 - wafer)
    return _backpropagate(cache, dldi, fields, 4.0 * alpha)


def pe_loss(wafer, target, window=None) -> float:
    """Loss::pe_loss (loss.cpp:16-23)."""
    d = (wafer - target) ** 2
    if window is not None:
        d = d * window
    return float(d.sum())


def background_penalty_loss(wafer, target, penalty_threshold: float, window=None) -> float:
    """Loss::background_penalty_loss (loss.cpp:53-64)."""
    excess = np.maximum(wafer - penalty_threshold, 0.0)
    w = 1.0 if window is None else window
    return float((w * (target < 0.5) * excess * excess).sum())


def pe_loss_with_penalty(wafer, target, penalty: float, penalty_threshold: float, window=None) -> float:
    """Loss::pe_loss_with_penalty (loss.cpp:25-51)."""
    return pe_loss(wafer, target, window) + penalty * background_penalty_loss(
        wafer, target, penalty_threshold, window)


def mask_from_theta(theta):
    """M = (1 + cos theta) / 2 (CTM_Optimizer.cpp:82)."""
    return 0.5 * (1.0 + np.cos(theta))


def ctm_evaluate(cache: Cache, theta, target, threshold: float, alpha: float):
    """One grad-eval of CTM_Optimizer::_evaluate_theta_objective
    (CTM_Optimizer.cpp:175-187): returns (pe, grad_theta, mask, aerial, wafer, grad_mask)."""
    mask = mask_from_theta(theta)
    aerial, wafer, fields = forward(cache, mask, threshold, alpha)
    gm = pe_gradient(cache, wafer, target, fields, alpha)
    gt = -0.5 * gm * np.sin(theta)
    return pe_loss(wafer, target), gt, mask, aerial, wafer, gm
