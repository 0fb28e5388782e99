"""Optical model and one-time SOCS kernel precompute, FP64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, in numpy, the reference files
  source/litho_model/grid.cpp, pupil.cpp, zernike.cpp, source.cpp,
  litho_prepare.cpp.
Every IEEE operation that decides a discrete outcome (the frequency axis, the
source-point coordinates, the floor() in the compact SOCS window) is written in
the same order as the reference so that the window and the point set are
reproduced exactly.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# --------------------------------------------------------------------------- grid
def fft_freq_axis(n: int, pixel_nm: float) -> np.ndarray:
    """Centred physical frequency axis, DC at index n//2 (grid.cpp:38-48)."""
    df = 1.0 / (n * pixel_nm)
    return (np.arange(n, dtype=np.float64) - (n // 2)) * df


def grid_axes(n: int, pixel_nm: float, na: float, wavelength_nm: float):
    """Grid::Grid (grid.cpp:4-36).

    Returns (coord, f_norm): pixel-centre coordinates (i-(n-1)/2)*dx in nm and
    the centred frequency axis divided by NA/lambda.  Rows are y, columns x.
    """
    coord = (np.arange(n, dtype=np.float64) - (n - 1) * 0.5) * pixel_nm
    f_norm = fft_freq_axis(n, pixel_nm) / (na / wavelength_nm)
    return coord, f_norm


# ------------------------------------------------------------------------ zernike
def zernike(index: int, f: np.ndarray, g: np.ndarray) -> np.ndarray:
    """Fringe Zernike Z1..Z11 on normalised pupil coordinates (zernike.cpp:11-43)."""
    r2 = f * f + g * g
    s3, s5, s6, s8 = math.sqrt(3.0), math.sqrt(5.0), math.sqrt(6.0), math.sqrt(8.0)
    table = {
        1: lambda: np.ones_like(f),
        2: lambda: 2.0 * f,
        3: lambda: 2.0 * g,
        4: lambda: s3 * (2.0 * r2 - 1.0),
        5: lambda: s6 * (2.0 * f * g),
        6: lambda: s6 * (f * f - g * g),
        7: lambda: s8 * (3.0 * r2 - 2.0) * g,
        8: lambda: s8 * (3.0 * r2 - 2.0) * f,
        9: lambda: s8 * (3.0 * f * f * g - g ** 3),
        10: lambda: s8 * (f ** 3 - 3.0 * f * g * g),
        11: lambda: s5 * (6.0 * r2 * r2 - 6.0 * r2 + 1.0),
    }
    return table.get(index, lambda: np.zeros_like(f))()


def pupil(f_norm: np.ndarray, aberrations: dict | None = None) -> np.ndarray:
    """Pupil P = [rho^2 <= 1] * exp(i 2 pi W) on the N x N centred grid.

    pupil.cpp:14-23 (amplitude), :41-46 (phase), wavefront from
    zernike.cpp:46-53 (zero coefficients are skipped).
    """
    fx = f_norm[None, :].repeat(f_norm.size, axis=0)
    fy = f_norm[:, None].repeat(f_norm.size, axis=1)
    amp = ((fx * fx + fy * fy) <= 1.0).astype(np.float64)
    w = np.zeros_like(fx)
    for idx, c in (aberrations or {}).items():
        if c == 0.0:
            continue
        w = w + c * zernike(int(idx), fx, fy)
    return amp * np.exp(1j * 2.0 * math.pi * w)


# ------------------------------------------------------------------------- source
def _eigen_linspaced(num: int, low: float, high: float) -> np.ndarray:
    """Eigen::VectorXd::LinSpaced for floating scalars (low + i*step, last = high)."""
    if num == 1:
        return np.array([high], dtype=np.float64)
    step = (high - low) / float(num - 1)
    if abs(high) < abs(low):  # Eigen's flipped evaluation
        i = np.arange(num, dtype=np.float64)
        out = high - (num - 1 - i) * step
        out[0] = low
        return out
    out = low + np.arange(num, dtype=np.float64) * step
    out[-1] = high
    return out


def source_map(freq_coords: np.ndarray, sigma_in: float, sigma_out: float,
               upsample: int = 10, smoothing: float = 0.0):
    """Source::compute_source_map (source.cpp:114-158).

    Keeps the axis points inside [-2, 2], evaluates the annulus on a
    upsample-times linspace grid, average-pools back, no normalisation.
    Returns (coords, weight_map) with weight_map[row=fy, col=fx].
    """
    ff = freq_coords[(freq_coords >= -2.0) & (freq_coords <= 2.0)]
    num = ff.size
    hi = _eigen_linspaced(num * upsample, ff[0], ff[-1])
    ffx = hi[None, :]
    ffy = hi[:, None]
    rho = np.sqrt(ffx * ffx + ffy * ffy)

    def smooth_step(width, x):  # source.cpp:105-111
        if width == 0.0:
            return (x > 0.0).astype(np.float64)
        from scipy.special import erf
        return 0.5 * (1.0 + erf(x / width))

    src_hi = smooth_step(smoothing, sigma_out - rho) * smooth_step(smoothing, rho - sigma_in)
    if upsample > 1:
        m = num
        src = src_hi[: m * upsample, : m * upsample].reshape(m, upsample, m, upsample)
        src = src.sum(axis=(1, 3)) / float(upsample * upsample)
    else:
        src = src_hi
    src = np.maximum(src, 0.0)
    if src.size == 0 or src.max() <= 0.0:
        raise RuntimeError("source map is all zero")
    return ff.copy(), src


def source_points(coords: np.ndarray, weight_map: np.ndarray, min_weight: float = 1e-8):
    """LithoPrepare::_extract_source_points (litho_prepare.cpp:65-102).

    Row-major scan (row = fy, col = fx); returns fs, gs, ws for w > min_weight.
    """
    rows, cols = np.nonzero(weight_map > min_weight)   # C order == row-major scan
    if rows.size == 0:
        raise RuntimeError("source has no points")
    return coords[cols].copy(), coords[rows].copy(), weight_map[rows, cols].copy()


# ------------------------------------------------------------------- SOCS precompute
@dataclass
class SocsKernels:
    """The subset of ImagingCache (litho_prepare.h:18-47) the hot path reads.

    kernels[k] is the W_r x W_c frequency window of SOCS kernel k in the
    *centred* frequency layout: element (r, c) sits at centred index
    (win_y0 + r, win_x0 + c); DC-corner index = (centred - N//2) mod N.
    """
    n: int
    win_y0: int
    win_x0: int
    kernels: np.ndarray            # (K, Wr, Wc) complex128
    weights: np.ndarray            # (K,) lambda_k (SOCS) or w_i (Abbe)
    weight_sum: float              # sum of source weights
    singular_values: np.ndarray = field(default_factory=lambda: np.zeros(0))
    num_source_points: int = 0

    @property
    def k(self) -> int:
        return int(self.kernels.shape[0])

    @property
    def win_rows(self) -> int:
        return int(self.kernels.shape[1])

    @property
    def win_cols(self) -> int:
        return int(self.kernels.shape[2])

    def dense(self, k: int) -> np.ndarray:
        """N x N DC-corner kernel = ifftshift(zero-padded window)
        (litho_prepare.cpp:330,343-345)."""
        full = np.zeros((self.n, self.n), dtype=np.complex128)
        full[self.win_y0:self.win_y0 + self.win_rows,
             self.win_x0:self.win_x0 + self.win_cols] = self.kernels[k]
        return np.fft.ifftshift(full)

    def row_freqs(self) -> np.ndarray:
        """DC-corner frequency index of each window row."""
        return (self.win_y0 + np.arange(self.win_rows) - self.n // 2) % self.n

    def col_freqs(self) -> np.ndarray:
        return (self.win_x0 + np.arange(self.win_cols) - self.n // 2) % self.n


def socs_window(p: np.ndarray, fs, gs, dfx: float, dfy: float, compact: bool = True):
    """Compact support rectangle of all shifted pupils (litho_prepare.cpp:238-256)."""
    n = p.shape[0]
    if not compact:
        return 0, 0, n, n
    ys, xs = np.nonzero(p != 0)
    y0, y1, x0, x1 = int(ys.min()), int(ys.max()), int(xs.min()), int(xs.max())
    origin_y = max(0, y0 - int(math.floor(float(np.max(gs)) / dfy)) - 1)
    origin_x = max(0, x0 - int(math.floor(float(np.max(fs)) / dfx)) - 1)
    end_y = min(n, y1 - int(math.floor(float(np.min(gs)) / dfy)) + 1)
    end_x = min(n, x1 - int(math.floor(float(np.min(fs)) / dfx)) + 1)
    return origin_y, origin_x, end_y, end_x


def shifted_pupils(p: np.ndarray, fs, gs, ws, dfx: float, dfy: float,
                   origin_y: int, origin_x: int, nr: int, nc: int) -> np.ndarray:
    """Columns of A: sqrt(w_i) * bilinear P(f + f_i) on the window
    (litho_prepare.cpp:266-294).  Returns (Ns, nr, nc) complex128."""
    n = p.shape[0]
    ns = len(fs)
    out = np.empty((ns, nr, nc), dtype=np.complex128)
    pad = np.zeros((n + 2, n + 2), dtype=np.complex128)
    for i in range(ns):
        dc = fs[i] / dfx
        dr = gs[i] / dfy
        c0 = int(math.floor(dc))
        r0 = int(math.floor(dr))
        tc = dc - c0
        tr = dr - r0
        ra = origin_y + np.arange(nr) + r0
        ca = origin_x + np.arange(nc) + c0

        def at(rr, cc):
            vr = (rr >= 0) & (rr < n)
            vc = (cc >= 0) & (cc < n)
            blk = p[np.clip(rr, 0, n - 1)][:, np.clip(cc, 0, n - 1)]
            return blk * (vr[:, None] & vc[None, :])

        sp = ((1 - tr) * (1 - tc) * at(ra, ca) + (1 - tr) * tc * at(ra, ca + 1)
              + tr * (1 - tc) * at(ra + 1, ca) + tr * tc * at(ra + 1, ca + 1))
        out[i] = math.sqrt(ws[i]) * sp
    del pad
    return out


def socs_kernels(n: int, pixel_nm: float, wavelength_nm: float, na: float,
                 sigma_in: float, sigma_out: float, num_kernels: int,
                 aberrations: dict | None = None, compact: bool = True) -> SocsKernels:
    """LithoPrepare(use_socs=true) with K as a parameter (litho_prepare.cpp:16-54,225-385).

    Thin SVD of A (numpy LAPACK); lambda_k = sigma_k^2, kernel_k = U[:, k].
    """
    _, f_norm = grid_axes(n, pixel_nm, na, wavelength_nm)
    p = pupil(f_norm, aberrations)
    coords, wmap = source_map(f_norm, sigma_in, sigma_out)
    fs, gs, ws = source_points(coords, wmap)
    dfx = f_norm[1] - f_norm[0]
    dfy = f_norm[1] - f_norm[0]
    oy, ox, ey, ex = socs_window(p, fs, gs, dfx, dfy, compact)
    nr, nc = ey - oy, ex - ox
    cols = shifted_pupils(p, fs, gs, ws, dfx, dfy, oy, ox, nr, nc)
    a = cols.reshape(len(fs), nr * nc).T            # (nr*nc, Ns)
    u, s, _ = np.linalg.svd(a, full_matrices=False)
    k_use = min(num_kernels if num_kernels > 0 else len(fs), s.size)
    kern = u[:, :k_use].T.reshape(k_use, nr, nc).copy()
    return SocsKernels(n=n, win_y0=oy, win_x0=ox, kernels=kern,
                       weights=(s[:k_use] ** 2).copy(), weight_sum=float(ws.sum()),
                       singular_values=s, num_source_points=len(fs))


def abbe_kernels(n: int, pixel_nm: float, wavelength_nm: float, na: float,
                 sigma_in: float, sigma_out: float,
                 aberrations: dict | None = None) -> SocsKernels:
    """Abbe mode: one shifted pupil per source point, weight w_i
    (litho_prepare.cpp:154-223).  Window = full grid."""
    _, f_norm = grid_axes(n, pixel_nm, na, wavelength_nm)
    p = pupil(f_norm, aberrations)
    coords, wmap = source_map(f_norm, sigma_in, sigma_out)
    fs, gs, ws = source_points(coords, wmap)
    dfx = f_norm[1] - f_norm[0]
    cols = shifted_pupils(p, fs, gs, np.ones_like(ws), dfx, dfx, 0, 0, n, n)
    return SocsKernels(n=n, win_y0=0, win_x0=0, kernels=cols, weights=ws.copy(),
                       weight_sum=float(ws.sum()), num_source_points=len(fs))


def source_summary(n: int, pixel_nm: float, wavelength_nm: float, na: float,
                   sigma_in: float, sigma_out: float):
    """(window side rows, cols, Ns, weight_sum) without the SVD."""
    _, f_norm = grid_axes(n, pixel_nm, na, wavelength_nm)
    p = pupil(f_norm)
    coords, wmap = source_map(f_norm, sigma_in, sigma_out)
    fs, gs, ws = source_points(coords, wmap)
    dfx = f_norm[1] - f_norm[0]
    oy, ox, ey, ex = socs_window(p, fs, gs, dfx, dfx, True)
    return ey - oy, ex - ox, len(fs), float(ws.sum()), oy, ox
