"""CTM -> binary hand-off, Imbinarize::binarize_advanced (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates source/utils/imbinarize.cpp:84-190 and its helpers
(_quantile :44-58, _peak_values :63-80).  The reference runs OpenCV
(dilate, connectedComponentsWithStats, getStructuringElement); OpenCV is not
available here, so those operations are restated from their definitions:
* dilation with a structuring element: the maximum over the element's
  footprint centred on the pixel, pixels outside the image ignored (OpenCV's
  default border for morphology);
* 8-connected components and their pixel areas (scipy.ndimage.label with a
  3 x 3 structure);
* the MORPH_ELLIPSE element of OpenCV's getStructuringElement.
Pinned in tests/test_oracle_binarize.py by brute force (flood fill, explicit
neighbourhood maxima, sorted-list quantiles) and the known 11 x 11 ellipse.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import ndimage


def ellipse_element(ksize: int) -> np.ndarray:
    """getStructuringElement(MORPH_ELLIPSE, Size(ksize, ksize)): row i has ones
    on [max(c - dx, 0), min(c + dx + 1, ksize)) with r = c = ksize // 2 and
    dx = round(c * sqrt((r^2 - dy^2) / r^2)), dy = i - r (rounded to nearest,
    ties to even, as cvRound)."""
    if ksize == 1:
        return np.ones((1, 1), np.uint8)
    r = c = ksize // 2
    inv_r2 = 1.0 / (r * r) if r else 0.0
    el = np.zeros((ksize, ksize), np.uint8)
    for i in range(ksize):
        dy = i - r
        if abs(dy) <= r:
            dx = int(round(c * math.sqrt((r * r - dy * dy) * inv_r2)))
            el[i, max(c - dx, 0):min(c + dx + 1, ksize)] = 1
    return el


def dilate(mask: np.ndarray, element: np.ndarray) -> np.ndarray:
    """Binary dilation (uint8 0/1) with a centred element; outside the image is 0."""
    return ndimage.binary_dilation(mask > 0, structure=element > 0, border_value=0).astype(np.uint8)


def peak_values(img: np.ndarray, threshold_abs: float) -> np.ndarray:
    """_peak_values (imbinarize.cpp:63-80): pixels >= their 3 x 3 neighbourhood
    maximum (image border ignored) and > threshold_abs; their values in row
    order."""
    dil = ndimage.maximum_filter(img, size=3, mode="constant", cval=-np.inf)
    is_peak = (img >= dil) & (img > threshold_abs)
    return img[is_peak]


def quantile(vals, q: float) -> float:
    """_quantile (imbinarize.cpp:44-58): linear interpolation between the
    floor and ceil order statistics at q (n - 1)."""
    v = np.sort(np.asarray(vals, dtype=np.float64))
    if v.size == 0:
        return 0.0
    q = min(max(q, 0.0), 1.0)
    pos = q * (v.size - 1)
    lo, hi = int(math.floor(pos)), int(math.ceil(pos))
    if lo == hi:
        return float(v[lo])
    return float(v[lo] + (v[hi] - v[lo]) * (pos - lo))


def remove_small_components(mask: np.ndarray, area_threshold: int) -> np.ndarray:
    """connectedComponentsWithStats(8-connectivity) and zeroing every component
    with area < area_threshold (imbinarize.cpp:125-133, 170-180)."""
    lab, num = ndimage.label(mask > 0, structure=np.ones((3, 3), int))
    area = np.bincount(lab.ravel(), minlength=num + 1)
    keep = area >= area_threshold
    keep[0] = False
    return keep[lab].astype(np.uint8)


def binarize_advanced(gray: np.ndarray, target: np.ndarray, peak_min_intensity: float = 0.1,
                      quantile_val: float = 0.2, area_threshold: int = 50, min_spacing_pixels: int = 15) -> dict:
    """Imbinarize::binarize_advanced (imbinarize.cpp:84-190).  Returns the six
    0/1 masks of BinarizeResult plus the adaptive threshold and peak count."""
    g = np.asarray(gray, dtype=np.float64)
    t = np.asarray(target, dtype=np.float64)
    z = np.zeros(g.shape, np.uint8)
    peaks = peak_values(g, peak_min_intensity)
    if peaks.size == 0:
        return dict(initial=z, filtered=z, sraf_initial=z, sraf_final=z, final=z, keep_out=z, threshold=0.0,
                    num_peaks=0)
    thr = quantile(peaks, quantile_val)
    initial = (g >= thr).astype(np.uint8)
    filtered = remove_small_components(initial, area_threshold)
    tgt = (t > 0.5).astype(np.uint8)
    sraf_initial = filtered * (1 - tgt)
    keep_out = z
    sraf_final = sraf_initial
    if min_spacing_pixels > 0:
        keep_out = dilate(tgt, ellipse_element(2 * min_spacing_pixels + 1))
        sraf_final = sraf_initial * (1 - keep_out)
    final = remove_small_components(tgt | sraf_final, area_threshold)
    return dict(initial=initial, filtered=filtered, sraf_initial=sraf_initial, sraf_final=sraf_final, final=final,
                keep_out=keep_out, threshold=thr, num_peaks=int(peaks.size))
