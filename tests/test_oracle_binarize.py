"""Pins of the binarisation oracle (no GPU): the OpenCV operations that
Imbinarize::binarize_advanced uses, checked by brute force and known
answers; the pipeline's stage relations on random fields."""
from collections import deque

import numpy as np
import pytest

from oracle import binarize as ob


def test_ellipse_11_known_pattern():
    # OpenCV's 11 x 11 MORPH_ELLIPSE: one-pixel tips, rows of 7 and 9, five full rows
    want_widths = [1, 7, 9, 11, 11, 11, 11, 11, 9, 7, 1]
    el = ob.ellipse_element(11)
    assert [int(r.sum()) for r in el] == want_widths
    for r in el:                                   # each row is one centred run
        idx = np.flatnonzero(r)
        assert idx[0] + idx[-1] == 10 and np.all(np.diff(idx) == 1)
    assert ob.ellipse_element(3).tolist() == [[0, 1, 0], [1, 1, 1], [0, 1, 0]]
    assert ob.ellipse_element(1).tolist() == [[1]]


def _flood_components(mask):
    """8-connected components by breadth-first search."""
    h, w = mask.shape
    lab = np.zeros((h, w), int)
    areas = [0]
    for y in range(h):
        for x in range(w):
            if mask[y, x] and lab[y, x] == 0:
                areas.append(0)
                k = len(areas) - 1
                q = deque([(y, x)])
                lab[y, x] = k
                while q:
                    cy, cx = q.popleft()
                    areas[k] += 1
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            ny, nx = cy + dy, cx + dx
                            if 0 <= ny < h and 0 <= nx < w and mask[ny, nx] and lab[ny, nx] == 0:
                                lab[ny, nx] = k
                                q.append((ny, nx))
    return lab, areas


@pytest.mark.parametrize("seed,p,thr", [(0, 0.45, 4), (1, 0.3, 2), (2, 0.55, 9)])
def test_remove_small_components_brute_force(seed, p, thr):
    m = (np.random.default_rng(seed).random((37, 41)) < p).astype(np.uint8)
    lab, areas = _flood_components(m)
    want = np.array([a >= thr for a in areas])[lab] & (m > 0)
    assert np.array_equal(ob.remove_small_components(m, thr), want.astype(np.uint8))


def test_diagonal_pixels_are_connected():
    m = np.zeros((6, 6), np.uint8)
    m[1, 1] = m[2, 2] = m[3, 3] = 1                # one 8-connected component of area 3
    assert ob.remove_small_components(m, 3).sum() == 3
    assert ob.remove_small_components(m, 4).sum() == 0


def test_peaks_brute_force_with_plateaus_and_border():
    rng = np.random.default_rng(5)
    img = np.round(rng.random((23, 19)) * 8) / 8     # many ties: plateau pixels count as peaks
    got = ob.peak_values(img, 0.25)
    want = []
    for y in range(23):
        for x in range(19):
            nb = img[max(y - 1, 0):y + 2, max(x - 1, 0):x + 2]
            if img[y, x] >= nb.max() and img[y, x] > 0.25:
                want.append(img[y, x])
    assert np.array_equal(got, np.array(want))


def test_quantile_order_statistics():
    v = [5.0, 1.0, 3.0, 2.0, 4.0]
    assert ob.quantile(v, 0.0) == 1.0 and ob.quantile(v, 1.0) == 5.0
    assert ob.quantile(v, 0.5) == 3.0
    assert ob.quantile(v, 0.1) == pytest.approx(1.4, abs=1e-15)   # pos 0.4 between 1 and 2
    assert ob.quantile(v, 2.0) == 5.0 and ob.quantile([], 0.3) == 0.0


def test_dilate_brute_force():
    rng = np.random.default_rng(9)
    m = (rng.random((30, 26)) < 0.02).astype(np.uint8)
    el = ob.ellipse_element(7)
    got = ob.dilate(m, el)
    want = np.zeros_like(m)
    for y, x in zip(*np.nonzero(m)):
        for i in range(7):
            for j in range(7):
                if el[i, j] and 0 <= y + i - 3 < 30 and 0 <= x + j - 3 < 26:
                    want[y + i - 3, x + j - 3] = 1
    assert np.array_equal(got, want)


def test_pipeline_stage_relations():
    rng = np.random.default_rng(3)
    from scipy import ndimage
    gray = ndimage.gaussian_filter(rng.random((96, 96)), 2.0)
    gray = (gray - gray.min()) / (gray.max() - gray.min())
    tgt = np.zeros((96, 96))
    tgt[20:40, 20:60] = 1
    r = ob.binarize_advanced(gray, tgt, 0.1, 0.2, 5, 4)
    assert r["num_peaks"] > 0
    assert np.array_equal(r["initial"], (gray >= r["threshold"]).astype(np.uint8))
    assert np.all(r["filtered"] <= r["initial"])
    assert np.array_equal(r["sraf_initial"], r["filtered"] & (1 - (tgt > 0.5)))
    assert np.all(r["sraf_final"] <= r["sraf_initial"]) and not np.any(r["sraf_final"] & r["keep_out"])
    assert np.all(r["final"] <= ((tgt > 0.5) | r["sraf_final"]))
    # no peak above the floor: every mask empty, target included (imbinarize.cpp:108-117)
    e = ob.binarize_advanced(gray * 0.05, tgt, 0.1, 0.2, 5, 4)
    assert e["num_peaks"] == 0 and e["final"].sum() == 0
