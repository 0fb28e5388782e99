"""Seeded synthetic Manhattan targets (the reference ships only a few BMPs).

Rows of axis-aligned rectangles on a periodic tile: every row has one height
drawn uniformly from [min_side, max_side] pixels; along a row the rectangles
have widths drawn from the same range, separated by gaps drawn from
[spacing, max_gap]; consecutive rows are `spacing` pixels apart.  The first
row and the first rectangle of each row start at a random offset in
[0, spacing], and the last row / rectangle stops at least `spacing` before
the periodic wrap, so every pair of rectangles is at least `spacing` apart
also across the tile boundary.  At 4 nm/px the defaults (10..30 px sides,
10..20 px gaps) give 40-120 nm features snapped to 4 nm, >= 40 nm spacing and
~36 % pattern density.  Pure integer arithmetic on numpy's PCG64 stream, so
the same seed gives the same target everywhere.
"""
from __future__ import annotations

import numpy as np


def manhattan_target(n: int, seed: int, min_side: int = 10, max_side: int = 30, spacing: int = 10,
                     max_gap: int = 20, height: int | None = None) -> np.ndarray:
    """Binary (0/1) float32 target of shape (height or n, n)."""
    h = n if height is None else height
    w = n
    rng = np.random.default_rng(seed)
    out = np.zeros((h, w), dtype=np.float32)
    y_first = int(rng.integers(0, spacing + 1))
    y = y_first
    while True:
        rh = int(rng.integers(min_side, max_side + 1))
        if y + rh + spacing > h + y_first:
            break
        x_first = int(rng.integers(0, spacing + 1))
        x = x_first
        while True:
            rw = int(rng.integers(min_side, max_side + 1))
            if x + rw + spacing > w + x_first:
                break
            out[np.ix_(np.arange(y, y + rh) % h, np.arange(x, x + rw) % w)] = 1.0
            x += rw + int(rng.integers(spacing, max_gap + 1))
        y += rh + spacing
    return out


def density(t: np.ndarray) -> float:
    return float(t.mean())
