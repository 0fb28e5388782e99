"""The seeded Manhattan target generator (no GPU): rectangles with sides of
10..30 px (40-120 nm at 4 nm/px), at least 10 px (40 nm) apart also across the
periodic tile boundary, ~35 % density, and the same bits for the same seed."""
import numpy as np
import pytest
from scipy import ndimage

from conftest import targets


@pytest.mark.parametrize("n,seed", [(128, 0), (200, 7), (256, 3)])
def test_rectangles_sides_and_spacing(n, seed):
    t = targets().manhattan_target(n, seed)
    big = np.tile(t, (3, 3))                     # periodic neighbours around the centre copy
    lab, num = ndimage.label(big > 0)
    objs = ndimage.find_objects(lab)
    checked = 0
    for i, sl in enumerate(objs, start=1):
        ys, xs = sl
        if not (n <= ys.start < 2 * n and n <= xs.start < 2 * n):
            continue                             # judge each rectangle once, by its copy in the centre
        comp = lab[sl] == i
        assert comp.all(), "component is not a filled rectangle"
        h, w = comp.shape
        assert 10 <= h <= 30 and 10 <= w <= 30
        # no other rectangle within 9 px (L-inf): spacing >= 10 px
        y0, y1 = max(ys.start - 9, 0), min(ys.stop + 9, 3 * n)
        x0, x1 = max(xs.start - 9, 0), min(xs.stop + 9, 3 * n)
        around = lab[y0:y1, x0:x1]
        assert set(np.unique(around)) <= {0, i}
        checked += 1
    assert checked > 0


def test_density_and_determinism():
    tg = targets()
    d = [tg.density(tg.manhattan_target(512, s)) for s in range(4)]
    assert all(0.30 <= x <= 0.42 for x in d), d
    assert np.array_equal(tg.manhattan_target(512, 11), tg.manhattan_target(512, 11))
    assert not np.array_equal(tg.manhattan_target(512, 11), tg.manhattan_target(512, 12))
    assert tg.manhattan_target(96, 1, height=64).shape == (64, 96)
