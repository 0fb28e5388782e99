"""Pins of the oracle's CTM optimizers (no GPU): the reference's
optimizer-control invariants (tests/test_ctm_optimizer.cpp:181-196) and
properties of the restated LBFGS++ solver.  The line-search trajectory is
PARITY UNPINNED (LBFGS++ is not vendored in the reference)."""
import numpy as np
import pytest

from oracle import ctm as oc
from oracle import imaging as im
from oracle import optics as op

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)


@pytest.fixture(scope="module")
def case64():
    n = 64
    cache = im.Cache.from_socs(op.socs_kernels(n, **OPT, num_kernels=36))
    t = np.zeros((n, n))
    t[23:42, 23:42] = 1.0
    return cache, t


def test_initialize_theta():
    t = np.array([[0.0, 1.0], [0.2, 0.0]])
    th = oc.initialize_theta(t)
    assert np.allclose(th, [[0.7 * np.pi, 0.3 * np.pi], [0.3 * np.pi, 0.7 * np.pi]])


def test_tolerance_stops_after_one_evaluation(case64):
    cache, t = case64
    gd_mask, gd_hist, _ = oc.gradient_descent(cache, t, 0.25, 85.0, 3, 0.9, 1e100)
    lb_mask, lb_hist, _, info = oc.lbfgs(cache, t, 0.25, 85.0, 3, 10, 1e100)
    assert len(gd_hist) == 1 and len(lb_hist) == 1
    assert np.allclose(gd_mask, lb_mask, rtol=1e-12, atol=0)


def test_one_iteration_lbfgs(case64):
    cache, t = case64
    mask, hist, _, info = oc.lbfgs(cache, t, 0.25, 85.0, 1, 10, 0.0)
    assert len(hist) >= 2 and hist[-1] <= hist[0] and np.all(np.isfinite(mask))


def test_lbfgs_on_quadratic():
    # on a convex quadratic the solver converges and every accepted step satisfies strong Wolfe
    rng = np.random.default_rng(0)
    a = rng.standard_normal((30, 30))
    q = a @ a.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    xs = np.linalg.solve(q, b)

    def f(x):
        return 0.5 * x @ q @ x - b @ x, q @ x - b

    its, x, fx, gn, status = oc.lbfgs_minimize(f, np.zeros(30), oc.LBFGSParams(m=5, epsilon=1e-7,
                                                                             max_iterations=200))
    assert status == "converged" and np.allclose(x, xs, atol=1e-6)


def test_zoom_trial_is_cubic_minimiser():
    # exact minimiser of a cubic through the bracket ends
    f = lambda a: (a - 0.3) ** 2 * (a + 2.0)
    df = lambda a: 2 * (a - 0.3) * (a + 2.0) + (a - 0.3) ** 2
    a = oc.zoom_trial(0.0, f(0.0), df(0.0), 1.0, f(1.0), df(1.0))
    assert abs(a - 0.3) < 1e-12
    # outside the safeguard interval -> bisection
    assert oc.zoom_trial(0.0, 0.0, -1.0, 1.0, -10.0, -5.0) == 0.5


def test_gd_history_matches_manual_loop(case64):
    cache, t = case64
    best, hist, theta = oc.gradient_descent(cache, t, 0.25, 85.0, 3, 0.9, 0.0)
    th = oc.initialize_theta(t)
    man = []
    for _ in range(3):
        pe, gt, *_ = im.ctm_evaluate(cache, th, t, 0.25, 85.0)
        man.append(pe)
        th = th - 0.9 * gt
    assert np.allclose(hist, man, rtol=1e-14, atol=0)
    assert np.allclose(theta, th, rtol=0, atol=1e-14)
