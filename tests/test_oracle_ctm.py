"""Pins of the oracle's CTM optimizers (no GPU): the reference's
optimizer-control invariants (tests/test_ctm_optimizer.cpp:181-196) and what
the mathematics fixes for the restated LBFGS++ solver:
* the two-loop direction equals -H g with H the explicit BFGS inverse Hessian
  built from the kept pairs (H0 = s'y / y'y I of the newest);
* every accepted step of a real CTM trajectory meets the strong-Wolfe
  conditions (ftol 1e-4, wolfe 0.9);
* 1-D line-search traces follow the quadratic-interpolation rule, evaluated
  here in exact rational arithmetic, with the bracket updates of Nocedal &
  Wright Alg. 3.5 / 3.6 written out trial by trial;
* when the trials run out, the best point of sufficient decrease is returned.
Agreement with LBFGS++ itself stays unpinned (its source is not vendored)."""
from fractions import Fraction

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


def _explicit_bfgs_inverse(pairs, n):
    """H from H0 = gamma I (gamma = s'y / y'y of the newest pair) and the BFGS
    update H <- (I - rho s y') H (I - rho y s') + rho s s' per pair, oldest
    first (Nocedal & Wright eq. 6.17)."""
    s_l, y_l = pairs[-1]
    h = (s_l @ y_l) / (y_l @ y_l) * np.eye(n)
    for s, y in pairs:
        rho = 1.0 / (y @ s)
        v = np.eye(n) - rho * np.outer(y, s)
        h = v.T @ h @ v + rho * np.outer(s, s)
    return h


@pytest.mark.parametrize("n,m,count", [(6, 3, 3), (12, 5, 9), (12, 10, 4)])
def test_two_loop_equals_explicit_bfgs(n, m, count):
    rng = np.random.default_rng(n * 100 + count)
    a = rng.standard_normal((n, n))
    q = a @ a.T + n * np.eye(n)           # SPD: every pair has s'y > 0
    hist = oc.LbfgsHistory(m)
    kept = []
    for i in range(count):
        s = rng.standard_normal(n)
        y = q @ s + 0.1 * rng.standard_normal(n)
        if s @ y <= 0:
            continue
        assert hist.add(s, y)
        kept.append((s, y))
        if i == 1:                         # a pair with s'y < 0 is rejected and changes nothing
            assert not hist.add(s, -y)
    g = rng.standard_normal(n)
    want = -_explicit_bfgs_inverse(kept[-m:], n) @ g
    got = hist.direction(g)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    # no pair: steepest descent (theta = 1)
    assert np.array_equal(oc.LbfgsHistory(m).direction(g), -g)


def test_accepted_steps_meet_strong_wolfe(case64):
    cache, t = case64
    steps = []

    def evalf(theta):
        pe, gt, *_ = im.ctm_evaluate(cache, theta, t, 0.225, 50.0)
        return pe, gt

    def on_step(step, xp, fp, gp, d, x, f, g):
        dg0 = float(np.dot(gp.ravel(), d.ravel()))
        dg = float(np.dot(g.ravel(), d.ravel()))
        steps.append((step, fp, f, dg0, dg))

    p = oc.LBFGSParams(m=10, epsilon=0.0, max_iterations=12)
    oc.lbfgs_minimize(evalf, oc.initialize_theta(t), p, on_step=on_step)
    assert len(steps) == 12
    for step, fp, f, dg0, dg in steps:
        assert dg0 < 0.0
        assert f - fp <= step * p.ftol * dg0            # sufficient decrease
        assert abs(dg) <= -p.wolfe * dg0                # curvature (strong)


def _quartic(x):
    return float(x[0] ** 4 / 4 - x[0]), np.array([x[0] ** 3 - 1.0])


def _fq(a):
    return a ** 4 / 4 - a


def _dq(a):
    return a ** 3 - 1


def _vertex(lo, hi):
    """Minimiser of the quadratic through (lo, f(lo)) with slope f'(lo) and
    through (hi, f(hi)), exact: lo - f'(lo) (hi-lo)^2 / (2 (f(hi) - f(lo) - f'(lo) (hi-lo)))."""
    w = hi - lo
    return lo - _dq(lo) * w * w / (2 * (_fq(hi) - _fq(lo) - _dq(lo) * w))


def _run_trace(a0, wolfe):
    p = oc.LBFGSParams()
    p.wolfe = wolfe
    tr = []
    step, _, _, _ = oc.line_search(_quartic, np.array([0.0]), 0.0, np.array([-1.0]), np.array([1.0]), a0, p,
                                   trace=tr)
    return tr, step


def _check_zoom(trials, lo, hi, script):
    """Each zoom trial against the exact vertex of its bracket.  script: one
    entry per zoom trial, 'v' (vertex) or 'm' (midpoint: the vertex lies within
    1 % of the bracket width of an end), then the bracket update: 'lo' (trial
    becomes lo), 'swap' (hi <- lo, then lo <- trial), 'hi' (sufficient decrease
    fails: hi <- trial) or 'accept'.  Bracket ends are the trials the search
    actually made (exact rationals of the floats), so the rational sizes stay
    bounded."""
    assert len(trials) == len(script)
    for t, (kind, upd) in zip(trials, script):
        v = _vertex(lo, hi)
        if kind == "m":
            assert min(abs(v - lo), abs(v - hi)) < Fraction(1, 100) * abs(hi - lo)
            want = (lo + hi) / 2
        else:
            assert min(abs(v - lo), abs(v - hi)) >= Fraction(1, 100) * abs(hi - lo)
            want = v
        assert abs(Fraction(t) - want) <= Fraction(1, 10 ** 12) * abs(want)
        a = Fraction(t)
        if upd == "lo":
            assert _fq(a) < _fq(lo) and _dq(a) * (hi - lo) < 0
            lo = a
        elif upd == "swap":
            assert _fq(a) < _fq(lo) and _dq(a) * (hi - lo) >= 0
            lo, hi = a, lo
        elif upd == "hi":
            assert _fq(a) > a * Fraction(-1, 10 ** 4) or _fq(a) >= _fq(lo)
            hi = a


def test_line_search_trace_decrease_failure_then_zoom():
    # a0 = 4 fails sufficient decrease: bracket (lo 0, hi 4); the zoom walks lo up by
    # vertices until a vertex lies within 1 % of the bracket width of an end (midpoint
    # taken, which fails decrease and becomes hi), then vertices again until |f'| <= 0.1
    script = [("v", "lo")] * 11 + [("m", "hi"), ("v", "lo"), ("v", "lo"), ("v", "accept")]
    tr, step = _run_trace(4.0, 0.1)
    assert tr[0] == 4.0
    _check_zoom(tr[1:], Fraction(0), Fraction(4), script)
    assert step == tr[-1] and abs(_dq(step)) <= 0.1
    # a cubic (Hermite) trial would use f'(hi) = 63 and land elsewhere than 1/8
    assert tr[1] == 0.125


def test_line_search_trace_reversed_bracket():
    # a0 = 1.5 passes decrease but f' > 0: bracket (lo 1.5, hi 0) reversed; the
    # first vertex has f' < 0 with f'(a)(hi - lo) >= 0, so hi takes the old lo
    script = [("v", "swap"), ("v", "lo"), ("v", "accept")]
    tr, step = _run_trace(1.5, 0.1)
    assert tr[0] == 1.5
    _check_zoom(tr[1:], Fraction(3, 2), Fraction(0), script)
    assert step == tr[-1]


def test_line_search_expansion():
    # decrease holds and f' < 0 stays: the step doubles until |f'| <= 0.9 f'(0)
    tr, step = _run_trace(0.05, 0.9)
    assert tr == [0.05, 0.1, 0.2, 0.4, 0.8] and step == 0.8


def test_line_search_exhaustion_returns_best_decrease_point():
    # wolfe 1e-9 cannot be met: after 1 bracketing and 20 zoom trials the search
    # returns the best trial of sufficient decrease, not the last trial, and no error
    p = oc.LBFGSParams()
    p.wolfe = 1e-9
    tr = []
    step, x, f, g = oc.line_search(_quartic, np.array([0.0]), 0.0, np.array([-1.0]), np.array([1.0]), 4.0, p,
                                   trace=tr)
    assert len(tr) == 21
    ok = [(_fq(a), a) for a in tr if _fq(a) <= a * p.ftol * -1.0]
    assert (f, step) == min(ok)
    assert x[0] == step and g[0] == step ** 3 - 1.0


def test_line_search_rejects_ascent_direction():
    with pytest.raises(oc.LineSearchError):
        oc.line_search(_quartic, np.array([0.0]), 0.0, np.array([-1.0]), np.array([-1.0]), 1.0, oc.LBFGSParams())


def test_quad_interp_guards():
    # non-finite candidate, a candidate outside the bracket, and one too close to an end -> midpoint
    assert oc.quad_interp(0.0, 1.0, 0.0, 0.0, 0.0) == 0.5
    assert oc.quad_interp(0.0, 1.0, 0.0, -10.0, -1.0) == 0.5
    assert oc.quad_interp(0.0, 1.0, 0.0, 1e6, -1.0) == 0.5
    # reversed bracket: q(a) = (a - 2) + 3/4 (a - 2)^2 passes (2, 0) with slope 1 and (0, 1); vertex 4/3
    assert oc.quad_interp(2.0, 0.0, 0.0, 1.0, 1.0) == pytest.approx(4.0 / 3.0, rel=1e-15)


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
