"""CTM optimizer (theta init, gradient descent, L-BFGS), FP64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates source/optimizer/CTM_Optimizer.cpp.  The L-BFGS solver is the
reference's third-party LBFGS++ (yixuan/LBFGSpp @ c524a407, "v0.4.0",
CMakeLists.txt:98-115), which is not vendored.  CTM_Optimizer.cpp:194-200
constructs LBFGSSolver<double> with the default line-search class,
LineSearchNocedalWright, and LBFGSParam defaults apart from m, epsilon,
epsilon_rel = 0 and max_iterations.  This file restates that published
algorithm: two-loop recursion with H0 = (s'y / y'y) I of the newest kept pair,
pairs kept only when s'y > eps*y'y, first step 1/||d||, later steps 1, and the
Nocedal-Wright strong-Wolfe search with expansion 2, quadratic-interpolation
zoom, ftol 1e-4, wolfe 0.9, max_linesearch 20.
Pinned here by mathematics (tests/test_oracle_ctm.py): the two-loop direction
equals the explicit BFGS inverse-Hessian product, every accepted step meets
the strong-Wolfe conditions, and a 1-D trace follows the interpolation rule
computed in exact rational arithmetic.  PARITY UNPINNED against LBFGS++ itself
(its source is absent, so the restatement cannot be run against it).
"""
from __future__ import annotations

import math

import numpy as np

from .imaging import Cache, ctm_evaluate, forward, mask_from_theta, pe_gradient, pe_loss

PI = 3.14159265358979323846


def initialize_theta(pattern: np.ndarray, tbias: float = 0.3) -> np.ndarray:
    """CTM_Optimizer::initialize_theta (CTM_Optimizer.cpp:66-73):
    0.3*pi where pattern > 0, 0.7*pi elsewhere."""
    return np.where(pattern > 0.0, tbias * PI, (1.0 - tbias) * PI)


def gradient_descent(cache: Cache, target, threshold, alpha, iterations: int, lr: float, tol: float = 0.0,
                     theta0=None):
    """CTM_Optimizer::_optimize_gradient_descent (CTM_Optimizer.cpp:109-173).

    Returns (best_mask, history, theta_final).
    """
    theta = initialize_theta(target) if theta0 is None else np.array(theta0, dtype=np.float64)
    mask = mask_from_theta(theta)
    best = mask.copy()
    min_pe = np.finfo(np.float64).max
    history = []
    for _ in range(iterations):
        aerial, wafer, fields = forward(cache, mask, threshold, alpha)
        gm = pe_gradient(cache, wafer, target, fields, alpha)
        pe = pe_loss(wafer, target)
        history.append(pe)
        if pe < min_pe:
            min_pe = pe
            best = mask.copy()
        gt = -0.5 * gm * np.sin(theta)
        if float(np.linalg.norm(gt)) <= tol:
            break
        theta = theta - lr * gt
        mask = mask_from_theta(theta)
    return best, history, theta


class LineSearchError(RuntimeError):
    pass


class LBFGSParams:
    """LBFGS++ LBFGSParam defaults used by CTM_Optimizer (CTM_Optimizer.cpp:194-198)."""

    def __init__(self, m=10, epsilon=0.0, max_iterations=50):
        self.m = m
        self.epsilon = epsilon
        self.epsilon_rel = 0.0
        self.max_iterations = max_iterations
        self.max_linesearch = 20
        self.ftol = 1e-4
        self.wolfe = 0.9


def quad_interp(step_lo, step_hi, fx_lo, fx_hi, dg_lo):
    """Zoom trial of LBFGS++'s LineSearchNocedalWright (v0.4.0, restated):
    the minimiser of the quadratic through (step_lo, fx_lo) with slope dg_lo
    at step_lo and through (step_hi, fx_hi), written as
        (fdiff * step_lo - smid * sdiff * dg_lo) / (fdiff - sdiff * dg_lo)
    with fdiff = fx_hi - fx_lo, sdiff = step_hi - step_lo, smid the midpoint.
    The midpoint (bisection) is taken instead when that value is not finite,
    lies on or outside the bracket, or within 1 % of the bracket width of
    either end."""
    fdiff = fx_hi - fx_lo
    sdiff = step_hi - step_lo
    smid = (step_hi + step_lo) / 2.0
    num = fdiff * step_lo - smid * sdiff * dg_lo
    den = fdiff - sdiff * dg_lo
    if den == 0.0 or not math.isfinite(num) or not math.isfinite(den):
        cand = math.nan
    else:
        cand = num / den
    if not math.isfinite(cand):
        return smid
    end_dist = min(abs(cand - step_lo), abs(cand - step_hi))
    near_end = end_dist < 0.01 * abs(sdiff)
    if cand <= min(step_lo, step_hi) or cand >= max(step_lo, step_hi) or near_end:
        return smid
    return cand


def line_search(evalf, xp, fx_init, g_init, d, step, p: LBFGSParams, trace=None):
    """LineSearchNocedalWright::LineSearch restated (LBFGS++ v0.4.0; Nocedal &
    Wright Alg. 3.5 bracketing with expansion 2, Alg. 3.6 zoom with
    quad_interp trials).

    evalf(x) -> (f, g).  Returns (step, x, f, g).  The search ends at the first
    trial meeting the strong-Wolfe conditions.  When max_linesearch
    non-terminal trials are used up it returns the best point of sufficient
    decrease seen (step_lo), as LBFGS++ does; it raises LineSearchError where
    LBFGS++ throws (ascent direction, no step of sufficient decrease, a zoom
    trial equal to a bracket end).  trace (a list, optional) receives every
    trial step."""
    if step <= 0.0:
        raise ValueError("'step' must be positive")
    dg_init = float(np.dot(g_init.ravel(), d.ravel()))
    if dg_init > 0.0:
        raise LineSearchError("the moving direction increases the objective function value")
    test_decr = p.ftol * dg_init
    test_curv = -p.wolfe * dg_init
    step_lo, fx_lo, dg_lo, x_lo, g_lo = 0.0, fx_init, dg_init, xp, g_init
    step_hi = fx_hi = None

    def trial(alpha):
        if trace is not None:
            trace.append(alpha)
        x = xp + alpha * d
        f, g = evalf(x)
        return x, f, g, float(np.dot(g.ravel(), d.ravel()))

    it = 0
    # bracketing (Alg. 3.5)
    while True:
        x, fx, g, dg = trial(step)
        if fx - fx_init > step * test_decr or (0.0 < step_lo and fx >= fx_lo):
            step_hi, fx_hi = step, fx
            break
        if abs(dg) <= test_curv:
            return step, x, fx, g
        step_hi, fx_hi = step_lo, fx_lo
        step_lo, fx_lo, dg_lo, x_lo, g_lo = step, fx, dg, x, g
        if dg >= 0.0:
            break
        it += 1
        if it >= p.max_linesearch:
            return step_lo, x_lo, fx_lo, g_lo
        step *= 2.0
    # zoom (Alg. 3.6)
    while True:
        step = quad_interp(step_lo, step_hi, fx_lo, fx_hi, dg_lo)
        x, fx, g, dg = trial(step)
        if fx - fx_init > step * test_decr or fx >= fx_lo:
            if step == step_hi:
                raise LineSearchError("the line search routine failed, possibly due to insufficient numeric "
                                      "precision")
            step_hi, fx_hi = step, fx
        else:
            if abs(dg) <= test_curv:
                return step, x, fx, g
            if dg * (step_hi - step_lo) >= 0.0:
                step_hi, fx_hi = step_lo, fx_lo
            if step == step_lo:
                raise LineSearchError("the line search routine failed, possibly due to insufficient numeric "
                                      "precision")
            step_lo, fx_lo, dg_lo, x_lo, g_lo = step, fx, dg, x, g
        it += 1
        if it >= p.max_linesearch:
            if step_lo <= 0.0:
                raise LineSearchError("the line search routine failed, unable to sufficiently decrease the "
                                      "function value")
            return step_lo, x_lo, fx_lo, g_lo


class LbfgsHistory:
    """BFGSMat restated (LBFGS++ BFGSMat.h): at most m correction pairs,
    add_correction keeps a pair only when s'y > eps y'y (the check in
    LBFGSSolver::minimize) and drops the oldest beyond m; theta = y'y / s'y of
    the newest kept pair (1 before the first); apply_Hv(g, -1) is the two-loop
    recursion (Nocedal & Wright Alg. 7.4) with H0 = I / theta."""

    def __init__(self, m):
        self.m = m
        self.s, self.y = [], []
        self.theta = 1.0

    def add(self, s, y) -> bool:
        s = np.asarray(s, dtype=np.float64).ravel()
        y = np.asarray(y, dtype=np.float64).ravel()
        ys = float(np.dot(s, y))
        yy = float(np.dot(y, y))
        if not ys > np.finfo(np.float64).eps * yy:
            return False
        self.s.append(s)
        self.y.append(y)
        if len(self.s) > self.m:
            self.s.pop(0)
            self.y.pop(0)
        self.theta = yy / ys
        return True

    def direction(self, g):
        """d = -H g."""
        q = -np.asarray(g, dtype=np.float64).ravel().copy()
        ys = [float(np.dot(s, y)) for s, y in zip(self.s, self.y)]
        alphas = [0.0] * len(self.s)
        for i in range(len(self.s) - 1, -1, -1):
            alphas[i] = float(np.dot(self.s[i], q)) / ys[i]
            q -= alphas[i] * self.y[i]
        q /= self.theta
        for i in range(len(self.s)):
            beta = float(np.dot(self.y[i], q)) / ys[i]
            q += (alphas[i] - beta) * self.s[i]
        return q


def lbfgs_minimize(evalf, x, p: LBFGSParams, on_step=None):
    """LBFGSSolver::minimize restated.  evalf(x) -> (f, g); returns
    (iterations, x, f, gnorm, status) where status is 'converged',
    'max_iterations' or 'line_search_failed' (where LBFGS++ throws).
    on_step(step, xp, fp, gp, d, x, f, g), optional, sees every accepted step."""
    f, g = evalf(x)
    gnorm = float(np.linalg.norm(g))
    if gnorm <= p.epsilon or gnorm <= p.epsilon_rel * float(np.linalg.norm(x)):
        return 1, x, f, gnorm, "converged"
    d = -g
    step = 1.0 / float(np.linalg.norm(d))
    hist = LbfgsHistory(p.m)
    k = 1
    while True:
        xp, gp, fp = x, g, f
        try:
            step, x, f, g = line_search(evalf, xp, f, gp, d, step, p)
        except LineSearchError:
            return k, xp, fp, gnorm, "line_search_failed"
        if on_step is not None:
            on_step(step, xp, fp, gp, d, x, f, g)
        gnorm = float(np.linalg.norm(g))
        if gnorm <= p.epsilon or gnorm <= p.epsilon_rel * float(np.linalg.norm(x)):
            return k, x, f, gnorm, "converged"
        if p.max_iterations != 0 and k >= p.max_iterations:
            return k, x, f, gnorm, "max_iterations"
        hist.add(x - xp, g - gp)
        d = hist.direction(g).reshape(g.shape)
        step = 1.0
        k += 1


def lbfgs(cache: Cache, target, threshold, alpha, iterations: int, m: int = 10, tol: float = 0.0,
          theta0=None):
    """CTM_Optimizer::optimize_lbfgs (CTM_Optimizer.cpp:189-227).

    Returns (final_mask, history, theta_final, info).
    """
    history = []

    def evalf(theta):
        pe, gt, *_ = ctm_evaluate(cache, theta, target, threshold, alpha)
        history.append(pe)
        return pe, gt

    x0 = initialize_theta(target) if theta0 is None else np.array(theta0, dtype=np.float64)
    p = LBFGSParams(m=m, epsilon=tol, max_iterations=iterations)
    its, x, f, gnorm, status = lbfgs_minimize(evalf, x0, p)
    return mask_from_theta(x), history, x, {"iterations": its, "final_pe": f, "grad_norm": gnorm,
                                             "status": status}
