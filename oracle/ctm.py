"""CTM optimizer (theta init, gradient descent, L-BFGS), FP64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates source/optimizer/CTM_Optimizer.cpp.  The L-BFGS solver is the
reference's third-party LBFGS++ (yixuan/LBFGSpp @ c524a407, "v0.4.0"), which is
not vendored: this file restates its published algorithm (two-loop recursion
with H0 = (s'y / y'y) I, pairs kept only when s'y > eps*y'y, first step
1/||d||, later steps 1, strong-Wolfe line search of Nocedal & Wright
Alg. 3.5/3.6 with ftol 1e-4, wolfe 0.9, max 20 trials).  PARITY UNPINNED for
the line-search trajectory beyond the reference invariants of
tests/test_ctm_optimizer.cpp:181-196.
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
        self.min_step = 1e-20
        self.max_step = 1e20
        self.ftol = 1e-4
        self.wolfe = 0.9


def _cubic_min(a_lo, f_lo, g_lo, a_hi, f_hi, g_hi):
    """Minimiser of the cubic interpolating (a, f, f') at both ends
    (Nocedal & Wright eq. 3.59); None when it does not exist."""
    d1 = g_lo + g_hi - 3.0 * (f_lo - f_hi) / (a_lo - a_hi)
    rad = d1 * d1 - g_lo * g_hi
    if rad < 0.0:
        return None
    d2 = math.copysign(math.sqrt(rad), a_hi - a_lo)
    den = g_hi - g_lo + 2.0 * d2
    if den == 0.0:
        return None
    return a_hi - (a_hi - a_lo) * (g_hi + d2 - d1) / den


def zoom_trial(a_lo, f_lo, g_lo, a_hi, f_hi, g_hi):
    """Safeguarded cubic trial inside the bracket (shared rule of oracle and host)."""
    lo, hi = min(a_lo, a_hi), max(a_lo, a_hi)
    width = hi - lo
    a = _cubic_min(a_lo, f_lo, g_lo, a_hi, f_hi, g_hi)
    if a is None or not math.isfinite(a) or a < lo + 0.1 * width or a > hi - 0.1 * width:
        a = 0.5 * (lo + hi)
    return a


def line_search(evalf, x0, f0, g0, d, step, p: LBFGSParams):
    """Strong-Wolfe line search (Nocedal & Wright Alg. 3.5 + 3.6).

    evalf(x) -> (f, g).  Returns (step, x, f, g).  Raises LineSearchError when
    max_linesearch trials pass without an acceptable point.
    """
    dg0 = float(np.dot(g0.ravel(), d.ravel()))
    if dg0 >= 0.0:
        raise LineSearchError("the moving direction does not decrease the objective function value")
    a_prev, f_prev, dg_prev = 0.0, f0, dg0
    a = step
    trials = 0

    def trial(alpha):
        x = x0 + alpha * d
        f, g = evalf(x)
        return x, f, g, float(np.dot(g.ravel(), d.ravel()))

    def zoom(a_lo, f_lo, g_lo, a_hi, f_hi, g_hi, trials):
        while trials < p.max_linesearch:
            aj = zoom_trial(a_lo, f_lo, g_lo, a_hi, f_hi, g_hi)
            x, f, g, dg = trial(aj)
            trials += 1
            if f > f0 + p.ftol * aj * dg0 or f >= f_lo:
                a_hi, f_hi, g_hi = aj, f, dg
            else:
                if abs(dg) <= -p.wolfe * dg0:
                    return aj, x, f, g
                if dg * (a_hi - a_lo) >= 0.0:
                    a_hi, f_hi, g_hi = a_lo, f_lo, g_lo
                a_lo, f_lo, g_lo = aj, f, dg
        raise LineSearchError("the line search routine reached the maximum number of iterations")

    while trials < p.max_linesearch:
        if a < p.min_step:
            raise LineSearchError("the line search step became smaller than the minimum value allowed")
        x, f, g, dg = trial(a)
        trials += 1
        if f > f0 + p.ftol * a * dg0 or (trials > 1 and f >= f_prev):
            return zoom(a_prev, f_prev, dg_prev, a, f, dg, trials)
        if abs(dg) <= -p.wolfe * dg0:
            return a, x, f, g
        if dg >= 0.0:
            return zoom(a, f, dg, a_prev, f_prev, dg_prev, trials)
        a_prev, f_prev, dg_prev = a, f, dg
        a = min(2.0 * a, p.max_step)
    raise LineSearchError("the line search routine reached the maximum number of iterations")


def lbfgs_minimize(evalf, x, p: LBFGSParams):
    """LBFGSSolver::minimize restated.  evalf(x) -> (f, g); returns
    (iterations, x, f, gnorm, status) where status is 'converged',
    'max_iterations' or 'line_search_failed'."""
    f, g = evalf(x)
    gnorm = float(np.linalg.norm(g))
    if gnorm <= p.epsilon or gnorm <= p.epsilon_rel * float(np.linalg.norm(x)):
        return 1, x, f, gnorm, "converged"
    d = -g
    step = 1.0 / float(np.linalg.norm(d))
    s_list, y_list, ys_list = [], [], []
    theta = 1.0
    eps = np.finfo(np.float64).eps
    k = 1
    while True:
        xp, gp = x, g
        try:
            step, x, f, g = line_search(evalf, xp, f, gp, d, step, p)
        except LineSearchError:
            return k, xp, f, gnorm, "line_search_failed"
        gnorm = float(np.linalg.norm(g))
        if gnorm <= p.epsilon or gnorm <= p.epsilon_rel * float(np.linalg.norm(x)):
            return k, x, f, gnorm, "converged"
        if p.max_iterations != 0 and k >= p.max_iterations:
            return k, x, f, gnorm, "max_iterations"
        s = (x - xp).ravel()
        y = (g - gp).ravel()
        ys = float(np.dot(s, y))
        yy = float(np.dot(y, y))
        if ys > eps * yy:
            s_list.append(s)
            y_list.append(y)
            ys_list.append(ys)
            if len(s_list) > p.m:
                s_list.pop(0)
                y_list.pop(0)
                ys_list.pop(0)
            theta = yy / ys
        # two-loop recursion, H0 = I / theta
        q = -g.ravel().copy()
        alphas = [0.0] * len(s_list)
        for i in range(len(s_list) - 1, -1, -1):
            alphas[i] = float(np.dot(s_list[i], q)) / ys_list[i]
            q -= alphas[i] * y_list[i]
        q /= theta
        for i in range(len(s_list)):
            beta = float(np.dot(y_list[i], q)) / ys_list[i]
            q += (alphas[i] - beta) * s_list[i]
        d = q.reshape(g.shape)
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
