"""Optimization::minimizeOneDimen and brent_opt (utils/optimization.cpp:
185-373), restated for the parity tests of the gamma-shape optimisation: the
bracket around the guess (eps = 50 * tol * guess, the whole interval when
the guess is not lower than both neighbours), Brent's method (Brent 1973 in
the Numerical Recipes form: ITMAX 100, CGOLD 0.3819660, ZEPS 1e-10) and the
"keep the guess if Brent ended higher" rule. Test infrastructure: only the
tests call it, with the oracle's likelihood as the function.

RateGamma::optimizeParameters (model/rategamma.cpp:232-246) is
minimise(f, 0.02, shape, 1000, max(eps, 0.001)) followed by one more
evaluation at the optimum."""
import math

ITMAX = 100
CGOLD = 0.3819660
ZEPS = 1.0e-10
MIN_GAMMA_SHAPE = 0.02      # utils/tools.h:600-602
MAX_GAMMA_SHAPE = 1000.0
TOL_GAMMA_SHAPE = 0.001


def _sign(a, b):
    return abs(a) if b >= 0.0 else -abs(a)


def brent_opt(f, ax, bx, cx, tol, fa, fb, fc):
    """Returns (x, f(x)) of brent_opt."""
    a, b = min(ax, cx), max(ax, cx)
    x, fx = bx, fb
    if fa < fc:
        w, fw, v, fv = ax, fa, cx, fc
    else:
        w, fw, v, fv = cx, fc, ax, fa
    d = e = 0.0
    for _ in range(ITMAX):
        xm = 0.5 * (a + b)
        tol1 = tol * abs(x) + ZEPS
        tol2 = 2.0 * tol1
        if abs(x - xm) <= tol2 - 0.5 * (b - a):
            return x, fx
        if abs(e) > tol1:
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2.0 * (q - r)
            if q > 0.0:
                p = -p
            q = abs(q)
            etemp, e = e, d
            if abs(p) >= abs(0.5 * q * etemp) or p <= q * (a - x) or p >= q * (b - x):
                e = a - x if x >= xm else b - x
                d = CGOLD * e
            else:
                d = p / q
                u = x + d
                if u - a < tol2 or b - u < tol2:
                    d = _sign(tol1, xm - x)
        else:
            e = a - x if x >= xm else b - x
            d = CGOLD * e
        u = x + d if abs(d) >= tol1 else x + _sign(tol1, d)
        fu = f(u)
        if fu <= fx:
            if u >= x:
                a = x
            else:
                b = x
            v, w, x = w, x, u
            fv, fw, fx = fw, fx, fu
        else:
            if u < x:
                a = u
            else:
                b = u
            if fu <= fw or w == x:
                v, w, fv, fw = w, u, fw, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    return x, fx


def minimize_one_dimen(f, xmin, xguess, xmax, tol):
    """Returns (x, f(x)) of minimizeOneDimen."""
    xguess = min(max(xguess, xmin), xmax)
    eps = xguess * tol * 50.0
    ax, bx, cx = max(xguess - eps, xmin), xguess, min(xguess + eps, xmax)
    fb = f(bx)
    fa = f(ax) if ax < bx else fb
    fc = f(cx) if bx < cx else fb
    if fa < fb or fc < fb:
        if ax != xmin:
            fa = f(xmin)
        if cx != xmax:
            fc = f(xmax)
        ax, cx = xmin, xmax
    x, fx = brent_opt(f, ax, bx, cx, tol, fa, fb, fc)
    if fx > fb:
        return bx, f(bx)
    return x, fx


class Counted:
    """f with a count of its evaluations."""

    def __init__(self, f):
        self.f, self.n = f, 0

    def __call__(self, x):
        self.n += 1
        return self.f(x)


def optimize_gamma_shape(neg_lnl, shape, eps=0.0):
    """RateGamma::optimizeParameters: (optimal shape, lnL there, evaluations)."""
    f = Counted(neg_lnl)
    x, _ = minimize_one_dimen(f, MIN_GAMMA_SHAPE, shape, MAX_GAMMA_SHAPE, max(eps, TOL_GAMMA_SHAPE))
    lnl = -f(x)
    return x, lnl, f.n
