"""Macroscopic dielectric function on the grid frequencies (oracle restatement).

Follows ``interpolate_dielec_func`` (src/core/dielecmodel.cpp:65-104), which maps the values
eps_M(i omega) a caller gives on its own imaginary frequencies to the minimax nodes before the
head of the dielectric matrix is replaced:

* option 0: the values as given (one per grid frequency);
* option 1: ``interp_cubic_spline`` (src/math/interpolate.cpp:9-153). Its steps are kept as they
  are, including two that differ from a textbook natural spline: the right-hand side holds the
  slopes 3 (y[i+1] - y[i]) / h[i] themselves, and the back substitution runs inside the forward
  sweep, so it reads entries of z that the sweep has not yet written (zero). Outside the data the
  value is the cubic Hermite extrapolation of the first or last interval, with slope 0 at the
  last point;
* option 2: the double Havriliak-Negami model (dielecmodel.cpp:33-63) fitted by
  ``LevMarqFitting::fit`` (src/math/fitting.cpp:80-196): Levenberg-Marquardt on the summed
  squared residuals, damping 0.1 divided or multiplied by 10, the normal matrix factored by
  Cholesky, at most 10000 iterations, stop when the error falls by less than 1e-8; all
  parameters start at 1 except the two static constants, which start at the first datum.

Both static constants start equal and the two terms receive equal gradients, so the fitted two
terms stay equal: the reference's double model acts as one Havriliak-Negami term of twice the
strength.

Pins: tests/test_oracle_ref_dielec.py holds both against the reference's own interpolate.cpp
and fitting.cpp (oracle/ref_dielec.cpp); the spline passes through the data with continuous value
and second derivative (its first derivative jumps at the knots, so unlike a textbook spline it does
not reproduce a straight line), and the fit reproduces data that two equal terms represent exactly
(1 + A / (B + u^2)). tests/test_oracle_known_answers.py runs the reference's own known answer for
the fit (src/test/test_fitting.cpp:9-40: sin(a x) from a = 1 recovers a to 1e-8).
"""
from __future__ import annotations

import math
from typing import List, Sequence


def spline(xs: Sequence[float], ys: Sequence[float], xs_new: Sequence[float]) -> List[float]:
    """interp_cubic_spline (interpolate.cpp:9-153), step for step."""
    n = len(xs)
    if n < 2:
        raise ValueError("cubic spline needs at least 2 points")
    h = [xs[i + 1] - xs[i] for i in range(n - 1)]
    alpha = [3 * (ys[i + 1] - ys[i]) / h[i] for i in range(n - 1)]
    l = [0.0] * n
    mu = [0.0] * (n - 1)
    l[0] = 1.0
    mu[0] = 0.0
    for i in range(1, n - 1):
        l[i] = 2 * (xs[i + 1] - xs[i - 1]) - h[i - 1] * mu[i - 1]
        mu[i] = h[i] / l[i]
    l[n - 1] = 1.0
    z = [0.0] * n
    c = [0.0] * n
    for i in range(1, n):
        z[i] = (alpha[i - 1] - h[i - 1] * z[i - 1]) / l[i]
        c[n - i - 1] = z[n - i] - mu[n - i - 1] * c[n - i]
    a, b, d = [0.0] * n, [0.0] * n, [0.0] * n
    for i in range(n - 1):
        a[i] = ys[i]
        b[i] = (ys[i + 1] - ys[i]) / h[i] - h[i] * (c[i + 1] + 2 * c[i]) / 3
        d[i] = (c[i + 1] - c[i]) / (3 * h[i])
    a[n - 1] = ys[n - 1]

    def at(x):
        if x < xs[0] or x > xs[n - 1]:
            k = 0 if x < xs[0] else n - 2
            x0, x1 = xs[k], xs[k + 1]
            y0, y1, m0, m1 = a[k], a[k + 1], b[k], b[k + 1]
            hh = x1 - x0
            dx, dy, dm = x - x0, y1 - y0, m1 - m0
            return (y0 + m0 * dx
                    + ((3 * dy / math.pow(hh, 2)) - (2 * m0 / hh) - (dm / hh)) * math.pow(dx, 2)
                    + ((-2 * dy / math.pow(hh, 3)) + (m0 / math.pow(hh, 2)) + (dm / math.pow(hh, 2)))
                    * math.pow(dx, 3))
        for i in range(n - 1):
            if xs[i] <= x <= xs[i + 1]:
                dx = x - xs[i]
                return a[i] + b[i] * dx + c[i] * math.pow(dx, 2) + d[i] * math.pow(dx, 3)
        return 0.0

    return [at(x) for x in xs_new]


def dhn(u: float, p: Sequence[float]) -> float:
    """Double Havriliak-Negami eps(i u) (dielecmodel.cpp:33-39)."""
    return (1.0 + (p[0] - 1.0) / math.pow(1.0 + math.pow(u * p[3], p[1]), p[2])
            + (p[4] - 1.0) / math.pow(1.0 + math.pow(u * p[7], p[5]), p[6]))


def dhn_grad(u: float, p: Sequence[float]) -> List[float]:
    """Its gradient in the parameters (dielecmodel.cpp:41-63)."""
    pw, lg = math.pow, math.log
    return [
        1.0 / pw(1.0 + pw(u * p[3], p[1]), p[2]),
        (p[0] - 1.0) * (-p[2]) / pw(1.0 + pw(u * p[3], p[1]), p[2] + 1) * lg(u * p[3]) * pw(u * p[3], p[1]),
        (1.0 - p[0]) * lg(1.0 + pw(u * p[3], p[1])) / pw(1.0 + pw(u * p[3], p[1]), p[2]),
        (p[0] - 1.0) * (-p[2]) / pw(1.0 + pw(u * p[3], p[1]), p[2] + 1) * p[1] / p[3] * pw(u * p[3], p[1]),
        1.0 / pw(1.0 + pw(u * p[7], p[5]), p[6]),
        (p[4] - 1.0) * (-p[6]) / pw(1.0 + pw(u * p[7], p[5]), p[6] + 1) * lg(u * p[7]) * pw(u * p[7], p[5]),
        (1.0 - p[4]) * lg(1.0 + pw(u * p[7], p[5])) / pw(1.0 + pw(u * p[7], p[5]), p[6]),
        (p[4] - 1.0) * (-p[6]) / pw(1.0 + pw(u * p[7], p[5]), p[6] + 1) * p[5] / p[7] * pw(u * p[7], p[5]),
    ]


def _sq_error(p, xs, ys, func):
    e = 0.0
    for x, y in zip(xs, ys):
        r = func(x, p) - y
        e += r * r
    return e


def _cholesky(a, n):
    """Lower Cholesky factor of the lower triangle of a, or None below the 1e-30 pivot floor."""
    lo = [[0.0] * n for _ in range(n)]
    for i in range(n):
        for j in range(i):
            s = 0.0
            for k in range(j):
                s += lo[i][k] * lo[j][k]
            lo[i][j] = (a[i][j] - s) / lo[j][j]
        s = 0.0
        for k in range(i):
            s += lo[i][k] * lo[i][k]
        s = a[i][i] - s
        if s < 1e-30:
            return None
        lo[i][i] = math.sqrt(s)
    return lo


def _solve(lo, b, n):
    x = [0.0] * n
    for i in range(n):
        s = 0.0
        for j in range(i):
            s += lo[i][j] * x[j]
        x[i] = (b[i] - s) / lo[i][i]
    for i in range(n - 1, -1, -1):
        s = 0.0
        for j in range(i + 1, n):
            s += lo[j][i] * x[j]
        x[i] = (x[i] - s) / lo[i][i]
    return x


def levmarq_fit(pars, xs, ys, func, grad, maxiter=10000, lam=0.1, up=10.0, down=10.0, target=1e-8):
    """LevMarqFitting::fit (fitting.cpp:80-182): returns the fitted parameters."""
    p = list(pars)
    npar = len(p)
    down = 1.0 / down
    err = _sq_error(p, xs, ys, func)
    h = [[0.0] * npar for _ in range(npar)]
    newpar = [0.0] * npar
    newerr = 0.0
    derr = 0.0
    it = 0
    while it < maxiter:
        d = [0.0] * npar
        for i in range(npar):
            for j in range(i + 1):
                h[i][j] = 0.0
        for x, y in zip(xs, ys):
            g = grad(x, p)
            for i in range(npar):
                d[i] += (y - func(x, p)) * g[i] * 1.0
                for j in range(i + 1):
                    h[i][j] += g[i] * g[j] * 1.0
        mult = 1 + lam
        ill = True
        while ill and it < maxiter:
            for i in range(npar):
                h[i][i] = h[i][i] * mult
            lo = _cholesky(h, npar)
            ill = lo is None
            if not ill:
                delta = _solve(lo, d, npar)
                newpar = [p[i] + delta[i] for i in range(npar)]
                newerr = _sq_error(newpar, xs, ys, func)
                derr = newerr - err
                ill = derr > 0
            if ill:
                mult = (1 + lam * up) / (1 + lam)
                lam *= up
                it += 1
        p = list(newpar)
        err = newerr
        lam *= down
        if (not ill) and (-derr < target):
            break
        it += 1
    return p


def interpolate(option: int, om_in, df_in, om_target):
    """interpolate_dielec_func (dielecmodel.cpp:65-104)."""
    if option == 0:
        if len(om_in) != len(om_target):
            raise ValueError("option 0 needs one value per grid frequency")
        return list(df_in)
    if option == 1:
        return spline(list(om_in), list(df_in), list(om_target))
    if option == 2:
        p0 = [1.0] * 8
        p0[0] = p0[4] = df_in[0]
        p = levmarq_fit(p0, list(om_in), list(df_in), dhn, lambda x, q: dhn_grad(x, q))
        return [dhn(x, p) for x in om_target]
    raise ValueError(f"unsupported option {option}")
