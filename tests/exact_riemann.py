"""Exact solution of the 1D Euler Riemann problem (ideal gas).

Textbook algorithm (Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", 3rd ed., ch. 4): Newton iteration for the star pressure with the
two-rarefaction initial guess, then sampling of the self-similar solution.
Used to pin the oracle's shock-tube results; independent of the code under
test.
"""
import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (B + p))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    pr = p / pk  # rarefaction
    f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    z = (g - 1.0) / (2.0 * g)
    p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
    p = max(p, 1e-12)
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(p - dp, 1e-12)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, g):
    """Density, velocity, pressure at positions x (array) at time t."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((3, len(x)))
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    w = (rl, ul, pl)
                else:
                    r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                    w = (r, us, ps)
            else:
                shl = ul - cl
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cs
                if s <= shl:
                    w = (rl, ul, pl)
                elif s > stl:
                    w = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    w = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    w = (rr, ur, pr)
                else:
                    r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                    w = (r, us, ps)
            else:
                shr = ur + cr
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + cs
                if s >= shr:
                    w = (rr, ur, pr)
                elif s <= str_:
                    w = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    w = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
        out[:, n] = w
    return out
