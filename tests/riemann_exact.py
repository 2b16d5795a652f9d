"""Exact solution of the 1D Riemann problem for the ideal-gas Euler equations.

Textbook algorithm (E. F. Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", 3rd ed., ch. 4): Newton iteration for the star-region pressure with the
two-rarefaction initial guess, then sampling of the self-similar solution.  Used by the
tests as an independent closed form for the first-order scheme's convergence; it shares
no code or formula with the scheme under test.
"""
from __future__ import annotations

import math

import numpy as np


def _fk(p, rho_k, p_k, c_k, g):
    if p > p_k:  # shock
        A = 2.0 / ((g + 1.0) * rho_k)
        B = (g - 1.0) / (g + 1.0) * p_k
        q = math.sqrt(A / (p + B))
        return (p - p_k) * q, q * (1.0 - 0.5 * (p - p_k) / (B + p))
    r = p / p_k  # rarefaction
    return (2.0 * c_k / (g - 1.0)) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), \
        (1.0 / (rho_k * c_k)) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g=1.4, tol=1e-14):
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    # two-rarefaction guess
    z = (g - 1.0) / (2.0 * g)
    p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
    p = max(p, 1e-12)
    for _ in range(100):
        fl, dl = _fk(p, rl, pl, cl, g)
        fr, dr = _fk(p, rr, pr, cr, g)
        f = fl + fr + (ur - ul)
        p_new = max(p - f / (dl + dr), 1e-14)
        if abs(p_new - p) <= tol * 0.5 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(xi, rl, ul, pl, rr, ur, pr, g=1.4):
    """(rho, u, p) at similarity coordinate xi = (x - x0) / t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    if xi <= us:  # left of contact
        if ps > pl:  # left shock
            s = ul - cl * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
            if xi <= s:
                return rl, ul, pl
            return rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps
        csl = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
        if xi <= ul - cl:
            return rl, ul, pl
        if xi >= us - csl:
            return rl * (ps / pl) ** (1.0 / g), us, ps
        u = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + xi)
        c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - xi))
        return rl * (c / cl) ** (2.0 / (g - 1.0)), u, pl * (c / cl) ** (2.0 * g / (g - 1.0))
    if ps > pr:  # right shock
        s = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
        if xi >= s:
            return rr, ur, pr
        return rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps
    csr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
    if xi >= ur + cr:
        return rr, ur, pr
    if xi <= us + csr:
        return rr * (ps / pr) ** (1.0 / g), us, ps
    u = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + xi)
    c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - xi))
    return rr * (c / cr) ** (2.0 / (g - 1.0)), u, pr * (c / cr) ** (2.0 * g / (g - 1.0))


def profile(x, x0, t, left, right, g=1.4):
    out = np.array([sample((xx - x0) / t, *left, *right, g) for xx in x])
    return out[:, 0], out[:, 1], out[:, 2]
