"""Exact solution of the 1-D Riemann problem of the Euler equations for an ideal gas.

Test infrastructure: the shock-tube tests compare the oracle and the HIP path with it.  The
algorithm is the textbook one (E. F. Toro, Riemann Solvers and Numerical Methods for Fluid
Dynamics, 3rd ed., chapter 4): the star-region pressure solves f_L(p) + f_R(p) + u_R - u_L = 0
by Newton iteration from the two-rarefaction guess, with f_K the shock (Rankine-Hugoniot) or
rarefaction (isentropic) branch; the solution at x/t is then sampled wave by wave.
"""
import math

import numpy as np


def _fk(p, rk, pk, ck, g):
    """f_K(p) and its derivative (Toro eqs. 4.6-4.7, 4.37)."""
    if p > pk:   # shock
        ak = 2.0 / ((g + 1.0) * rk)
        bk = (g - 1.0) / (g + 1.0) * pk
        s = math.sqrt(ak / (p + bk))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (bk + p))
    pr = p / pk   # rarefaction
    f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    return f, 1.0 / (rk * ck) * pr ** (-(g + 1.0) / (2.0 * g))


def star_state(left, right, gamma=1.4, tol=1e-15):
    """(p*, u*) for primitive states (rho, u, p) left and right."""
    rl, ul, pl = left
    rr, ur, pr = right
    g = gamma
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    if 2.0 * (cl + cr) / (g - 1.0) <= ur - ul:
        raise ValueError("vacuum is generated")
    z = (g - 1.0) / (2.0 * g)
    p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
    for _ in range(100):
        fl, dl = _fk(p, rl, pl, cl, g)
        fr, dr = _fk(p, rr, pr, cr, g)
        pn = max(p - (fl + fr + ur - ul) / (dl + dr), 1e-14)
        if abs(pn - p) <= tol * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(left, right, s, gamma=1.4):
    """Primitive (rho, u, p) at the similarity coordinates s = (x - x0) / t (array)."""
    rl, ul, pl = left
    rr, ur, pr = right
    g = gamma
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    ps, us = star_state(left, right, g)
    s = np.asarray(s, dtype=np.float64)
    out = np.empty(s.shape + (3,))
    gm = (g - 1.0) / (g + 1.0)
    for idx, x in np.ndenumerate(s):
        if x <= us:   # left of the contact
            if ps > pl:
                sl = ul - cl * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                w = (rl, ul, pl) if x <= sl else (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:
                shl = ul - cl
                cs = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                stl = us - cs
                if x <= shl:
                    w = (rl, ul, pl)
                elif x >= stl:
                    w = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - x))
                    u = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + x)
                    r = rl * (c / cl) ** (2.0 / (g - 1.0))
                    w = (r, u, pl * (c / cl) ** (2.0 * g / (g - 1.0)))
        else:         # right of the contact
            if ps > pr:
                sr = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                w = (rr, ur, pr) if x >= sr else (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:
                shr = ur + cr
                cs = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                str_ = us + cs
                if x >= shr:
                    w = (rr, ur, pr)
                elif x <= str_:
                    w = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - x))
                    u = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + x)
                    r = rr * (c / cr) ** (2.0 / (g - 1.0))
                    w = (r, u, pr * (c / cr) ** (2.0 * g / (g - 1.0)))
        out[idx] = w
    return out
