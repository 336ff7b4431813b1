"""Exact solutions used as physics oracles.

sod_exact: exact solution of the 1D Riemann problem for an ideal gas (Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", ch. 4: Newton iteration for the star pressure, then sampling
of the self-similar solution).  vortex_exact: Shu's isentropic vortex translated by the freestream.
"""
import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        f = (p - pk) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
    else:  # rarefaction
        f = 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)
        df = 1 / (rho * ck) * (p / pk) ** (-(g + 1) / (2 * g))
    return f, df


def star_state(WL, WR, g=1.4):
    rl, ul, pl = WL
    rr, ur, pr = WR
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-12)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sod_exact(x, t, x0=0.5, WL=(1.0, 0.0, 1.0), WR=(0.125, 0.0, 0.1), g=1.4):
    """(rho, u, p) at positions x and time t > 0."""
    ps, us = star_state(WL, WR, g)
    rl, ul, pl = WL
    rr, ur, pr = WR
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((len(x), 3))
    for i, xi in enumerate(np.asarray(x)):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    out[i] = (rl, ul, pl)
                else:
                    r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                    out[i] = (r, us, ps)
            else:  # left rarefaction
                shl = ul - cl
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - csl
                if s <= shl:
                    out[i] = (rl, ul, pl)
                elif s > stl:
                    out[i] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[i] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:  # right of contact
            if ps > pr:  # right shock
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    out[i] = (rr, ur, pr)
                else:
                    r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                    out[i] = (r, us, ps)
            else:
                shr = ur + cr
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + csr
                if s >= shr:
                    out[i] = (rr, ur, pr)
                elif s <= str_:
                    out[i] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[i] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out


def vortex_exact(ccen, t=0.0, beta=5.0, mach=0.5, gamma=1.4, L=20.0):
    """Shu's isentropic vortex at time t on the periodic box [-L/2, L/2]^2 (conserved variables)."""
    uinf = mach * math.sqrt(gamma)
    x = ccen[:, 0] - uinf * t
    x = (x + L / 2) % L - L / 2
    y = ccen[:, 1]
    r2 = x * x + y * y
    e = np.exp(0.5 * (1 - r2))
    du = -beta / (2 * math.pi) * y * e
    dv = beta / (2 * math.pi) * x * e
    T = 1 - (gamma - 1) * beta ** 2 / (8 * gamma * math.pi ** 2) * np.exp(1 - r2)
    rho = T ** (1 / (gamma - 1))
    p = rho * T
    u, v = uinf + du, dv
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)], axis=1)
