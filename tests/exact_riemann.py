"""Exact solution of the 1D Euler Riemann problem (ideal gas), the textbook
method of Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics",
ch. 4: Newton iteration on f_L(p) + f_R(p) + du = 0, then sampling.
Used as an accuracy reference (not a parity oracle) for the Sod runs; pinned
against the star states the reference's own tests quote
(Tests/UnitDomainTest.H:99-361).
"""
import math

import numpy as np


def _f(p, rho, pk, g):
    c = math.sqrt(g * pk / rho)
    if p > pk:
        A = 2 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
    return (2 * c / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1),
            (p / pk) ** (-(g + 1) / (2 * g)) / (rho * c))


def star_state(left, right, g=1.4):
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, g)
        fr, dr = _f(p, rr, pr, g)
        pn = max(1e-14, p - (fl + fr + ur - ul) / (dl + dr))
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, g)
    fr, _ = _f(p, rr, pr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    def rho_star(rk, pk):
        if p > pk:
            r = (g - 1) / (g + 1)
            return rk * (p / pk + r) / (r * p / pk + 1)
        return rk * (p / pk) ** (1 / g)
    return p, u, rho_star(rl, pl), rho_star(rr, pr)


def sample(left, right, xi, g=1.4):
    """Primitive solution (rho, u, p) at similarity coordinate xi = x/t."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us, rsl, rsr = star_state(left, right, g)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    out = np.zeros((3,) + np.shape(xi))
    for idx, s in np.ndenumerate(np.asarray(xi, dtype=float)):
        if s <= us:  # left of contact
            if ps > pl:
                S = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                st = (rl, ul, pl) if s <= S else (rsl, us, ps)
            else:
                sh = ul - cl
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                st_ = us - cs
                if s <= sh:
                    st = (rl, ul, pl)
                elif s >= st_:
                    st = (rsl, us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    r = rl * (c / cl) ** (2 / (g - 1))
                    st = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                S = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                st = (rr, ur, pr) if s >= S else (rsr, us, ps)
            else:
                sh = ur + cr
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                st_ = us + cs
                if s >= sh:
                    st = (rr, ur, pr)
                elif s <= st_:
                    st = (rsr, us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    r = rr * (c / cr) ** (2 / (g - 1))
                    st = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
        for n in range(3):
            out[(n,) + idx] = st[n]
    return out
