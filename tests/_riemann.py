"""Exact solutions used by the physics checks (test infrastructure).

* `riemann_exact`: the exact solution of the 1-D Riemann problem for a
  polytropic gas, by the pressure-function method of E. F. Toro, "Riemann
  Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4:
  Newton iteration on f(p) = f_L(p) + f_R(p) + (u_R - u_L) = 0 for the star
  pressure, then sampling of the self-similar solution at x/t.
* `oblique_shock`: the weak oblique-shock solution of the theta-beta-M
  relation for a wedge of half-angle theta, with the Rankine-Hugoniot jumps
  across it.
"""
import math

import numpy as np


def _f_and_df(p, rho, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        s = math.sqrt(a / (p + b))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (b + p))
    # rarefaction
    r = p / pk
    return (2.0 * ck / (g - 1.0)) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), \
        (1.0 / (rho * ck)) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(left, right, g):
    """(p*, u*) of the Riemann problem with left/right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    p = max(1e-12 * min(pl, pr), 0.5 * (pl + pr))  # arithmetic-mean start (Toro 4.3.2)
    for _ in range(100):
        fl, dl = _f_and_df(p, rl, pl, cl, g)
        fr, dr = _f_and_df(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(p - dp, 1e-12 * p)
        if abs(pn - p) <= 1e-14 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f_and_df(p, rl, pl, cl, g)
    fr, _ = _f_and_df(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def riemann_exact(left, right, g, xi):
    """Density, velocity and pressure of the exact solution at speeds xi = x/t
    (array), left/right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    ps, us = star_state(left, right, g)
    gm = (g - 1.0) / (g + 1.0)
    out = np.zeros((3, len(xi)))
    for i, s in enumerate(np.asarray(xi, dtype=float)):
        if s <= us:  # left of the contact
            if ps > pl:  # left shock
                sl = ul - cl * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                if s <= sl:
                    out[:, i] = (rl, ul, pl)
                else:
                    out[:, i] = (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:  # left rarefaction
                shl = ul - cl
                csl = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                stl = us - csl
                if s <= shl:
                    out[:, i] = (rl, ul, pl)
                elif s >= stl:
                    out[:, i] = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - s))
                    u = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + s)
                    out[:, i] = (rl * (c / cl) ** (2.0 / (g - 1.0)), u, pl * (c / cl) ** (2.0 * g / (g - 1.0)))
        else:  # right of the contact
            if ps > pr:  # right shock
                sr = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                if s >= sr:
                    out[:, i] = (rr, ur, pr)
                else:
                    out[:, i] = (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:  # right rarefaction
                shr = ur + cr
                csr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                stt = us + csr
                if s >= shr:
                    out[:, i] = (rr, ur, pr)
                elif s <= stt:
                    out[:, i] = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - s))
                    u = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + s)
                    out[:, i] = (rr * (c / cr) ** (2.0 / (g - 1.0)), u, pr * (c / cr) ** (2.0 * g / (g - 1.0)))
    return out


def oblique_shock(mach, theta_deg, g):
    """Weak-shock angle beta (degrees) and the jumps p2/p1, rho2/rho1 behind
    an attached oblique shock on a wedge of angle theta (theta-beta-M relation,
    solved by bisection between the Mach angle and the detachment angle)."""
    th = math.radians(theta_deg)

    def theta_of(b):
        m2 = (mach * math.sin(b)) ** 2
        return math.atan(2.0 / math.tan(b) * (m2 - 1.0) / (mach * mach * (g + math.cos(2.0 * b)) + 2.0))

    mu = math.asin(1.0 / mach)
    bs = np.linspace(mu + 1e-9, 0.5 * math.pi - 1e-9, 20001)
    bmax = bs[int(np.argmax([theta_of(b) for b in bs]))]
    lo, hi = mu + 1e-12, bmax
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if theta_of(mid) < th:
            lo = mid
        else:
            hi = mid
    b = 0.5 * (lo + hi)
    m2 = (mach * math.sin(b)) ** 2
    p21 = 1.0 + 2.0 * g / (g + 1.0) * (m2 - 1.0)
    r21 = (g + 1.0) * m2 / ((g - 1.0) * m2 + 2.0)
    return math.degrees(b), p21, r21
