"""Set-ups of the physics checks (shared by the CPU and GPU tests): a Sod
shock tube in a flat channel and the supersonic compression wedge."""
import math
import os

import numpy as np

import supersonicflowsolver_amd as ss
from _riemann import oblique_shock, riemann_exact, star_state

# Sod's ratios (density 8, pressure 10) at engineering values: left 1e5 Pa /
# 300 K, right 1e4 Pa / 240 K, gas at rest, diaphragm at x = 1.5 in a 3 x 1
# channel of 600 x 2 cells with slip walls.  The left end is the fixed-state
# inlet (u = 0), the right end an extrapolation outlet; neither is reached
# by a wave in 350 iterations.
SOD_NX, SOD_NY, SOD_X0 = 600, 2, 1.5
SOD_LEFT, SOD_RIGHT = (1.0e5, 300.0), (1.0e4, 240.0)  # (p, T)
# uniform time step CFL * sqrt(V_min) / c_inlet (psstUniformTime::initialize)
# = 0.02 * 0.05 / 347 m/s = 2.88e-6 s, a Courant number of about 0.37 on the
# fastest wave; 350 iterations reach t = 1.0e-3 s
SOD_CFL, SOD_ITERS = 0.02, 350


def sod_case(tmp, flux=3):
    p = ss.wedge_params(nx=SOD_NX, ny=SOD_NY, ramp_deg=0.0)
    mesh = os.path.join(tmp, "tube.msh")
    ss.write_wedge_msh(p, mesh)
    prefix = os.path.join(tmp, "tube")
    ss.write_case_files(prefix, cfl=SOD_CFL, time_mode=0, iterations=SOD_ITERS, rk_levels=4, flux=flux, order=2,
                        grad=1, limiter=2, mach=0.0, p=SOD_LEFT[0], T=SOD_LEFT[1])
    case = ss.Case.open(mesh, prefix)
    s = case.settings()
    N, G = case.n_cells, case.n_ghosts
    x = np.ctypeslib.as_array(case.mesh().cell_centre, shape=(3 * (N + G),)).reshape(3, N + G)[0]
    prim = np.zeros((5, N + G))
    left = x < SOD_X0
    prim[2] = np.where(left, SOD_LEFT[0], SOD_RIGHT[0])
    prim[3] = np.where(left, SOD_LEFT[1], SOD_RIGHT[1])
    prim[4] = prim[2] / (s.rm * prim[3])
    return mesh, prefix, case, prim, x[:N]


def sod_errors(case, x, q):
    """L1 errors (density, pressure, velocity; normalised by the left state's
    density, pressure and sound speed, averaged over the tube) against the
    exact solution, the mean pressure between the contact and the shock
    against p*, and the largest transverse velocity."""
    s = case.settings()
    g, rm = s.gamma, s.rm
    N = case.n_cells
    t = SOD_ITERS * s.uniform_dt
    rl = SOD_LEFT[0] / (rm * SOD_LEFT[1])
    rr = SOD_RIGHT[0] / (rm * SOD_RIGHT[1])
    left, right = (rl, 0.0, SOD_LEFT[0]), (rr, 0.0, SOD_RIGHT[0])
    ex = riemann_exact(left, right, g, (x - SOD_X0) / t)
    cl = math.sqrt(g * SOD_LEFT[0] / rl)
    ps, us = star_state(left, right, g)
    xc = SOD_X0 + us * t
    window = (x > xc + 0.06) & (x < xc + 0.15)  # between the smeared contact and the shock
    return {
        "rho": float(np.abs(q[4, :N] - ex[0]).mean() / rl),
        "p": float(np.abs(q[2, :N] - ex[2]).mean() / SOD_LEFT[0]),
        "u": float(np.abs(q[0, :N] - ex[1]).mean() / cl),
        "p_star": float(q[2, :N][window].mean() / ps - 1.0),
        "v_max": float(np.abs(q[1, :N]).max() / cl),
    }


# The default generated wedge: 120 x 40 cells, M 2.5, 15 degree ramp from
# x = 0.5 to 1.1, HLLC + DM/MinMod + Green-Gauss + RK4, time step mode 2.
WEDGE_ITERS = 3000


def wedge_case(tmp):
    p = ss.wedge_params(nx=120, ny=40)
    mesh = os.path.join(tmp, "w.msh")
    ss.write_wedge_msh(p, mesh)
    prefix = os.path.join(tmp, "w")
    ss.write_case_files(prefix, cfl=0.8, time_mode=2, iterations=WEDGE_ITERS, rk_levels=4, flux=3, order=2, grad=1,
                        limiter=2)
    return mesh, prefix, ss.Case.open(mesh, prefix)


def wedge_plateau(case, q):
    """Pressure and density ratios and the flow angle (degrees) averaged over
    the cells between the ramp and the oblique shock (x in [0.75, 1.05], at
    least 0.04 above the ramp and 0.06 below the exact shock), the exact
    values, and the shock's x at y = 0.1, 0.2, 0.3 (first cell past half the
    pressure jump) against the exact shock line."""
    g = case.settings().gamma
    N = case.n_cells
    c = np.ctypeslib.as_array(case.mesh().cell_centre, shape=(3 * (N + case.n_ghosts),)).reshape(3, -1)
    x, y = c[0, :N], c[1, :N]
    beta, p21, r21 = oblique_shock(2.5, 15.0, g)
    tb, tt = math.tan(math.radians(beta)), math.tan(math.radians(15.0))
    sel = (x > 0.75) & (x < 1.05) & (y > (x - 0.5) * tt + 0.04) & (y < (x - 0.5) * tb - 0.06)
    p_inf = q[2, :N][x < 0.3].mean()
    r_inf = q[4, :N][x < 0.3].mean()
    shock = []
    for yl in (0.1, 0.2, 0.3):
        row = np.abs(y - yl) < 0.5 / 40
        xs = np.sort(x[row][q[2, :N][row] > 0.5 * (1.0 + p21) * p_inf])
        shock.append((float(xs[0]), 0.5 + yl / tb))
    return {
        "cells": int(sel.sum()),
        "p21": float(q[2, :N][sel].mean() / p_inf), "p21_exact": p21,
        "r21": float(q[4, :N][sel].mean() / r_inf), "r21_exact": r21,
        "angle": float(np.degrees(np.arctan2(q[1, :N][sel], q[0, :N][sel])).mean()),
        "shock_x": shock,
    }
