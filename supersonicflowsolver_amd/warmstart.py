"""Learned warm start of the compression-ramp solves (``include/psst_mlp.h``).

A coordinate MLP maps (x, y, Mach, ramp angle) to (rho, u, v, p).  It is
trained on converged fields that the batched solver produces, and a new
solve starts from its prediction instead of the freestream.  The plain
solver still runs to the same residual tolerance, so it stays the arbiter of
the surrogate's error; the residual drop is normalised by the freestream
start's first residual, so the warm start cannot lower its own denominator.
"""
import ctypes as C
import time

import numpy as np

from . import Batch, PsstError, _track
from ._lib import lib
from .sweep import build_cases, draw_instances, solve_to_tolerance

P_INF, T_INF = 101325.0, 300.0


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """The network on the GPU: BF16 MFMA forward/backward kernels, Adam on
    FP32 master weights (``psst_mlp_*``)."""

    def __init__(self, n_in=4, width=256, depth=5, n_out=4, seed=0):
        h = C.c_void_p()
        st = lib().psst_mlp_create(n_in, width, depth, n_out, C.c_uint64(seed), C.byref(h))
        if st != 0:
            raise PsstError(st, lib().psst_mlp_last_error(None).decode())
        self._h = h
        _track(self)
        self.dims = (n_in, width, depth, n_out)
        self.n_params = int(lib().psst_mlp_param_count(h))

    def _check(self, st):
        if st != 0:
            raise PsstError(st, lib().psst_mlp_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().psst_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib().psst_mlp_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.n_params,)
        self._check(lib().psst_mlp_set_params(self._h, _fp(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.dims[3]), dtype=np.float32)
        self._check(lib().psst_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double(0)
        self._check(lib().psst_mlp_loss_grad(self._h, _fp(x), _fp(y), x.shape[0], C.byref(loss), _fp(g)))
        return float(loss.value), g

    def set_data(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        self._check(lib().psst_mlp_set_data(self._h, _fp(x), _fp(y), x.shape[0]))

    def train(self, steps, batch=16384, lr=2e-3, lr_end=2e-5, seed=0, history=True):
        hist = np.zeros(steps, dtype=np.float32) if history else None
        self._check(lib().psst_mlp_train(self._h, int(steps), int(batch), float(lr), float(lr_end), C.c_uint64(seed),
                                         _fp(hist) if history else None))
        return hist

    def train_stats(self):
        ms = C.c_double(0)
        fl = C.c_double(0)
        self._check(lib().psst_mlp_train_stats(self._h, C.byref(ms), C.byref(fl)))
        return {"ms_per_step": ms.value, "gemm_flop_per_step": fl.value}


# ---------------------------------------------------------------------------
# Features, targets, initial states
# ---------------------------------------------------------------------------
def freestream(case, mach):
    s = case.settings()
    rho = P_INF / (s.rm * T_INF)
    u = mach * np.sqrt(s.gamma * s.rm * T_INF)
    return rho, u, P_INF


def centres(case):
    """Interior cell centroids [N, 2] (canonical order)."""
    m = case.mesh()
    NT = case.n_cells + case.n_ghosts
    c = np.ctypeslib.as_array(m.cell_centre, shape=(case.dim * NT,)).reshape(case.dim, NT)
    return c[:2, :case.n_cells].T.copy()


def features(xy, mach, theta):
    """Network inputs scaled to about [-1, 1]: the channel is 3 x 1, Mach in
    [1.5, 4], ramp angle in [5, 25] degrees."""
    n = xy.shape[0]
    x = np.empty((n, 4), dtype=np.float32)
    x[:, 0] = (xy[:, 0] - 1.5) / 1.5
    x[:, 1] = (xy[:, 1] - 0.5) / 0.5
    x[:, 2] = (mach - 2.75) / 1.25
    x[:, 3] = (theta - 15.0) / 10.0
    return x


def targets(case, prim, mach):
    """(rho, u, v, p) of the interior cells, relative to the freestream."""
    rho_i, u_i, p_i = freestream(case, mach)
    N = case.n_cells
    y = np.empty((N, 4), dtype=np.float32)
    y[:, 0] = prim[4, :N] / rho_i - 1.0
    y[:, 1] = prim[0, :N] / u_i - 1.0
    y[:, 2] = prim[1, :N] / u_i
    y[:, 3] = prim[2, :N] / p_i - 1.0
    return y


def warm_state(case, mlp, mach, theta, ramp_start=0.5):
    """Initial primitive state from the network, projected onto states the
    steady flow can take:

    * ghosts are the case's own initial state (the supersonic inlet's ghosts
      hold the exact inflow; the others are rewritten by the boundary
      conditions after the first stage);
    * cells upstream of the ramp (x < ramp_start) are exactly the
      freestream: with a supersonic inflow nothing reaches them;
    * density and pressure stay within [0.3, 40] and [0.3, 60] times the
      freestream (two oblique shocks and the expansion at the ramp's end stay
      inside these for Mach 1.5-4 and ramps up to 25 degrees), the speed at
      most 1.05 times the inflow speed, and the temperature follows from
      p / (rho rm), so it is positive."""
    prim = case.initial_state()
    s = case.settings()
    rho_i, u_i, p_i = freestream(case, mach)
    xy = centres(case)
    y = mlp.forward(features(xy, mach, theta)).astype(np.float64)
    N = case.n_cells
    rho = np.clip((y[:, 0] + 1.0) * rho_i, 0.3 * rho_i, 40.0 * rho_i)
    p = np.clip((y[:, 3] + 1.0) * p_i, 0.3 * p_i, 60.0 * p_i)
    u = (y[:, 1] + 1.0) * u_i
    v = y[:, 2] * u_i
    speed = np.hypot(u, v)
    scale = np.minimum(1.0, 1.05 * u_i / np.maximum(speed, 1e-300))
    u, v = u * scale, v * scale
    up = xy[:, 0] < ramp_start
    rho[up], p[up], u[up], v[up] = rho_i, p_i, u_i, 0.0
    prim[0, :N] = u
    prim[1, :N] = v
    prim[2, :N] = p
    prim[3, :N] = p / (rho * s.rm)
    prim[4, :N] = rho
    # the freestream cells are the case's own values, bit for bit
    prim[:, :N][:, up] = case.initial_state()[:, :N][:, up]
    return prim


# ---------------------------------------------------------------------------
# Data generation, training, A/B evaluation
# ---------------------------------------------------------------------------
def converged_fields(instances, workdir, nx, ny, tol=1e-8, budget_factor=20, tile=None, log=None):
    """Solves every instance (FP64 batch) to `tol`; returns the training set
    (X, Y) of the converged ones, and per-instance outcomes."""
    inst = [dict(d, nx=nx, ny=ny) for d in instances]
    cases = build_cases(inst, workdir, prefix="t", cfl=0.8, time_mode=2, rk_levels=4, flux=3, order=2, grad=1,
                        limiter=2)
    kw = {} if tile is None else {"tile_cells": tile[0], "tile_threads": tile[1]}
    t0 = time.perf_counter()
    b = Batch(cases, precision=64, **kw)
    r = solve_to_tolerance(b, tol=tol, budget=[budget_factor * nx] * len(cases), log=log)
    xs, ys = [], []
    for i, c in enumerate(cases):
        if r["outcome"][i] != 0:
            continue
        prim = b.state(i)
        xs.append(features(centres(c), inst[i]["mach"], inst[i]["theta"]))
        ys.append(targets(c, prim, inst[i]["mach"]))
    b.close()
    wall = time.perf_counter() - t0
    return np.concatenate(xs), np.concatenate(ys), r, wall


def ab_test(mlp, instances, workdir, nx, ny, tol=1e-8, budget_factor=20, tile=None, log=None):
    """Time to a `tol` residual drop from the freestream (A) and from the
    network's prediction (B), both batched, for the same instances; the drop
    of both is normalised by the freestream start's first residual.  A warm
    start that goes non-finite or diverges is reset to the freestream in the
    same batch (the plain solver is the fallback) and the iterations it had
    spent count against the warm arm.  Returns both arms' outcomes, the
    surrogate's error against the plain arm's converged field, and the
    difference between the two arms' converged fields."""
    inst = [dict(d, nx=nx, ny=ny) for d in instances]
    cases = build_cases(inst, workdir, prefix="e", cfl=0.8, time_mode=2, rk_levels=4, flux=3, order=2, grad=1,
                        limiter=2)
    kw = {} if tile is None else {"tile_cells": tile[0], "tile_threads": tile[1]}
    b = Batch(cases, precision=64, **kw)
    n = len(cases)
    first = b.run(1, history=True, allow_failures=True)
    res0 = first["hist"][0].copy()
    fresh = [c.initial_state() for c in cases]
    budget = [budget_factor * nx] * n
    for i in range(n):
        b.set_state(i, fresh[i])
        b.set_residual_reference(i, res0[i])
    if log:
        log("plain (freestream) starts of %d instances" % n)
    ra = solve_to_tolerance(b, tol=tol, budget=budget, log=log)
    conv_a = [b.state(i) for i in range(n)]
    t0 = time.perf_counter()
    warm = [warm_state(cases[i], mlp, inst[i]["mach"], inst[i]["theta"]) for i in range(n)]
    t_pred = time.perf_counter() - t0
    for i in range(n):
        b.set_state(i, warm[i])
        b.set_residual_reference(i, res0[i])
    if log:
        log("warm starts (network prediction, %.2f s)" % t_pred)
    rb = solve_to_tolerance(b, tol=tol, budget=budget, log=log, fallback={i: fresh[i] for i in range(n)})
    conv_b = [b.state(i) for i in range(n)]
    b.close()
    # surrogate error against the converged field: RMS over the interior
    # cells of (u, v, p, rho) differences, each scaled by its freestream
    # magnitude (v by the inflow speed); the same measure between the two
    # arms' converged fields
    err = np.zeros(n)
    field_diff = np.full(n, np.nan)
    for i in range(n):
        N = cases[i].n_cells
        rho_i, u_i, p_i = freestream(cases[i], inst[i]["mach"])
        scale = np.array([u_i, u_i, p_i, rho_i])[:, None]
        d = (warm[i][[0, 1, 2, 4], :N] - conv_a[i][[0, 1, 2, 4], :N]) / scale
        err[i] = np.sqrt(np.mean(d * d))
        if ra["outcome"][i] == 0 and rb["outcome"][i] == 0:
            d = (conv_b[i][[0, 1, 2, 4], :N] - conv_a[i][[0, 1, 2, 4], :N]) / scale
            field_diff[i] = np.sqrt(np.mean(d * d))
    return {"plain": ra, "warm": rb, "surrogate_rel_l2": err, "converged_field_diff": field_diff, "predict_s": t_pred,
            "cells": np.array([c.n_cells for c in cases])}
