"""Benchmark of the explicit finite-volume Euler step on one MI355X.

Workload (the reference path at scale): 2D supersonic compression-ramp
channel, domain 3 x 1 with a 15 degree ramp from x = 0.5 to 1.1, quad mesh
3840 x 1280 = 4,915,200 cells, interior nodes jittered +-10% of the local
spacing (seed 1), cells numbered in reverse Cuthill-McKee order; inflow
M 2.5, p 101325 Pa, T 300 K, gamma 1.4; HLLC flux, DM/MUSCL reconstruction
with the MinMod limiter, Green-Gauss gradients, 4-stage RK, local time step
(mode 2) at CFL 0.8.  A step is one solver iteration (one gradient pass and
four RK stages over every cell).  Inputs are resident in HBM before timing.

Prints one JSON line: cell-iterations/s (unit "cell-updates/s", one update
= one full iteration of a cell), the dominant kernel's achieved HBM
bandwidth against the 8 TB/s peak (HIP events on the solver's stream), and
the CPU oracle's serial throughput on a bounded sample of the same workload.
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s (% HBM roofline); time-to-1e-8 residual: NN warm vs plain"
HBM_PEAK_GBS = 8000.0
FLUX_IDS = {"ausm": 1, "hll": 2, "hllc": 3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx3", type=int, default=768)
    ap.add_argument("--ny3", type=int, default=256)
    ap.add_argument("--nz3", type=int, default=64)
    ap.add_argument("--tets", type=int, default=0, help="wedge3d: six tetrahedra per hexahedron")
    ap.add_argument("--nx", type=int, default=3840)
    ap.add_argument("--ny", type=int, default=1280)
    ap.add_argument("--flux", default="hllc", choices=sorted(FLUX_IDS))
    ap.add_argument("--precision", type=int, default=64, choices=(32, 64))
    ap.add_argument("--numbering", default="rcm", choices=("rcm", "random", "column"))
    ap.add_argument("--time-mode", type=int, default=2)
    ap.add_argument("--fused", type=int, default=-1, help="-1: library default")
    ap.add_argument("--cpu-iters", type=int, default=3,
                    help="iterations of the CPU oracle sample, about 14 s on one host core (0: skip)")
    ap.add_argument("--ref-iters", type=int, default=5,
                    help="iterations of the reference build's sample (a 4x coarser wedge, about 12 s; 0: skip)")
    ap.add_argument("--also-fp32", type=int, default=1, help="also time the FP32 mode")
    ap.add_argument("--timing-iters", type=int, default=-1,
                    help="-1: replay the timed region from its starting state with per-kernel HIP events (the "
                         "kernels' mean durations over the same iterations); N > 0: N more iterations after it; "
                         "0: no kernel timing")
    ap.add_argument("--profile", action="store_true", help="short run for rocprof (no CPU baseline, no JSON extras)")
    ap.add_argument("--tile-cells", type=int, default=0)
    ap.add_argument("--tile-threads", type=int, default=0)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="profile mode: first copy a buffer of this many MiB (a kernel of known traffic that "
                         "calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    # batched sweep workload
    ap.add_argument("--workload", default="wedge",
                    choices=("wedge", "wedge3d", "sweep", "warmstart", "refcase", "fp32field", "perturb"),
                    help="wedge: the 4.9M-cell iteration throughput; sweep: batched solves of many wedges to a "
                         "residual tolerance")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--sweep-tune", default="largest", choices=("median", "largest"),
                    help="sweep: the instance the autotuner times (the largest is bandwidth-bound like the batch; "
                         "the median one, about 0.1 ms per iteration, is launch-bound)")
    ap.add_argument("--tune-iters", type=int, default=30, help="sweep: autotuner trial iterations per tile shape")
    ap.add_argument("--sweep-seed", type=int, default=3)
    ap.add_argument("--cells-min", type=int, default=20000)
    ap.add_argument("--cells-max", type=int, default=400000)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--switch-tol", type=float, default=1e-3)
    ap.add_argument("--max-iter", type=int, default=60000, help="sweep: iteration budget cap per instance "
                    "(each instance gets min(this, 20 nx))")
    ap.add_argument("--sweep-precision", type=int, default=64, choices=(32, 64),
                    help="32: FP32 phase, then FP64 to the tolerance; 64: FP64 throughout")
    ap.add_argument("--compare-fp64", type=int, default=0, help="sweep: also solve the batch in FP64 throughout")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--budget-nx", type=int, default=20, help="sweep: iteration budget per instance, this x nx")
    ap.add_argument("--sweep-table", default="", help="sweep: write a per-instance CSV table to this path")
    # FP32 converged-field workload
    ap.add_argument("--field-config", default="c1", choices=("c1", "c2", "c3", "c3tet"),
                    help="fp32field: c1 120x40 wedge, c2 3840x1280 jittered wedge, c3 768x256x64 hexahedra, "
                         "c3tet the same split into tetrahedra")
    ap.add_argument("--field-max-iter", type=int, default=60000)
    ap.add_argument("--field-mesh", default="", help="fp32field, 2D configs: nx x ny instead of the config's mesh")
    ap.add_argument("--switch-tols", default="", help="fp32field: comma-separated FP32 -> FP64 switch drops "
                                                      "(default: --switch-tol)")
    # learned warm start workload
    ap.add_argument("--train-instances", type=int, default=256)
    ap.add_argument("--train-mesh", default="384x128", help="nx x ny of the training solves")
    ap.add_argument("--eval-instances", type=int, default=128)
    ap.add_argument("--eval-mesh", default="960x320", help="nx x ny of the held-out solves")
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--train-batch", type=int, default=16384)
    ap.add_argument("--budget-factor", type=int, default=20, help="iteration budget per solve: this x nx")
    return ap.parse_args()


def run_warmstart(a):
    """Learned warm start: training solves (seed 4) -> MLP training on the
    GPU -> held-out solves (seed 5) from the freestream and from the MLP's
    prediction, both to the same residual tolerance."""
    import numpy as np

    from supersonicflowsolver_amd.sweep import draw_instances
    from supersonicflowsolver_amd.warmstart import MLP, ab_test, converged_fields

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    tnx, tny = (int(v) for v in a.train_mesh.split("x"))
    enx, eny = (int(v) for v in a.eval_mesh.split("x"))
    tmp = tempfile.mkdtemp(prefix="psst_warm_")
    tr = draw_instances(a.train_instances, 4)
    X, Y, rt, wall_data = converged_fields(tr, tmp, tnx, tny, tol=a.tol, budget_factor=a.budget_factor, log=log)
    n_ok = int(np.sum(rt["outcome"] == 0))
    log("training data: %d points from %d/%d converged %dx%d solves, %.1f s" % (len(X), n_ok, len(tr), tnx, tny,
                                                                               wall_data))
    net = MLP(4, 256, 5, 4, seed=1)
    net.set_data(X, Y)
    t0 = time.perf_counter()
    hist = net.train(a.train_steps, batch=a.train_batch, lr=2e-3, lr_end=2e-5, seed=2)
    t_train = time.perf_counter() - t0
    st = net.train_stats()
    log("MLP training: %d steps in %.1f s (%.3f ms/step, hidden GEMMs %.1f TFLOP/s), loss %.5f -> %.6f" %
        (a.train_steps, t_train, st["ms_per_step"], st["gemm_flop_per_step"] / st["ms_per_step"] / 1e9,
         float(hist[:50].mean()), float(hist[-200:].mean())))
    ev = draw_instances(a.eval_instances, 5)
    r = ab_test(net, ev, tmp, enx, eny, tol=a.tol, budget_factor=a.budget_factor, log=log)
    pa, pb = r["plain"], r["warm"]
    ok = (pa["outcome"] == 0) & (pb["outcome"] == 0)
    cells = r["cells"]
    # The headline is the ratio of solver work (cell-iterations) to reach the tolerance, plain over warm, over
    # the instances that converge from both starts; batch wall times also include instances that never
    # converge (they run their budget or diverge) and are reported beside it.
    ratio = float(np.sum(cells[ok] * pa["iters64"][ok]) / max(1, np.sum(cells[ok] * pb["iters64"][ok])))
    out = {"metric": "time-to-%g residual: NN warm vs plain (held-out Mach/angle, instances converging in both)"
                     % a.tol,
           "value": round(ratio, 4), "unit": "x (plain cell-iterations / warm cell-iterations)", "n_gpus": 1,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 solver, bf16 MLP",
           "data": "synthetic",
           "config": {"workload": "learned warm start", "train": "%d solves at %dx%d (seed 4)" % (len(tr), tnx, tny),
                      "eval": "%d solves at %dx%d (seed 5)" % (len(ev), enx, eny),
                      "mlp": "4 -> 256 x 5 (SiLU) -> 4, BF16 MFMA, FP32 accumulation, Adam",
                      "train_steps": a.train_steps, "train_batch": a.train_batch,
                      "budget": "%d x nx iterations per solve" % a.budget_factor,
                      "scheme": "HLLC, DM/MinMod, Green-Gauss, RK4, time step mode 2, CFL 0.8, FP64, to a %g drop" % a.tol},
           "plain_wall_s": round(pa["wall"], 3), "warm_wall_s": round(pb["wall"], 3),
           "both_converged": int(ok.sum()),
           "outcomes_plain": {k: int(np.sum(pa["outcome"] == v)) for k, v in (("converged", 0), ("budget", 1),
                                                                              ("diverged", 2))},
           "outcomes_warm": {k: int(np.sum(pb["outcome"] == v)) for k, v in (("converged", 0), ("budget", 1),
                                                                            ("diverged", 2))},
           "iters_plain": {"median": int(np.median(pa["iters64"][ok])), "sum": int(pa["iters64"][ok].sum())},
           "iters_warm": {"median": int(np.median(pb["iters64"][ok])), "sum": int(pb["iters64"][ok].sum())},
           "wall_ratio_all_instances": round(pa["wall"] / pb["wall"], 4),
           "surrogate_rms_err": {"median": round(float(np.median(r["surrogate_rel_l2"][ok])), 5),
                                 "max": round(float(np.max(r["surrogate_rel_l2"][ok])), 5)},
           "warm_restarted_from_freestream": int(np.sum(pb["restarted"])),
           "converged_fields_differ_max": (float(np.nanmax(r["converged_field_diff"][ok])) if ok.any() else None),
           "predict_s": round(r["predict_s"], 3), "train_s": round(t_train, 2),
           "train_ms_per_step": round(st["ms_per_step"], 4),
           "train_gemm_tflops": round(st["gemm_flop_per_step"] / st["ms_per_step"] / 1e9, 2),
           "train_data_s": round(wall_data, 2)}
    print(json.dumps(out), flush=True)


def run_sweep(a):
    """Batched sweep: `instances` wedges (cells log-uniform, Mach and ramp
    angle uniform, seeded), all solved to a `tol` residual drop in one batch
    (one workgroup per active tile, hipGraph replay, autotuned tiles)."""
    import numpy as np

    import supersonicflowsolver_amd as ss
    from supersonicflowsolver_amd.sweep import (build_cases, draw_instances, reflection_regime, solve_to_tolerance,
                                                theta_max)

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    inst = draw_instances(a.instances, a.sweep_seed, cells=(a.cells_min, a.cells_max))
    tmp = tempfile.mkdtemp(prefix="psst_sweep_")
    t0 = time.perf_counter()
    cases = build_cases(inst, tmp, cfl=0.8, time_mode=2, rk_levels=4, flux=3, order=2, grad=1, limiter=2)
    N = np.array([c.n_cells for c in cases], dtype=np.int64)
    log("%d instances, %d cells (min %d, max %d), host setup %.1f s" %
        (len(cases), N.sum(), N.min(), N.max(), time.perf_counter() - t0))
    if a.tile_cells and a.tile_threads:
        best = (a.tile_cells, a.tile_threads)
        log("tile fixed at %s (--tile-cells, --tile-threads)" % (best,))
    else:
        rep = int(np.argmax(N)) if a.sweep_tune == "largest" else int(np.argsort(N)[len(N) // 2])
        t0 = time.perf_counter()
        best, table = ss.autotune(cases[rep], trial_iters=a.tune_iters, precision=a.sweep_precision)
        log("autotune on instance %d (%d cells): best tile %s, %.1f s; %s" %
            (rep, N[rep], best, time.perf_counter() - t0, [(t[0], round(t[1], 4)) for t in table]))

    def solve(precision):
        t0 = time.perf_counter()
        b = ss.Batch(cases, precision=precision, tile_cells=best[0], tile_threads=best[1])
        b.set_graph(a.graph)
        log("batch context (precision %d): %s, create %.1f s" % (precision, b.info(), time.perf_counter() - t0))
        budget = [min(a.max_iter, a.budget_nx * d["nx"]) for d in inst]
        r = solve_to_tolerance(b, tol=a.tol, switch_tol=a.switch_tol, budget=budget, log=log)
        b.close()
        return r

    r = solve(a.sweep_precision)
    it_tot = r["iters32"] + r["iters64"]
    conv = int(np.sum(r["outcome"] == 0))
    cell_its = float(np.sum(N * it_tot))
    out = {"metric": "batched sweep: wall time until every instance reaches a %g residual drop" % a.tol,
           "value": round(r["wall"], 3), "unit": "s", "n_gpus": 1, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64" if a.sweep_precision == 32 else "f64", "data": "synthetic",
           "config": {"workload": "sweep of %d compression-ramp wedges" % len(cases), "instances": len(cases),
                      "cells_total": int(N.sum()), "cells_range": [a.cells_min, a.cells_max],
                      "mach_range": [1.5, 4.0], "ramp_deg_range": [5.0, 25.0], "seed": a.sweep_seed,
                      "scheme": "HLLC, DM/MinMod, Green-Gauss, RK4, local time step 2, CFL 0.8",
                      "tile": list(best), "graph": bool(a.graph)},
           "converged": conv, "diverged": int(np.sum(r["outcome"] == 2)),
           "budget_spent": int(np.sum(r["outcome"] == 1)), "cell_updates_per_s": round(cell_its / r["wall"], 1),
           "wall_fp32_phase_s": round(r["wall32"], 3), "wall_fp64_phase_s": round(r["wall64"], 3),
           "iters_fp32": {"median": int(np.median(r["iters32"])), "max": int(r["iters32"].max())},
           "iters_fp64": {"median": int(np.median(r["iters64"])), "max": int(r["iters64"].max())}}
    # outcome against how each instance's ramp shock reflects from the top
    # wall (regular reflection, Mach reflection, or a detached ramp shock)
    tmax = np.array([theta_max(d["mach"]) for d in inst])
    regime = np.array([reflection_regime(d["mach"], d["theta"]) for d in inst])
    out["outcomes_by_shock"] = {
        kind: {"instances": int(sel.sum()), "converged": int(np.sum(sel & (r["outcome"] == 0))),
               "budget_spent": int(np.sum(sel & (r["outcome"] == 1))), "diverged": int(np.sum(sel & (r["outcome"] == 2)))}
        for kind, sel in (("regular reflection", regime == "regular"), ("Mach reflection", regime == "mach"),
                          ("detached ramp shock", regime == "detached"))}
    out["config"]["budget"] = "%d x nx iterations" % a.budget_nx
    if a.sweep_table:
        with open(a.sweep_table, "w") as f:
            f.write("instance,cells,nx,ny,mach,theta_deg,theta_max_deg,regime,outcome,iterations,final_drop\n")
            for i, d in enumerate(inst):
                f.write("%d,%d,%d,%d,%.4f,%.3f,%.3f,%s,%s,%d,%.3e\n" % (
                    i, N[i], d["nx"], d["ny"], d["mach"], d["theta"], tmax[i], regime[i],
                    ("converged", "budget", "diverged")[int(r["outcome"][i])], it_tot[i], r["drop"][i]))
    if a.compare_fp64 and a.sweep_precision == 32:
        r64 = solve(64)
        out["fp64_only"] = {"wall_s": round(r64["wall"], 3), "converged": int(np.sum(r64["outcome"] == 0)),
                            "iters": {"median": int(np.median(r64["iters64"])), "max": int(r64["iters64"].max())}}
    print(json.dumps(out), flush=True)


def run_fp32field(a):
    """Converged-field error of the FP32 mode: the case solved in FP64 to a
    `tol` residual drop; in FP32 for the same iterations (an FP32 state
    floors the drop near 1e-4); and in two phases, FP32 to a `switch_tol`
    drop and then FP64 to `tol` (both phases normalised by the FP64 run's
    first residual row).  Reports the relative L2 error of every field
    against the converged FP64 field (velocity components against the norm
    of the speed), the FP32 floor, and the iterations and wall time of the
    FP64 and the two-phase solves.  The FP64 run is bitwise the oracle's."""
    import numpy as np

    import supersonicflowsolver_amd as ss

    def log(msg):
        print("bench: " + msg, file=sys.stderr, flush=True)

    tmp = tempfile.mkdtemp(prefix="psst_field_")
    prefix = os.path.join(tmp, "case")
    cfgs = {"c1": ("2D wedge 120x40 quads, 15 degree ramp, column numbering; M 2.5, HLLC + DM/MinMod + "
                   "Green-Gauss, RK4, time step mode 2, CFL 0.8", dict(nx=120, ny=40)),
            "c2": ("2D wedge 3840x1280 quads, jitter 10%% seed 1, RCM numbering; M 2.5, HLLC + DM/MinMod + "
                   "Green-Gauss, RK4, time step mode 2, CFL 0.8", dict(nx=3840, ny=1280, jitter=0.1, jitter_seed=1,
                                                                   numbering=1)),
            "c3": ("3D wedge 768x256x64 hexahedra, 20 degree ramp, z-periodic; M 3, AUSM + DM/Venkatakrishnan + "
                   "least squares, RK4, time step mode 2, CFL 0.8", dict(nx=768, ny=256, nz=64, depth=1.0,
                                                                         ramp_deg=20.0, periodic_z=1)),
            "c3tet": ("3D wedge 768x256x64 hexahedra split into 6 tetrahedra each, 20 degree ramp, z-periodic; M 3, "
                      "AUSM + DM/Venkatakrishnan + least squares, RK4, time step mode 2, CFL 0.8",
                      dict(nx=768, ny=256, nz=64, depth=1.0, ramp_deg=20.0, periodic_z=1, tets=1))}
    desc, wp = cfgs[a.field_config]
    dim = 3 if "nz" in wp else 2
    if a.field_mesh and dim == 2:
        nx, ny = (int(v) for v in a.field_mesh.split("x"))
        desc = desc.replace("%dx%d" % (wp["nx"], wp["ny"]), "%dx%d" % (nx, ny))
        wp = dict(wp, nx=nx, ny=ny)
    if dim == 3:
        ss.write_case_files(prefix, dim=3, cfl=0.8, time_mode=2, iterations=1000, rk_levels=4, flux=FLUX_IDS["ausm"],
                            order=2, grad=2, limiter=4, mach=3.0, p=101325.0, T=300.0)
    else:
        ss.write_case_files(prefix, cfl=0.8, time_mode=2, iterations=1000, rk_levels=4, flux=FLUX_IDS["hllc"],
                            order=2, grad=1, limiter=2, mach=2.5, p=101325.0, T=300.0)
    case = ss.Case.wedge(ss.wedge_params(**wp), prefix)
    N = case.n_cells
    log("%s: %d cells" % (a.field_config, N))

    def drops(rows, ref):
        """max_j res_j / ref_j per row, leaving out columns whose reference is
        round-off against the row's largest (as the library's tolerance test)."""
        keep = ref > 1e-12 * ref.max()
        return (rows[:, keep] / ref[keep]).max(axis=1)

    def run_chunked(s, n, tol, what):
        """s.run in chunks of at most 2000 iterations, logging each, until n
        iterations ran or the tolerance stopped a chunk; returns the rows."""
        rows, done = [], 0
        while done < n:
            k = min(2000, n - done)
            r = s.run(k, tol=tol, allow_no_convergence=True)
            rows.append(r[:s.iters_done])
            done += s.iters_done
            log("%s: %d iterations, drop %.2e" % (what, done, float(drops(rows[-1][-1:], rows[0][0])[0])))
            if s.iters_done < k:
                break
        return np.concatenate(rows), done

    def solve(precision, n, tol, res0=None):
        s = ss.Solver(case, precision=precision)
        if res0 is not None:
            s.set_residual_reference(res0)
        t = time.perf_counter()
        r, _ = run_chunked(s, n, tol, "FP%d" % precision)
        el = time.perf_counter() - t
        return s, r, el

    s64, r64, t64 = solve(64, a.field_max_iter, a.tol)
    it64 = len(r64)
    ref = s64.state()
    res0 = r64[0].copy()
    s64.close()
    drop64 = float(drops(r64[-1:], res0)[0])
    log("FP64: %d iterations to a %.2e drop, %.1f s" % (it64, drop64, t64))
    s32, r32, t32 = solve(32, it64, 0.0, res0)
    st32 = s32.state()
    s32.close()
    d32 = drops(r32, res0)
    log("FP32: %d iterations, %.1f s, lowest drop %.2e" % (it64, t32, d32.min()))
    # two phases: FP32 to a switch drop, then FP64 to tol
    two = []
    for sw in ([float(x) for x in a.switch_tols.split(",")] if a.switch_tols else [a.switch_tol]):
        st = ss.Solver(case, precision=32)
        st.set_residual_reference(res0)
        t = time.perf_counter()
        _, n1 = run_chunked(st, it64, sw, "two phases, FP32 to %.0e" % sw)
        st.set_precision(64)
        _, n2 = run_chunked(st, a.field_max_iter, a.tol, "two phases, FP64")
        t_two = time.perf_counter() - t
        two.append((sw, n1, n2, t_two, st.state()))
        st.close()
        log("two phases, switch at %.0e: %d FP32 + %d FP64 iterations, %.1f s" % (sw, n1, n2, t_two))

    def first_below(d, x):
        return int(np.argmax(d <= x)) + 1 if (d <= x).any() else None

    d64 = drops(r64, res0)
    levels = (1e-1, 1e-2, 1e-3, 1e-4, 1e-6, 1e-8)
    names = ["u", "v", "w", "p", "T", "rho"] if dim == 3 else ["u", "v", "p", "T", "rho"]

    def rel(st_):
        speed = np.linalg.norm(ref[:dim, :N])
        out = {}
        for k, nm in enumerate(names):
            scale = speed if k < dim else np.linalg.norm(ref[k, :N])
            out[nm] = float(np.linalg.norm(st_[k, :N] - ref[k, :N]) / scale)
        return out

    out = {"metric": "FP32 converged-field error against FP64 (relative L2 per field)",
           "value": max(rel(st32).values()), "unit": "relative L2 (max over fields)", "n_gpus": 1,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (geometry, norms f64)",
           "data": "synthetic (generated wedge mesh, freestream initial state)",
           "config": {"workload": desc, "cells": N, "tol": a.tol, "switch_tol": a.switch_tol},
           "fp64": {"iterations": it64, "final_drop": drop64, "wall_s": round(t64, 3),
                    "first_iteration_below": {"%g" % x: first_below(d64, x) for x in levels}},
           "fp32": {"iterations": it64, "wall_s": round(t32, 3), "lowest_drop": float(d32.min()),
                    "first_iteration_below": {"%g" % x: first_below(d32, x) for x in levels},
                    "rel_l2_vs_converged_fp64": rel(st32)},
           "two_phase": [{"switch_drop": sw, "fp32_iterations": n1, "fp64_iterations": n2, "wall_s": round(tt, 3),
                          "rel_l2_vs_converged_fp64": rel(stt), "speedup_vs_fp64": round(t64 / tt, 4)}
                         for sw, n1, n2, tt, stt in two]}
    shutil.rmtree(tmp, ignore_errors=True)
    print(json.dumps(out), flush=True)


def run_perturb(a):
    """Sensitivity of the last decades of an FP64 solve to the state: the 2D
    wedge (HLLC, DM/MinMod, Green-Gauss, RK4, time step mode 2, CFL 0.8;
    --field-mesh, default 960x320) is solved to a 1e-3 drop, then continued
    to 1e-8 from that state as it is, and from copies whose u, v, p, T are
    multiplied by 1 + eps N(0, 1) (rho from p / (rm T)), both normalised by
    the solve's first residual row.  Reports the iterations each needs."""
    import numpy as np

    import supersonicflowsolver_amd as ss

    nx, ny = (int(v) for v in (a.field_mesh or "960x320").split("x"))
    tmp = tempfile.mkdtemp(prefix="psst_perturb_")
    prefix = os.path.join(tmp, "case")
    ss.write_case_files(prefix, cfl=0.8, time_mode=2, iterations=10, rk_levels=4, flux=3, order=2, grad=1, limiter=2)
    case = ss.Case.wedge(ss.wedge_params(nx=nx, ny=ny, jitter=0.1, jitter_seed=1, numbering=1), prefix)
    N = case.n_cells
    with ss.Solver(case) as s:
        r = s.run(a.field_max_iter, tol=1e-3, allow_no_convergence=True)
        k = s.iters_done
        res0 = r[0].copy()
        st = s.state()
        s.run(a.field_max_iter, tol=a.tol, allow_no_convergence=True)
        base = s.iters_done
    print("bench: %dx%d: a 1e-3 drop at iteration %d, then %d more to %g" % (nx, ny, k, base, a.tol), file=sys.stderr,
          flush=True)
    rng = np.random.default_rng(0)
    rm = case.settings().rm
    out = {}
    for eps in (1e-12, 1e-9, 1e-6, 1e-3):
        p = st.copy()
        for row in range(4):
            p[row, :N] *= 1.0 + eps * rng.standard_normal(N)
        p[4, :N] = p[2, :N] / (rm * p[3, :N])
        with ss.Solver(case) as s:
            s.set_state(p)
            s.set_residual_reference(res0)
            s.run(a.field_max_iter, tol=a.tol, allow_no_convergence=True)
            out["%g" % eps] = s.iters_done
        print("bench: eps %.0e: %d iterations" % (eps, out["%g" % eps]), file=sys.stderr, flush=True)
    shutil.rmtree(tmp, ignore_errors=True)
    print(json.dumps({"metric": "iterations from a 1e-3 drop to %g, unperturbed and perturbed" % a.tol,
                      "value": base, "unit": "iterations", "n_gpus": 1, "higher_is_better": False,
                      "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "2D wedge %dx%d, HLLC + DM/MinMod + Green-Gauss, RK4, time step mode 2, "
                                             "CFL 0.8" % (nx, ny), "cells": N},
                      "iteration_of_1e-3_drop": k, "perturbed": out}), flush=True)


# The reference's only published timing: the elapsed-time column of its
# committed residual file, row 3000 (wedge_residual.dat:3000), on its
# author's PC (CPU not stated; VS2017 Release build).
REFCASE_PUBLISHED_S = 152.856


def run_refcase(a):
    """The reference's own case: its wedge80_240.msh (19,200 quads) with
    wedge.inp/.bvp/.btq/.par/.prop (HLL, first order, RK1, time step mode 1 in
    the reference's order, CFL 2.08, 3000 iterations, a save every 1000), kept
    under tests/golden/ref_wedge80x240.  psst_case_solve runs it as the
    reference's main() does and writes the same residual file, whose row-3000
    elapsed time (from the start of the solve, saves at 1000 and 2000
    included, in whole milliseconds) is compared with the published one."""
    import hashlib
    import lzma

    import supersonicflowsolver_amd as ss
    src = os.path.join(ROOT, "tests", "golden", "ref_wedge80x240")
    d = tempfile.mkdtemp(prefix="psst_refcase_")
    for ext in ("inp", "bvp", "btq", "par", "prop"):
        shutil.copy(os.path.join(src, "wedge." + ext), d)
    with lzma.open(os.path.join(src, "wedge80_240.msh.xz")) as f, open(os.path.join(d, "wedge80_240.msh"), "wb") as g:
        g.write(f.read())
    prefix = os.path.join(d, "wedge")
    case = ss.Case.open(os.path.join(d, "wedge80_240.msh"), prefix)
    print("bench: the reference's own case, %d cells, iterations %d-%d" % (case.n_cells, case.first_iter,
                                                                           case.last_iter), file=sys.stderr,
          flush=True)
    t = time.perf_counter()
    case.solve()
    wall = time.perf_counter() - t
    with open(prefix + "_residual.dat") as f:
        rows = [l.split() for l in f if l.strip()]
    t3000 = float(rows[2999][1])
    with lzma.open(os.path.join(src, "wedge.sav.xz")) as f:
        ref_sav = hashlib.sha256(f.read().replace(b"\r", b"")).hexdigest()
    with open(prefix + ".sav", "rb") as f:
        same_sav = hashlib.sha256(f.read()).hexdigest() == ref_sav
    shutil.rmtree(d, ignore_errors=True)
    out = {"metric": "wall time of the reference's own 3000-iteration case (elapsed-time column, row 3000)",
           "value": round(t3000, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": round(t3000 / REFCASE_PUBLISHED_S, 5), "dtype": "f64",
           "data": "the reference's own mesh and settings files",
           "config": {"workload": "wedge80_240.msh (19,200 quads), HLL, first order, RK1, time step mode 1 in the "
                                  "reference's order, CFL 2.08, 3000 iterations, a save every 1000"},
           "baseline": {"value": REFCASE_PUBLISHED_S, "unit": "s", "hardware": "the reference author's PC (CPU not "
                        "stated), VS2017 Release build", "source": "wedge_residual.dat:3000, column 2"},
           "whole_call_s": round(wall, 3), "sav_identical_to_reference": bool(same_sav)}
    print(json.dumps(out), flush=True)


def alg_bytes(kind, N, G, F, nnz, dim, b, last_stage=False):
    """Algorithmic HBM bytes of one launch, the scope's unique-touch model:
    b = state bytes, geometry 8, index 4, nv = dim + 2 stored state values
    (u_i, p, T).  Stage: N (nv b [q read] + nv dim b [gradient read] + nv b
    [w0 read] + nv b [q write] + (1 + dim) 8 [V, x]) + F ((dim + 1) 8 [n, A] +
    2 * 4 [c0, c1]).  Gradient pass: N (nv b + nv dim b + 8) + F ((dim + 1) 8
    + 8 + 8 [beta]).  Arrays a kernel reads beyond these (tile face lists,
    halo cells, time step face sums) count as traffic, not as algorithmic
    bytes."""
    nv = dim + 2  # u..., p, T
    if kind == "stage":
        return N * (nv * b + nv * dim * b + nv * b + nv * b + (1 + dim) * 8) + F * ((dim + 1) * 8 + 2 * 4)
    if kind == "grad":
        return N * (nv * b + nv * dim * b + 8) + F * ((dim + 1) * 8 + 8 + 8)
    if kind == "flux":  # unfused face kernel: the stage's face part plus the flux write
        return N * (nv * b + nv * dim * b + dim * 8) + G * (nv + 1) * b + F * ((dim + 1) * 8 + 2 * 4) + F * nv * b
    if kind == "update":  # unfused cell kernel: the stage's cell part plus the flux read
        return F * nv * b + N * (nv * b + nv * b + 8) + N * (nv + 1) * b + (N * nv * b if last_stage else 0)
    if kind == "bc":
        return G * (nv + 1) * b * 2 + G * 12
    raise ValueError(kind)


def reference_cpu_sample(ss, a, numbering):
    """The reference's own serial solver (oracle/_ref/psst_ref, g++ -O2 from
    its sources) on a bounded sample of the 2D workload: the same wedge,
    scheme and numbering at a quarter of the cells per direction, a.ref_iters
    iterations.  Per-iteration wall time comes from the reference's residual
    file, whose second column is the solve's elapsed steady_clock time in ms
    resolution (psstFiniteVolumeGrid.cpp:278, :371-383): (t_last - t_first) /
    (iterations - 1), so mesh reading, set-up and the final save are not
    timed.  Returns a cpu_baseline object, or {"error": ...}."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "psst_ref")
    if not os.path.exists(exe):
        return {"error": "oracle/_ref/psst_ref is not built"}
    nx, ny, n_it = a.nx // 4, a.ny // 4, a.ref_iters
    d = tempfile.mkdtemp(prefix="psst_ref_")
    ss.write_wedge_msh(ss.wedge_params(nx=nx, ny=ny, ramp_deg=15.0, jitter=0.1, jitter_seed=1, numbering=numbering,
                                       perm_seed=2), os.path.join(d, "mesh.msh"))
    ss.write_case_files(os.path.join(d, "case"), cfl=0.8, time_mode=a.time_mode, iterations=n_it,
                        save_step=10 ** 9, rk_levels=4, flux=FLUX_IDS[a.flux], order=2, grad=1, limiter=2, mach=2.5,
                        p=101325.0, T=300.0)
    with open(os.path.join(d, "problem_manager"), "w") as f:
        f.write("mesh.msh\ncase\n")
    print("bench: reference build, %d-cell sample, %d iterations" % (nx * ny, n_it), file=sys.stderr, flush=True)
    t = time.perf_counter()
    try:
        subprocess.run([exe], cwd=d, stdin=subprocess.DEVNULL, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=300, check=True)
        with open(os.path.join(d, "case_residual.dat")) as f:
            ts = [float(l.split()[1]) for l in f if l.strip()][:n_it]
    except Exception as e:  # reported, not fatal: the port's figure stands in
        return {"error": "reference build run failed: %s" % e}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    wall = time.perf_counter() - t
    per_it = (ts[-1] - ts[0]) / (len(ts) - 1)
    print("bench: reference build done (%.1f s wall, %.3f s per iteration)" % (wall, per_it), file=sys.stderr,
          flush=True)
    try:
        cpu_model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        cpu_model = "unknown"
    return {"value": round(nx * ny / per_it, 1), "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": "the reference's own solver (its sources, g++ -O2) on the same wedge, scheme and numbering at "
                      "%dx%d (%d cells, 1/16 of the workload), iterations 2-%d of %d timed by its residual file's "
                      "clock (mesh reading, set-up and output not timed); one host core (%s)"
                      % (nx, ny, nx * ny, len(ts), n_it, cpu_model)}


def model_bytes_per_cell_iteration(dim, b, N, F, stages=4):
    """Unique-touch model per cell-iteration (second order, RK): per stage
    q read + grad read + w0 read + q write + V, x; faces n, A, c0, c1; plus
    the gradient pass and the w0 snapshot."""
    nv = dim + 2
    stage = N * (nv * b + nv * dim * b + nv * b + nv * b + (1 + dim) * 8) + F * ((dim + 1) * 8 + 2 * 4)
    grad = N * (nv * b + nv * dim * b + 8) + F * ((dim + 1) * 8 + 8 + 8)
    snap = N * nv * b
    return (stages * stage + grad + snap) / N


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path's multi-GPU sharding is a design note); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.workload == "sweep":
        return run_sweep(a)
    if a.workload == "warmstart":
        return run_warmstart(a)
    if a.workload == "refcase":
        return run_refcase(a)
    if a.workload == "fp32field":
        return run_fp32field(a)
    if a.workload == "perturb":
        return run_perturb(a)
    import numpy as np

    import supersonicflowsolver_amd as ss

    numbering = {"column": 0, "rcm": 1, "random": 2}[a.numbering]
    tmp = tempfile.mkdtemp(prefix="psst_bench_")
    prefix = os.path.join(tmp, "case")
    three_d = a.workload == "wedge3d"
    dim = 3 if three_d else 2
    if three_d:
        # 20 degree ramp extruded along z with periodic side faces; M 3, AUSM, Venkatakrishnan, least squares
        ss.write_case_files(prefix, dim=3, wall_names=("symm", "wedge", "wall", "up"), cfl=0.8,
                            time_mode=a.time_mode, iterations=a.steps, rk_levels=4, flux=FLUX_IDS["ausm"], order=2,
                            grad=2, limiter=4, mach=3.0, p=101325.0, T=300.0)
        params = ss.wedge_params(nx=a.nx3, ny=a.ny3, nz=a.nz3, depth=1.0, ramp_deg=20.0, jitter=0.0,
                                 numbering=numbering, tets=a.tets, periodic_z=1)
    else:
        ss.write_case_files(prefix, cfl=0.8, time_mode=a.time_mode, iterations=a.steps, rk_levels=4,
                            flux=FLUX_IDS[a.flux], order=2, grad=1, limiter=2, mach=2.5, p=101325.0, T=300.0)
        params = ss.wedge_params(nx=a.nx, ny=a.ny, ramp_deg=15.0, jitter=0.1, jitter_seed=1, numbering=numbering,
                                 perm_seed=2)
    t0 = time.perf_counter()
    case = ss.Case.wedge(params, prefix)
    t_setup = time.perf_counter() - t0
    N, G, F = case.n_cells, case.n_ghosts, case.n_faces
    nnz = int(case.mesh().visit_offsets[N])
    print("bench: %s mesh, %d cells, %d faces, host setup %.1f s" % ("3D" if three_d else "2D", N, F, t_setup),
          file=sys.stderr, flush=True)

    # the reference itself (compiled from its own sources into oracle/_ref),
    # run as a child process before this process touches the GPU
    ref_cpu = None
    if a.ref_iters > 1 and not three_d and not a.profile:
        ref_cpu = reference_cpu_sample(ss, a, numbering)

    if a.profile and a.calib_copy_mb > 0:
        import ctypes
        ms = ctypes.c_double(0)
        rc = ss._lib.lib().psst_copy_calibration(a.calib_copy_mb << 20, 3, ctypes.byref(ms))
        if rc != 0:
            raise SystemExit("bench: calibration copy failed")
        print("bench: calibration copy (k_copy) of %d MiB: %.3f ms, %.0f GB/s" %
              (a.calib_copy_mb, ms.value, 2 * (a.calib_copy_mb << 20) / ms.value / 1e6), file=sys.stderr, flush=True)

    def timed(precision):
        kw = dict(precision=precision)
        if a.fused >= 0:
            kw["fused"] = a.fused
        if a.tile_cells:
            kw["tile_cells"] = a.tile_cells
        if a.tile_threads:
            kw["tile_threads"] = a.tile_threads
        t = time.perf_counter()
        s = ss.Solver(case, **kw)
        print("bench: FP%d context ready (%.1f s)" % (precision, time.perf_counter() - t), file=sys.stderr,
              flush=True)
        s.run(a.warmup)
        start = s.state() if a.timing_iters < 0 else None
        t = time.perf_counter()
        s.run(a.steps)
        el = time.perf_counter() - t
        print("bench: FP%d timed run done (%.3f s)" % (precision, el), file=sys.stderr, flush=True)
        # per-kernel mean durations (HIP events on the solver's stream).  The
        # kernels' cost follows the flow (limiter and wave-speed branches
        # diverge more once the shock has formed), so by default the timed
        # region is replayed from its starting state with the events on: the
        # same iterations as `value`, without the events' cost in `value`.
        stats, el_replay = {}, None
        if a.timing_iters != 0:
            if start is not None:
                s.set_state(start)
            s.kernel_timing(True)
            t = time.perf_counter()
            s.run(a.steps if start is not None else a.timing_iters)
            el_replay = time.perf_counter() - t
            for name in s.kernel_names():
                cnt, ms = s.kernel_stats(name)
                if cnt:
                    stats[name] = (cnt, ms)
            s.kernel_timing(False)
        state = s.state()
        s.close()
        return el, stats, state, el_replay

    print("bench: FP%d, %d warmup + %d timed iterations" % (a.precision, a.warmup, a.steps), file=sys.stderr,
          flush=True)
    el, stats, state, el_replay = timed(a.precision)
    if a.profile:
        print("bench: done (%.3f s)" % el, file=sys.stderr, flush=True)
        return
    if not np.isfinite(state).all():
        raise SystemExit("bench: non-finite state")
    b = 8 if a.precision == 64 else 4
    cell_its = N * a.steps / el
    # dominant kernel by total time per iteration
    timing_iters = a.steps if a.timing_iters < 0 else a.timing_iters
    per_iter = {k: cnt * ms / timing_iters for k, (cnt, ms) in stats.items()}
    dom = max(per_iter, key=per_iter.get)
    kind = {"flux": "flux", "update": "update", "grad_gg": "grad", "grad_lsq": "grad", "bc": "bc",
            "stage": "stage", "stage_last": "stage", "update_levels_block": "update",
            "update_levels_wg": "update", "update_level": "update",
            "update_chunks": "update"}.get(dom, dom)
    nbytes = alg_bytes(kind, N, G, F, nnz, dim, b, last_stage=(dom == "stage_last"))
    mean_ms = stats[dom][1]
    note = None
    if dom in ("update_levels_block", "update_levels_wg", "update_level", "update_chunks"):
        # time step mode 1 in the reference's order: each stage's update is one sweep over the cells'
        # dependency levels (one launch, or one per level); rate over a whole stage's sweep
        mean_ms = stats[dom][0] * stats[dom][1] / (timing_iters * 4)
        note = "reference-ordered update sweep (a chain of dependency levels): latency-bound, not HBM-bound"
    achieved = nbytes / (mean_ms * 1e-3) / 1e9
    # HBM traffic of the dominant kernel: rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs (each scaled by
    # the counters' reading for a copy kernel of known size), recorded under profiles/ for the default 2D
    # and 3D workloads in FP64; null for other configurations
    traffic = None
    if three_d:
        tpath = os.path.join(ROOT, "profiles", "traffic_wedge3d_768x256x64.json")
        recorded = (a.nx3, a.ny3, a.nz3, a.tets) == (768, 256, 64, 0)
    else:
        tpath = os.path.join(ROOT, "profiles", "traffic_wedge3840x1280.json")
        recorded = (a.nx, a.ny, a.flux) == (3840, 1280, "hllc")
    if (os.path.exists(tpath) and recorded and a.precision == 64 and a.numbering == "rcm"
            and a.time_mode == 2 and a.tile_cells == 0 and a.tile_threads == 0):
        with open(tpath) as f:
            tr = json.load(f).get(dom)
        if tr:
            traffic = tr["bytes_per_launch"]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "alg_bytes_per_launch": int(nbytes), "mean_ms": round(mean_ms, 4),
                "timed_over": ("a replay of the timed region (same start state, same %d iterations) with HIP events "
                               "around every launch; %.4f ms per step with the events" %
                               (a.steps, el_replay / a.steps * 1e3)) if a.timing_iters < 0 and el_replay else
                              ("%d iterations after the timed region" % timing_iters)}
    if note:
        roofline["note"] = note
    whole = model_bytes_per_cell_iteration(dim, b, N, F)
    out = {
        "metric": METRIC, "value": round(cell_its, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(el / a.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if a.precision == 64 else "f32",
        "data": "synthetic (generated wedge mesh, freestream initial state)",
        "config": {"workload": ("3D wedge %dx%dx%d %s (%d cells), 20 degree ramp, z-periodic, %s numbering; M 3, "
                                "AUSM + DM/Venkatakrishnan + least squares, RK4, time step mode %d, CFL 0.8" %
                                (a.nx3, a.ny3, a.nz3, "tetrahedra" if a.tets else "hexahedra", N, a.numbering,
                                 a.time_mode)) if three_d else
                               ("2D wedge %dx%d quads (%d cells), jitter 10%% seed 1, %s numbering; %s + DM/MinMod + "
                                "Green-Gauss, RK4, time step mode %d, CFL 0.8" % (a.nx, a.ny, N, a.numbering,
                                                                               a.flux.upper(), a.time_mode)),
                   "cells": N, "faces": F, "rk_stages": 4, "precision": a.precision},
        "roofline": roofline,
        "cell_stage_updates_per_s": round(cell_its * 4, 1),
        "whole_step_model": {"bytes_per_cell_iteration": round(whole, 1),
                             "achieved_GBs": round(whole * cell_its / 1e9, 1),
                             "frac": round(whole * cell_its / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v[1], 4) for k, v in stats.items()},
    }
    if a.also_fp32 and a.precision == 64:
        el32, stats32, state32, _ = timed(32)
        ci32 = N * a.steps / el32
        whole32 = model_bytes_per_cell_iteration(dim, 4, N, F)
        # relative L2 per field; the velocity components against the norm of
        # the speed (a component that is zero in FP64, such as w on the
        # extruded wedge, has no scale of its own)
        err = {}
        speed = np.linalg.norm(state[:dim, :N])
        for k, nm in enumerate(["u", "v", "w", "p", "T", "rho"] if three_d else ["u", "v", "p", "T", "rho"]):
            scale = speed if k < dim else np.linalg.norm(state[k, :N])
            err[nm] = float(np.linalg.norm(state32[k, :N] - state[k, :N]) / scale)
        out["fp32"] = {"value": round(ci32, 1), "ms_per_step": round(el32 / a.steps * 1e3, 4),
                       "whole_step_frac": round(whole32 * ci32 / 1e9 / HBM_PEAK_GBS, 4),
                       "kernels_ms": {k: round(v[1], 4) for k, v in stats32.items()},
                       "rel_l2_vs_fp64_after_steps": err}
    if a.cpu_iters > 0 and not three_d:
        from oracle.oracle import Oracle
        raw = case.raw()
        raw["dim"] = 2
        t = time.perf_counter()
        o = Oracle(None, prefix, raw=raw)
        t_open = time.perf_counter() - t
        t = time.perf_counter()
        o.run(a.cpu_iters)
        el_cpu = time.perf_counter() - t
        try:
            cpu_model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
        except Exception:
            cpu_model = "unknown"
        out["cpu_port"] = {"value": round(N * a.cpu_iters / el_cpu, 1), "unit": "cell-updates/s", "cores": 1,
                           "kind": "port",
                           "sample": "%d iterations of the same %d-cell FP64 workload, serial oracle on one "
                                     "host core (%s); setup %.1f s not timed" % (a.cpu_iters, N, cpu_model, t_open)}
    if ref_cpu and "value" in ref_cpu:
        out["cpu_baseline"] = ref_cpu
    elif "cpu_port" in out:
        out["cpu_baseline"] = out["cpu_port"]
        if ref_cpu:
            out["cpu_reference_error"] = ref_cpu["error"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
