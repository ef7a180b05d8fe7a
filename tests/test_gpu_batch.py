"""Batched solves (psst_batch_*): every instance of a batch must be bitwise
identical, in FP64, to the same case run alone through psst_ctx (which is
itself bitwise identical to the oracle): same residual rows, same stopping
iteration, same final state.  Covers the launch of one workgroup per active
tile (and its capacity shrinking with the active list), the per-instance
freeze at the tolerance, hipGraph replay against eager launches,
odd stage counts (alternating state buffers), the FP32 -> FP64 switch and a
non-finite instance.  Needs an MI355X."""
import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from oracle.oracle import Oracle
from supersonicflowsolver_amd.sweep import draw_instances, make_case

pytestmark = pytest.mark.gpu

SIZES = [(36, 12, 2.5, 15.0), (60, 20, 3.1, 9.0), (45, 15, 1.8, 21.0), (24, 8, 3.8, 6.0)]


def _cases(tmp_path, sizes=SIZES, **kw):
    out = []
    for k, (nx, ny, m, th) in enumerate(sizes):
        out.append(make_case({"nx": nx, "ny": ny, "mach": m, "theta": th}, str(tmp_path), "i%d" % k,
                             jitter=0.1, jitter_seed=k + 1, numbering=2 if k % 2 else 1, **kw))
    return out


def _alone(case, n, tol=0.0, **over):
    with ss.Solver(case, **over) as s:
        res = s.run(n, tol=tol, allow_no_convergence=True)
        return res, s.iters_done, s.state()


def _compare(b, cases, out, n, tol=0.0, **over):
    for i, c in enumerate(cases):
        res, done, st = _alone(c, n, tol, **over)
        print("instance %d: %d cells, alone %d iterations, batch %d" % (i, c.n_cells, done, out["iters"][i]),
              flush=True)
        assert out["iters"][i] == done
        assert np.array_equal(out["hist"][:done, i, :], res[:done])
        assert not out["hist"][done:, i, :].any()
        assert np.array_equal(b.state(i), st)


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_fixed_iterations_bitwise(tmp_path, graph):
    cases = _cases(tmp_path)
    with ss.Batch(cases, tile_cells=128, tile_threads=256) as b:
        b.set_graph(graph)
        out = b.run(40, history=True)
        _compare(b, cases, out, 40)


def test_batch_tolerance_freeze_bitwise(tmp_path):
    """Each instance stops at its own first iteration below the tolerance,
    exactly where a lone run stops; the others keep going."""
    cases = _cases(tmp_path)
    tol = 2e-2
    with ss.Batch(cases) as b:
        out = b.run(400, tol=tol, history=True, allow_no_convergence=True)
        assert len(set(out["iters"].tolist())) > 1, out["iters"]
        _compare(b, cases, out, 400, tol)


def test_batch_launch_capacity_follows_the_active_list(tmp_path):
    """32-cell tiles: about 110 tiles, of which the 60x20 instance holds a
    third.  Instances that stop within the first 256-iteration chunk shrink the
    active list, so later chunks launch a smaller grid from graphs captured for
    that capacity; every instance stays bitwise equal to its lone run."""
    sizes = [(36, 12, 3.8, 6.0), (60, 20, 1.8, 21.0), (45, 15, 3.5, 8.0), (24, 8, 3.8, 6.0), (30, 10, 3.1, 9.0)]
    cases = _cases(tmp_path, sizes)
    tol = 2e-2
    with ss.Batch(cases, tile_cells=32, tile_threads=256) as b:
        out = b.run(900, tol=tol, history=True, allow_no_convergence=True)
        print("iterations", out["iters"].tolist(), "tiles", b.info()["tiles"], flush=True)
        assert min(out["iters"]) < 256 < max(out["iters"]), out["iters"]
        _compare(b, cases, out, 900, tol, tile_cells=32, tile_threads=256)


def test_batch_odd_stage_count_and_continuation(tmp_path):
    """RK3 (odd: the state buffers alternate between iterations) with freezes,
    then a second run call continuing from the frozen states."""
    cases = _cases(tmp_path, rk_levels=3)
    with ss.Batch(cases) as b:
        o1 = b.run(300, tol=5e-2, history=True, allow_no_convergence=True)
        o2 = b.run(25, history=True)
        for i, c in enumerate(cases):
            with ss.Solver(c) as s:
                r1 = s.run(300, tol=5e-2, allow_no_convergence=True)
                d1 = s.iters_done
                assert d1 == o1["iters"][i]
                assert np.array_equal(o1["hist"][:d1, i], r1[:d1])
                r2 = s.run(25)
                assert np.array_equal(o2["hist"][:, i], r2)
                assert np.array_equal(b.state(i), s.state())


def test_batch_excluded_instance_keeps_its_state_with_odd_stage_count(tmp_path):
    """RK3 alternates the state buffers every iteration.  An instance
    excluded after an odd number of iterations keeps the state it had (not
    the other buffer's stage intermediate), and included again it continues
    from there: equal to a lone context run for the same iterations."""
    cases = _cases(tmp_path, rk_levels=3)[:2]
    with ss.Batch(cases) as b:
        b.run(1)
        b.set_enabled(0, False)
        b.run(1)
        with ss.Solver(cases[0]) as s0, ss.Solver(cases[1]) as s1:
            s0.run(1)
            s1.run(2)
            assert np.array_equal(b.state(0), s0.state())
            assert np.array_equal(b.state(1), s1.state())
            b.set_enabled(0, True)
            out = b.run(3, history=True)
            r0, r1 = s0.run(3), s1.run(3)
            assert np.array_equal(out["hist"][:, 0], r0) and np.array_equal(out["hist"][:, 1], r1)
            assert np.array_equal(b.state(0), s0.state())
            assert np.array_equal(b.state(1), s1.state())


def test_batch_precision_switch(tmp_path):
    """FP32 iterations, switch to FP64, more iterations: equal to a lone
    context doing the same switch."""
    cases = _cases(tmp_path)[:3]
    with ss.Batch(cases, precision=32) as b:
        b.run(30)
        b.set_precision(64)
        out = b.run(20, history=True)
        for i, c in enumerate(cases):
            with ss.Solver(c, precision=32) as s:
                s.run(30)
                s.set_precision(64)
                r = s.run(20)
                assert np.array_equal(out["hist"][:, i], r)
                assert np.array_equal(b.state(i), s.state())


def test_batch_nonfinite_instance_is_isolated(tmp_path):
    """A negative temperature gives NaN in one instance; it is frozen with
    PSST_NAN_DETECTED and the other instances are unaffected."""
    cases = _cases(tmp_path)[:3]
    with ss.Batch(cases) as b:
        bad = cases[1].initial_state()
        bad[3, : cases[1].n_cells // 2] = -300.0
        b.set_state(1, bad)
        with pytest.raises(ss.PsstError) as ei:
            b.run(10)
        assert ei.value.status == 3
        for i in (0, 2):
            res, done, st = _alone(cases[i], 10)
            assert np.array_equal(b.state(i), st)


def test_batch_sweep_draw_and_autotune(tmp_path):
    """A small seeded sweep (uneven sizes, Mach, angle) runs batched; the
    autotuner times every candidate tile shape and returns the fastest."""
    inst = draw_instances(6, seed=3, cells=(2000, 12000))
    cases = [make_case(d, str(tmp_path), "s%d" % k) for k, d in enumerate(inst)]
    best, table = ss.autotune(cases[-1], trial_iters=3)
    ran = [t for t in table if t[1] > 0]
    assert ran and best in [t[0] for t in ran]
    assert min(t[1] for t in ran) == dict(table)[best]
    with ss.Batch(cases, tile_cells=best[0], tile_threads=best[1]) as b:
        out = b.run(2000, tol=1e-3, allow_no_convergence=True)
        print("iterations per instance", out["iters"].tolist(), "drops", out["drop"].tolist(), flush=True)
        assert (out["status"] == 0).all()
        assert (out["drop"] <= 1e-3).all()


def test_run_with_tolerance_ends_on_the_stopping_iteration(tmp_path):
    """psst_run with a tolerance leaves the state of exactly the iteration
    that met it (equal to a plain run of that many iterations)."""
    case = _cases(tmp_path)[0]
    with ss.Solver(case) as s1, ss.Solver(case) as s2:
        r1 = s1.run(400, tol=2e-2)
        done = s1.iters_done
        assert 0 < done < 400 and not r1[done:].any()
        r2 = s2.run(done)
        assert np.array_equal(r1[:done], r2)
        assert np.array_equal(s1.state(), s2.state())


def _oracle(case, prefix):
    raw = case.raw()
    raw["dim"] = case.dim
    return Oracle(None, prefix, raw=raw)


def test_batch_fp64_instances_match_the_oracle(tmp_path):
    """Each FP64 instance of a batch is bitwise the CPU oracle's solve of
    the same case (residual rows and final state), for uneven meshes in RCM
    and random numbering."""
    cases = _cases(tmp_path)[:3]
    with ss.Batch(cases) as b:
        out = b.run(60, history=True)
        for i, c in enumerate(cases):
            o = _oracle(c, str(tmp_path / ("i%d" % i)))
            res = o.run(60)
            print("instance %d: %d cells, rows equal %s, state equal %s" %
                  (i, c.n_cells, np.array_equal(out["hist"][:, i], res), np.array_equal(b.state(i), o.state())),
                  flush=True)
            assert np.array_equal(out["hist"][:, i], res)
            assert np.array_equal(b.state(i), o.state())


def test_batch_fp32_converged_field_against_the_fp64_oracle(tmp_path):
    """An FP32 instance of a batch, run well past its residual floor, against
    the oracle's FP64 solve converged to a 1e-8 drop on the 120x40 wedge: the
    converged fields agree to a few 1e-6 (relative L2 per field, velocity
    against the speed), and the FP32 drop floors between 1e-5 and 1e-3."""
    wedge = make_case({"nx": 120, "ny": 40, "mach": 2.5, "theta": 15.0}, str(tmp_path), "c1", jitter=0.0,
                      numbering=0)
    other = _cases(tmp_path)[1]
    with ss.Batch([wedge, other], precision=32) as b:
        out = b.run(3500, history=True)
        st32 = b.state(0)
    o = _oracle(wedge, str(tmp_path / "c1"))
    res64 = o.run(3500)
    ref = o.state()
    drop64 = (res64[-1] / res64[0]).max()
    d32 = (out["hist"][:, 0] / out["hist"][0, 0]).max(axis=1)
    N = wedge.n_cells
    speed = np.linalg.norm(ref[:2, :N])
    err = {nm: float(np.linalg.norm(st32[k, :N] - ref[k, :N]) / (speed if k < 2 else np.linalg.norm(ref[k, :N])))
           for k, nm in enumerate(("u", "v", "p", "T", "rho"))}
    print("oracle FP64 drop %.2e after 3500 iterations; FP32 lowest drop %.2e; FP32 field error %s" %
          (drop64, d32.min(), {k: "%.2e" % v for k, v in err.items()}), flush=True)
    assert drop64 < 1e-8
    assert 1e-5 < d32.min() < 1e-3
    assert max(err.values()) < 5e-6
