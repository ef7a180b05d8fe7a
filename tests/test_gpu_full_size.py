"""The benchmark meshes at full size on the GPU against the CPU oracle.

* 2D, 3840 x 1280 = 4,915,200 jittered quads: one iteration bit-identical to
  the oracle for HLLC, HLL and AUSM (psstConvectiveFlow.cpp:106-309, :390-606,
  :673-871), RCM and random (seed 2) numbering, fused and unfused kernels;
  the FP32 mode within a few 1e-7 of the FP64 oracle after that iteration.
* 3D, 768 x 256 x 64 = 12,582,912 z-periodic hexahedra, AUSM + DM /
  Venkatakrishnan + least squares (psstGradCalculation.cpp:268-368 with the
  textbook 3D weights): one iteration bit-identical to the oracle.
* 3D, the same mesh split into 75,497,472 tetrahedra: properties the domain
  fixes, since the oracle would need hours -- every residual row and state
  finite and physical, and the inflow region upstream of the ramp, which a
  supersonic flow leaves untouched, still the freestream to round-off.
Needs an MI355X (and about 60 GB of host memory for the oracle's 3D run)."""
import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _first_mismatch(a, b):
    bad = np.argwhere(a != b)
    return None if bad.size == 0 else (tuple(bad[0]), a[tuple(bad[0])], b[tuple(bad[0])], len(bad))


def _oracle(case, prefix, dim):
    raw = case.raw()
    raw["dim"] = dim
    return Oracle(None, prefix, raw=raw)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("numbering,flux", [(1, 3), (2, 3), (1, 2), (1, 1)],
                         ids=["rcm-hllc", "random-hllc", "rcm-hll", "rcm-ausm"])
def test_full_size_wedge_one_iteration_matches_oracle(tmp_path, numbering, flux):
    p = ss.wedge_params(nx=3840, ny=1280, jitter=0.1, jitter_seed=1, numbering=numbering, perm_seed=2)
    prefix = str(tmp_path / "c")
    ss.write_case_files(prefix, time_mode=2, iterations=1, flux=flux, order=2, grad=1, limiter=2)
    case = ss.Case.wedge(p, prefix)
    print("full size: %d cells, numbering %d, flux %d" % (case.n_cells, numbering, flux), flush=True)
    o = _oracle(case, prefix, 2)
    ref_res = o.run(1)
    ref = o.state()
    o.close()
    for fused in ((1, 0) if (numbering, flux) == (1, 3) else (1,)):
        with ss.Solver(case, fused=fused) as s:
            res = s.run(1)
            assert np.array_equal(res, ref_res), fused
            assert _first_mismatch(s.state(), ref) is None, (fused, _first_mismatch(s.state(), ref))
    if (numbering, flux) == (1, 3):
        # FP32 state and fluxes, FP64 geometry and norms, against the FP64 oracle
        with ss.Solver(case, precision=32) as s:
            res32 = s.run(1)
            st32 = s.state()
        N = case.n_cells
        speed = np.linalg.norm(ref[:2, :N])
        for k, nm in enumerate(("u", "v", "p", "T", "rho")):
            rel = np.linalg.norm(st32[k, :N] - ref[k, :N]) / (speed if k < 2 else np.linalg.norm(ref[k, :N]))
            print("FP32 %s: relative L2 %.2e against the FP64 oracle" % (nm, rel), flush=True)
            assert rel < 1e-6
        assert np.allclose(res32, ref_res, rtol=1e-4)


def _wedge3d(tmp_path, tets):
    prefix = str(tmp_path / "w3")
    ss.write_case_files(prefix, dim=3, wall_names=("symm", "wedge", "wall", "up"), iterations=1, cfl=0.8,
                        time_mode=2, rk_levels=4, flux=1, order=2, grad=2, limiter=4, mach=3.0)
    p = ss.wedge_params(nx=768, ny=256, nz=64, depth=1.0, ramp_deg=20.0, jitter=0.0, numbering=1, tets=tets,
                        periodic_z=1)
    return ss.Case.wedge(p, prefix), prefix


@pytest.mark.timeout(900)
def test_full_size_3d_hexahedra_one_iteration_matches_oracle(tmp_path):
    case, prefix = _wedge3d(tmp_path, 0)
    print("3D hexahedra: %d cells, %d faces" % (case.n_cells, case.n_faces), flush=True)
    with ss.Solver(case) as s:
        res = s.run(1)
        st = s.state()
    o = _oracle(case, prefix, 3)
    ref_res = o.run(1)
    assert np.array_equal(res, ref_res)
    assert _first_mismatch(st, o.state()) is None, _first_mismatch(st, o.state())


@pytest.mark.timeout(900)
def test_full_size_3d_tetrahedra_properties(tmp_path):
    case, prefix = _wedge3d(tmp_path, 1)
    N = case.n_cells
    print("3D tetrahedra: %d cells, %d faces" % (N, case.n_faces), flush=True)
    init = case.initial_state()
    with ss.Solver(case) as s:
        res = s.run(30)
        st = s.state()
    print("residual rows 1, 10, 30:", res[0], res[9], res[29], flush=True)
    assert np.isfinite(res).all() and np.isfinite(st).all()
    assert (st[3, :N] > 0).all() and (st[4, :N] > 0).all() and (st[5, :N] > 0).all()
    # upstream of the ramp (x < 0.5) the supersonic inflow is never disturbed:
    # there the state stays the freestream up to the round-off of the
    # tetrahedra's face sums (sum_f n_f A_f = 0 only to rounding), a few 1e-13
    m = case.mesh()
    NT = N + case.n_ghosts
    x = np.ctypeslib.as_array(m.cell_centre, shape=(3 * NT,))[:N]
    up = x < 0.45
    assert up.sum() > 0.1 * N
    scale = np.abs(init[:, :N]).max(axis=1)
    scale[1] = scale[2] = scale[0]  # v, w against the inflow speed
    dev = float((np.abs(st[:, :N][:, up] - init[:, :N][:, up]) / scale[:, None]).max())
    print("upstream cells: %d, largest deviation from the freestream %.2e (relative)" % (up.sum(), dev), flush=True)
    assert dev < 1e-11
