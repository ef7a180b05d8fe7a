"""Physics validation of the HIP path (not parity): a Sod shock tube against
the exact Riemann solution and the supersonic wedge against the oblique-shock
(theta-beta-M) relation.  Needs an MI355X."""
import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from _physics import SOD_ITERS, WEDGE_ITERS, sod_case, sod_errors, wedge_case, wedge_plateau

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("flux,precision", [(3, 64), (2, 64), (1, 64), (3, 32)])
def test_sod_shock_tube_matches_the_exact_riemann_solution(tmp_path, flux, precision):
    """600 cells, second order (DM/MinMod, Green-Gauss), RK4 at a uniform
    time step.  The oracle's run of the same set-up is within 0.16% (L1
    density) of the exact solution and within 1e-4 of p*."""
    _, _, case, prim, x = sod_case(str(tmp_path), flux)
    with ss.Solver(case, precision=precision) as s:
        s.set_state(prim)
        s.run(SOD_ITERS)
        q = s.state()
    e = sod_errors(case, x, q)
    print("flux", flux, "precision", precision, e)
    assert e["rho"] < 3e-3 and e["p"] < 2e-3 and e["u"] < 3e-3, e
    assert abs(e["p_star"]) < 1e-3, e
    assert e["v_max"] < 1e-6, e


def test_oblique_shock_on_the_wedge_matches_theta_beta_mach(tmp_path):
    """M 2.5 over a 15 degree ramp: weak shock at 36.94 degrees, p2/p1 =
    2.4675, rho2/rho1 = 1.8665, flow turned by 15 degrees.  On 120 x 40 cells
    after 3000 iterations the plateau is within 1.5% and the shock within two
    cells (0.05) of the exact line."""
    _, _, case = wedge_case(str(tmp_path))
    with ss.Solver(case) as s:
        s.run(WEDGE_ITERS)
        q = s.state()
    r = wedge_plateau(case, q)
    print(r)
    assert r["cells"] >= 30
    assert abs(r["p21"] / r["p21_exact"] - 1.0) < 0.015, r
    assert abs(r["r21"] / r["r21_exact"] - 1.0) < 0.015, r
    assert abs(r["angle"] - 15.0) < 0.5, r
    for xs, xe in r["shock_x"]:
        assert abs(xs - xe) < 0.05, r
