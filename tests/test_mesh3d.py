"""3D extruded wedge meshes (hexahedra, six tetrahedra per hexahedron,
z-periodic or slip side walls) and the oracle's 3D paths that no reference
vector reaches.  CPU only.

Pins:
  * geometry: every cell closed (sum of outward n A = 0), positive volumes,
    total volume = profile area x depth, normals toward neighbour 0;
  * reduction to 2D: on a z-uniform state, a z-periodic two-layer extrusion's
    gradients and convective residual (oracle, 3D) equal the 2D ones of the
    same profile (oracle, 2D, itself bitwise the reference), with zero z
    components -- this pins the periodic face treatment and the 3D flux and
    gradient paths against the 2D ones;
  * least squares (3D, textbook QR weights) exact on linear fields, also
    across periodic faces (fields linear in x, y).
"""
import ctypes

import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from oracle.oracle import Oracle


def _arr(ptr, n, dtype=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype)


def _mesh_arrays(case):
    m = case.mesh()
    N, G, F, d = case.n_cells, case.n_ghosts, case.n_faces, case.dim
    NT = N + G
    vol = _arr(m.cell_volume, NT)
    cen = _arr(m.cell_centre, d * NT).reshape(d, NT)
    nrm = _arr(m.face_normal, d * F).reshape(d, F)
    area = _arr(m.face_area, F)
    fc = _arr(m.face_cells, 2 * F, np.int64).reshape(2, F)
    shift = _arr(m.face_shift, 3 * F).reshape(3, F) if m.face_shift else np.zeros((3, F))
    return vol, cen, nrm, area, fc, shift


def _wedge3d(tmp_path, name, **kw):
    prefix = str(tmp_path / name)
    ss.write_case_files(prefix, dim=3, wall_names=("symm", "wedge", "wall", "up") +
                        (() if kw.get("periodic_z") else ("frontAndBack",)), iterations=10, cfl=0.5,
                        **{k: v for k, v in kw.items() if k in ("flux", "order", "grad", "limiter", "time_mode")})
    p = ss.wedge_params(nx=kw.get("nx", 9), ny=kw.get("ny", 4), nz=kw.get("nz", 3), depth=kw.get("depth", 0.3),
                        ramp_deg=15.0, jitter=kw.get("jitter", 0.1), jitter_seed=5, numbering=kw.get("numbering", 0),
                        tets=kw.get("tets", 0), periodic_z=kw.get("periodic_z", 0))
    return p, prefix


@pytest.mark.parametrize("tets", [0, 1])
@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("numbering", [0, 1, 2])
def test_extruded_mesh_geometry(tmp_path, tets, periodic, numbering):
    p, prefix = _wedge3d(tmp_path, "g", tets=tets, periodic_z=periodic, numbering=numbering, depth=0.3)
    case = ss.Case.wedge(p, prefix)
    vol, cen, nrm, area, fc, shift = _mesh_arrays(case)
    N, G, F = case.n_cells, case.n_ghosts, case.n_faces
    assert N == 9 * 4 * 3 * (6 if tets else 1)
    assert (vol[:N] > 0).all()
    # profile area: 3 x 1 channel minus the ramp (0.6 long at 15 degrees) and the raised floor after it
    h = 0.6 * np.tan(np.radians(15.0))
    area2d = 3.0 - (0.5 * 0.6 * h + (3.0 - 1.1) * h)
    assert vol[:N].sum() == pytest.approx(area2d * 0.3, rel=1e-12)
    # closed cells: sum over faces of (+-) n A = 0
    acc = np.zeros((3, N + G))
    for k in range(3):
        np.add.at(acc[k], fc[0], -nrm[k] * area)
        np.add.at(acc[k], fc[1], nrm[k] * area)
    scale = area.max()
    assert np.abs(acc[:, :N]).max() < 1e-12 * scale
    # normals point toward neighbour 0 (interior faces: from centre 1 to centre 0)
    inner = (fc[0] < N) & (fc[1] < N)
    d = cen[:, fc[0, inner]] - (cen[:, fc[1, inner]] + shift[:, inner])
    assert (np.einsum("kf,kf->f", d, nrm[:, inner]) > 0).all()
    if periodic:
        assert np.count_nonzero(shift[2]) == 9 * 4 * (2 if tets else 1)
        assert np.all(shift[2][shift[2] != 0] == pytest.approx(-0.3))
    else:
        assert not shift.any()


def _column_map(c2, c3):
    """3D cell -> 2D cell with the same (x, y) centre."""
    key2 = {(round(x, 9), round(y, 9)): i for i, (x, y) in enumerate(c2.T)}
    return np.array([key2[(round(x, 9), round(y, 9))] for x, y in c3[:2].T])


@pytest.mark.parametrize("grad", [1, 2])
def test_periodic_extrusion_reduces_to_2d(tmp_path, grad):
    """Oracle: one gradient pass and one convective pass on a z-uniform state
    of a z-periodic two-layer hex extrusion equal the 2D ones on the same
    profile (per unit depth), with zero z components.  Green-Gauss is compared
    away from the walls (the reference's beta indexing, quirk Q1, gives
    boundary faces other faces' weights, which differ between the 2D and the
    3D face numbering); least squares everywhere."""
    kw = dict(flux=3, order=2, grad=grad, limiter=2, time_mode=2)
    pre2 = str(tmp_path / "p2")
    ss.write_case_files(pre2, iterations=10, cfl=0.5, **kw)
    p2 = ss.wedge_params(nx=24, ny=8, ramp_deg=15.0, jitter=0.1, jitter_seed=5)
    m2 = str(tmp_path / "m2.msh")
    ss.write_wedge_msh(p2, m2)
    dz = 0.1
    p3, pre3 = _wedge3d(tmp_path, "p3", nx=24, ny=8, nz=2, depth=2 * dz, periodic_z=1, **kw)
    m3 = str(tmp_path / "m3.msh")
    ss.write_wedge_msh(p3, m3)
    o2, o3 = Oracle(m2, pre2), Oracle(m3, pre3)
    o2.run(30)  # a state with a shock
    s2 = o2.state()
    c2, c3 = ss.Case.open(m2, pre2), ss.Case.open(m3, pre3)
    N2, N3 = c2.n_cells, c3.n_cells
    _, cen3, _, _, fc3, _ = _mesh_arrays(c3)
    _, cen2, _, _, fc2, _ = _mesh_arrays(c2)
    col = _column_map(cen2[:, :N2], cen3[:, :N3])
    # ghosts too: a 3D boundary face's ghost sits at its face centre, above the 2D ghost's
    gcol = _column_map(cen2[:, N2:], cen3[:, N3:]) + N2
    allcol = np.concatenate([col, gcol])
    s3 = c3.initial_state()
    for r3, r2 in ((0, 0), (1, 1), (3, 2), (4, 3), (5, 4)):
        s3[r3] = s2[r2, allcol]
    s3[2] = 0.0
    o3.set_state(s3)
    g2, g3 = o2.gradients(), o3.gradients()      # [(dim+2)][dim][N]
    R2, R3 = o2.residual(), o3.residual()        # [(dim+2)][N]
    if grad == 1:
        # 2D cells two layers away from every boundary face
        bnd = set(fc2[0][fc2[1] >= N2]) | set(fc2[1][fc2[0] >= N2])
        near = set(bnd)
        for f in range(fc2.shape[1]):
            a, b = fc2[0, f], fc2[1, f]
            if a in bnd and b < N2:
                near.add(b)
            if b in bnd and a < N2:
                near.add(a)
        keep = np.array([c not in near for c in col])
    else:
        keep = np.ones(N3, dtype=bool)
    assert keep.sum() > 20
    for v3, v2 in ((0, 0), (1, 1), (3, 2), (4, 3)):
        for k in range(2):
            a, b = g3[v3, k, keep], g2[v2, k, col[keep]]
            assert np.abs(a - b).max() <= 1e-8 * max(np.abs(b).max(), 1e-300), (v3, k)
        assert np.abs(g3[v3, 2, keep]).max() <= 1e-8 * max(np.abs(g2[v2, :, :]).max(), 1e-300)
    for v3, v2 in ((0, 0), (1, 1), (3, 2), (4, 3)):
        a, b = R3[v3, keep] / dz, R2[v2, col[keep]]
        assert np.abs(a - b).max() <= 1e-8 * np.abs(R2[v2]).max(), v3
    assert np.abs(R3[2, keep]).max() <= 1e-8 * np.abs(R3[0]).max()


@pytest.mark.parametrize("tets", [0, 1])
def test_least_squares_3d_exact_on_linear_fields(tmp_path, tets):
    p, prefix = _wedge3d(tmp_path, "l", tets=tets, periodic_z=1, nz=3, depth=0.3, flux=3, order=2, grad=2,
                         limiter=2, time_mode=2)
    mesh = str(tmp_path / "l.msh")
    ss.write_wedge_msh(p, mesh)
    o = Oracle(mesh, prefix)
    case = ss.Case.open(mesh, prefix)
    _, cen, _, _, _, _ = _mesh_arrays(case)
    q = case.initial_state().copy()
    x, y = cen[0], cen[1]
    coef = np.array([[3.0, -2.0], [0.5, 1.5], [-1.0, 4.0], [2.0, 0.25], [7.0, -3.0]])
    for r, (a, b) in enumerate(coef):
        q[r] = 100.0 + a * x + b * y
    q[5] = 1.0 + 0.1 * x
    o.set_state(q)
    g = o.gradients()  # [(dim+2)][dim][N]
    for r, (a, b) in enumerate(coef):
        assert np.allclose(g[r, 0], a, atol=1e-9) and np.allclose(g[r, 1], b, atol=1e-9)
        assert np.allclose(g[r, 2], 0.0, atol=1e-9)
