"""include/psst_ref_adapter.hpp inside the reference's own build: the
reference reads each mesh with its own Fluent reader and builds its own
psstGridFV (main.cpp:43-54, solver_parameters included); the adapter
flattens that grid (psstGridFV::face_storage()/volume_storage(),
psstFiniteVolumeGrid.h:59-65) and libpsst reads the settings files over the
flattened mesh (psst_case_from_mesh).  Every array of the mesh description,
the settings and the initial state must equal what the library's own reader
gives for the same files (psst_case_open), bit for bit, and the reference's
own initial field (its variables after solver_parameters) must equal the
library's.  The driver is built from the reference's sources where they lie
(oracle/Makefile, target ref_flatten), so the test is skipped where the
reference tree is absent.  CPU only."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from _fixtures import unpack

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
DRIVER = os.path.join(ROOT, "oracle", "_ref", "psst_ref_flatten")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="the reference tree is absent")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "ref_flatten"], check=True)
    return DRIVER


def _read(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            name = f.read(16)
            if not name:
                break
            kind = f.read(1)
            n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
            dt = np.float64 if kind == b"d" else np.int32
            out[name.rstrip(b"\0").decode()] = np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt)
    return out


# the reference's own 80x240 case, an adapted mesh of the reference tree
# (parent zones, trees), its 3D hexahedral wedge and a triangle mesh
CASES = ["ref_wedge80x240", "a_wedge3_hllc_o2_gg_mm_lc2_rk4", "r3d_wedge_hll_o2_gg_va_lc1_rk2",
         "t_hllc_o2_gg_mm_lc2_rk4"]


@pytest.mark.parametrize("name", CASES)
def test_flattened_reference_grid_equals_the_library_reader(driver, tmp_path, name):
    if not os.path.isdir(os.path.join(ROOT, "tests", "golden", name)):
        pytest.skip("fixture %s absent" % name)
    mesh, prefix, _, _ = unpack(name, str(tmp_path))
    out = str(tmp_path / "flat.bin")
    subprocess.run([driver, mesh, prefix, out], check=True, stdout=subprocess.DEVNULL, cwd=str(tmp_path))
    got = _read(out)
    case = ss.Case.open(mesh, prefix)
    m = case.mesh()
    dim, N, G, F = (int(x) for x in got["sizes"])
    assert (dim, N, G, F) == (case.dim, case.n_cells, case.n_ghosts, case.n_faces)
    NT = N + G

    def arr(ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    vo = arr(m.visit_offsets, N + 1)
    lo = arr(m.list_offsets, N + 1)
    expect = {
        "cell_volume": arr(m.cell_volume, NT), "cell_centre": arr(m.cell_centre, 3 * NT)[:dim * NT],
        "face_normal": arr(m.face_normal, 3 * F)[:dim * F], "face_area": arr(m.face_area, F),
        "face_cells": arr(m.face_cells, 2 * F), "face_beta": arr(m.face_beta, F), "visit_offsets": vo,
        "visit_faces": arr(m.visit_faces, vo[-1]), "list_offsets": lo, "list_faces": arr(m.list_faces, lo[-1]),
        "ghost_face": arr(m.ghost_face, G), "ghost_bc": arr(m.ghost_bc, G), "lsq_r": arr(m.lsq_r, 6 * N),
        "lsq_beta": arr(m.lsq_beta, N), "min_volume": np.array([m.min_volume])}
    for k, v in expect.items():
        assert np.array_equal(got[k], v), k
    s = case.settings()
    assert list(got["settings_i"]) == [s.flux, s.order, s.grad, s.limiter, s.time_mode, s.rk_levels]
    assert np.array_equal(got["settings_d"], np.array(list(s.rk_coeff) + [s.cfl, s.gamma, s.rm, s.cp, s.cv,
                                                                         s.uniform_dt]))
    bc = np.array([[b.dynamic_kind, b.thermal_kind, *b.velocity, b.pressure, b.temperature]
                   for b in (s.bc[i] for i in range(s.n_bc))], dtype=np.float64).ravel()
    assert np.array_equal(got["bc"], bc)
    init = case.initial_state().ravel()
    assert np.array_equal(got["prim_case"], init)
    # the reference's own initial field (interior and ghost states)
    assert np.array_equal(got["prim_ref"], init)
    # the state written back into the reference's variables and saved by the
    # reference's own save_data() is the file the library writes for it
    case.write_dat(case.initial_state(), str(tmp_path / "lib.dat"))
    with open(prefix + ".dat", "rb") as f1, open(str(tmp_path / "lib.dat"), "rb") as f2:
        assert f1.read().replace(b"\r", b"") == f2.read()
    print("%s: %d cells, %d ghosts, %d faces: the flattened reference grid equals the library's" % (name, N, G, F))


def test_integration_snippet_runs_up_to_the_device(driver, tmp_path):
    """INTEGRATION.md section 2's replacement of explicit_solve, compiled
    against the reference's headers and objects (oracle/ref_integration_main.cpp
    holds it verbatim): everything before the device runs; on a machine
    without a GPU it stops at psst_ctx_create with the HIP error."""
    mesh, prefix, _, _ = unpack("ref_wedge80x240", str(tmp_path))
    exe = os.path.join(ROOT, "oracle", "_ref", "psst_ref_integration")
    r = subprocess.run([exe, mesh, prefix], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       cwd=str(tmp_path))
    if os.path.exists("/dev/kfd"):  # a GPU: the whole case runs
        assert r.returncode == 0, r.stderr
    else:
        assert r.returncode == 1 and "hipStreamCreate" in r.stderr, r.stderr
