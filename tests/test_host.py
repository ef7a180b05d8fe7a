"""Host side of the library (CPU only): the C-ABI library loads and exports
every entry point of include/psst_gpu.h and include/psst_mlp.h; the host reader, grid flattening
and settings files reproduce the oracle's setup; the generator's meshes
round-trip through the Fluent writer and reader."""
import ctypes
import os
import re

import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from _fixtures import GOLDEN, ROOT, unpack
from oracle.oracle import Oracle


def test_library_exports_every_declared_symbol():
    text = ""
    for h in ("psst_gpu.h", "psst_mlp.h"):
        with open(os.path.join(ROOT, "include", h)) as f:
            text += f.read()
    declared = set(re.findall(r"\b(psst_[a-z_0-9]+)\s*\(", text))
    assert len(declared) >= 20
    L = ctypes.CDLL(ss._lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(ss._lib.exported_symbols())


@pytest.mark.parametrize("name", ["ref_wedge80x240", "c1_lc1", "m_hll_o2_lsq_mm_lc2_rk3", "m_ausm_o2_lsq_vl_lc0_rk2",
                                  "a_wedge3_ausm_o2_lsq_va_lc0_rk3", "a_wedge5_hllc_o2_gg_va_lc0_rk2",
                                  "restart_hllc_o2_gg_mm_lc2_rk4"])
def test_host_setup_matches_oracle(tmp_path, name):
    mesh, prefix, _, _ = unpack(name, str(tmp_path))
    case = ss.Case.open(mesh, prefix)
    o = Oracle(mesh, prefix)
    assert (case.dim, case.n_cells, case.n_ghosts, case.n_faces) == (o.dim, o.n_cells, o.n_ghosts, o.n_faces)
    assert np.array_equal(case.initial_state(), o.state())
    s = case.settings()
    assert (s.gamma, s.rm, s.cp, s.cv, s.cfl) == (o.gamma, o.rm, o.cp, o.cv, o.cfl)
    assert s.rk_levels == o.rk_levels
    if s.time_mode == 0:
        assert s.uniform_dt == o.uniform_dt
    assert case.mesh().min_volume == o.min_vol


def test_generator_roundtrip(tmp_path):
    p = ss.wedge_params(nx=40, ny=12, jitter=0.1, numbering=2)
    path = str(tmp_path / "w.msh")
    ss.write_wedge_msh(p, path)
    ss.write_case_files(str(tmp_path / "c"))
    a = ss.Case.open(path, str(tmp_path / "c")).raw()
    b = ss.Case.wedge(p, str(tmp_path / "c")).raw()
    for k in a:
        if k == "zone_names":
            assert a[k] == b[k]
        else:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("numbering", [0, 1, 2])
def test_generator_mesh_is_closed_and_oriented(tmp_path, numbering):
    """Every interior cell's faces close (sum of n A = 0), volumes are
    positive and sum to the channel area, and each normal points toward
    neighbour 0 (from centroid 1 to centroid 0)."""
    p = ss.wedge_params(nx=30, ny=10, jitter=0.1, numbering=numbering)
    ss.write_case_files(str(tmp_path / "c"))
    case = ss.Case.wedge(p, str(tmp_path / "c"))
    m = case.mesh()
    N, G, F = case.n_cells, case.n_ghosts, case.n_faces
    NT = N + G
    vol = np.ctypeslib.as_array(m.cell_volume, shape=(NT,))
    cx = np.ctypeslib.as_array(m.cell_centre, shape=(3 * NT,)).reshape(3, NT)
    fn = np.ctypeslib.as_array(m.face_normal, shape=(3 * F,)).reshape(3, F)
    fa = np.ctypeslib.as_array(m.face_area, shape=(F,))
    fc = np.ctypeslib.as_array(m.face_cells, shape=(2 * F,)).reshape(2, F)
    assert (vol[:N] > 0).all() and (vol[N:] == 0).all()
    area = 3.0 * 1.0 - 0.5 * 0.6 * 0.6 * np.tan(np.radians(15)) - 1.9 * 0.6 * np.tan(np.radians(15))
    assert vol[:N].sum() == pytest.approx(area, rel=1e-12)
    acc = np.zeros((2, NT))
    for k in range(2):
        np.add.at(acc[k], fc[0], -fn[k] * fa)
        np.add.at(acc[k], fc[1], fn[k] * fa)
    assert np.abs(acc[:, :N]).max() < 1e-14
    d = cx[:2, fc[0]] - cx[:2, fc[1]]
    assert ((d * fn[:2]).sum(axis=0) > 0).all()


# Two triangles filling the unit square: cell 1 = (0,0) (1,0) (1,1), cell 2 =
# (0,0) (1,1) (0,1).  Faces are "v0 v1 c0 c1" with the normal
# (y0 - y1, x1 - x0)/A pointing toward c0 (the ghost side on the boundary).
TRIANGLES_MSH = """(0 "two triangles")
(2 2)
(12 (0 1 2 0))
(12 (2 1 2 1 1))
(13 (0 1 5 0))
(13 (3 1 1 3 2)(
 2 1 0 1
)
)
(13 (4 2 2 3 2)(
 3 2 0 1
)
)
(13 (5 3 3 3 2)(
 4 3 0 2
)
)
(13 (6 4 4 3 2)(
 1 4 0 2
)
)
(13 (a 5 5 2 2)(
 3 1 1 2
)
)
(10 (0 1 4 1 2))
(10 (1 1 4 1 2)(
 0.0 0.0
 1.0 0.0
 1.0 1.0
 0.0 1.0
)
)
(39 (2 fluid fluid)())
(39 (3 wall wall)())
(39 (4 wall out)())
(39 (5 wall up)())
(39 (6 wall in)())
(39 (10 interior default-interior)())
"""


def test_triangle_geometry_known_answers(tmp_path):
    """Host geometry on triangles against closed forms: areas 1/2, centroids
    at the vertex means, the diagonal's length sqrt(2) and unit normal toward
    neighbour 0, outward unit normals on the boundary, ghosts of zero volume
    at their face centres, a Green-Gauss weight of 1/2 on the symmetric
    diagonal, and closed cells (sum of n A = 0)."""
    mesh = str(tmp_path / "tri.msh")
    with open(mesh, "w") as f:
        f.write(TRIANGLES_MSH)
    prefix = str(tmp_path / "c")
    ss.write_case_files(prefix, wall_names=("wall", "up"))
    case = ss.Case.open(mesh, prefix)
    N, G, F = case.n_cells, case.n_ghosts, case.n_faces
    assert (case.dim, N, G, F) == (2, 2, 4, 5)
    m = case.mesh()
    NT = N + G
    vol = np.ctypeslib.as_array(m.cell_volume, shape=(NT,))
    cx = np.ctypeslib.as_array(m.cell_centre, shape=(3 * NT,)).reshape(3, NT)[:2]
    fn = np.ctypeslib.as_array(m.face_normal, shape=(3 * F,)).reshape(3, F)[:2]
    fa = np.ctypeslib.as_array(m.face_area, shape=(F,))
    fc = np.ctypeslib.as_array(m.face_cells, shape=(2 * F,)).reshape(2, F)
    beta = np.ctypeslib.as_array(m.face_beta, shape=(F,))
    assert np.allclose(vol[:N], 0.5, rtol=0, atol=1e-15) and (vol[N:] == 0).all()
    assert m.min_volume == pytest.approx(0.5, abs=1e-15)
    cents = cx[:, :N].T[np.argsort(cx[0, :N])]
    assert np.allclose(cents, [(1 / 3, 2 / 3), (2 / 3, 1 / 3)], rtol=0, atol=1e-15)
    inner = [f for f in range(F) if fc[0, f] < N and fc[1, f] < N]
    assert len(inner) == 1
    d = inner[0]
    assert fa[d] == pytest.approx(np.sqrt(2.0), abs=1e-15)
    c0, c1 = fc[0, d], fc[1, d]
    assert np.dot(fn[:, d], cx[:, c0] - cx[:, c1]) > 0
    assert np.allclose(np.abs(fn[:, d]), np.sqrt(0.5), rtol=0, atol=1e-15)
    assert beta[d] == pytest.approx(0.5, abs=1e-14)
    # boundary faces: unit length, outward normal, ghost at the face centre
    expected = {(0.5, 0.0): (0.0, -1.0), (1.0, 0.5): (1.0, 0.0), (0.5, 1.0): (0.0, 1.0), (0.0, 0.5): (-1.0, 0.0)}
    for f in range(F):
        if f == d:
            continue
        ghost = fc[0, f] if fc[0, f] >= N else fc[1, f]
        centre = tuple(np.round(cx[:, ghost], 14) + 0.0)
        assert centre in expected, centre
        assert fa[f] == pytest.approx(1.0, abs=1e-15)
        assert np.allclose(fn[:, f], expected[centre], rtol=0, atol=1e-15)
    acc = np.zeros((2, NT))
    for k in range(2):
        np.add.at(acc[k], fc[0], -fn[k] * fa)
        np.add.at(acc[k], fc[1], fn[k] * fa)
    assert np.abs(acc[:, :N]).max() < 1e-15
    # the oracle reads the same mesh to the same set-up
    o = Oracle(mesh, prefix)
    assert (o.n_cells, o.n_ghosts, o.n_faces) == (N, G, F) and o.min_vol == m.min_volume
    assert np.array_equal(case.initial_state(), o.state())


def test_bad_inputs_raise(tmp_path):
    with pytest.raises(ss.PsstError):
        ss.Case.open(str(tmp_path / "missing.msh"), str(tmp_path / "c"))
    mesh, prefix, _, _ = unpack("m_hll_o1_lc0_rk1", str(tmp_path))
    with open(prefix + ".bvp", "a") as f:
        f.write("\nnosuchzone\t15")
    with pytest.raises(ss.PsstError, match="zone"):
        ss.Case.open(mesh, prefix)


def test_settings_with_a_boundary_table_of_another_size_are_rejected(tmp_path):
    mesh, prefix, _, _ = unpack("m_hll_o1_lc0_rk1", str(tmp_path))
    case = ss.Case.open(mesh, prefix)
    s = case.settings()
    n = s.n_bc
    table = (ss.BC * (n + 1))()
    for i in range(n):
        table[i] = s.bc[i]
    table[n] = s.bc[0]
    s.bc, s.n_bc = ctypes.cast(table, ctypes.POINTER(ss.BC)), n + 1
    with pytest.raises(ss.PsstError):
        case.set_settings(s)
    # unchanged, and a table of the right size is taken entry for entry
    t = case.settings()
    assert t.n_bc == n
    s.n_bc = n
    s.cfl = 0.5
    case.set_settings(s)
    assert case.settings().cfl == 0.5


@pytest.mark.parametrize("field,value,message", [("dim", 4, "dim"), ("n_cells", 0, "empty mesh"),
                                                  ("rk_levels", 9, "rk_levels"), ("flux", 7, "flux"),
                                                  ("order", 3, "order"), ("limiter", 0, "limiter")])
def test_context_creation_rejects_bad_arguments(tmp_path, field, value, message):
    """psst_ctx_create validates the mesh and settings before touching the
    device: PSST_BAD_ARG and a message, no exception across the ABI."""
    mesh, prefix, _, _ = unpack("m_hllc_o2_gg_mm_lc2_rk4", str(tmp_path))
    case = ss.Case.open(mesh, prefix)
    m, s = case.mesh(), case.settings()
    setattr(m if field in ("dim", "n_cells") else s, field, value)
    h = ctypes.c_void_p()
    L = ss._lib.lib()
    assert L.psst_ctx_create(ctypes.byref(m), ctypes.byref(s), ctypes.byref(h)) == 1
    assert not h.value
    assert message in L.psst_last_error(None).decode()
    assert L.psst_ctx_create(None, ctypes.byref(s), ctypes.byref(h)) == 1


def test_trailing_newline_in_boundary_files_is_accepted(tmp_path):
    """The reference crashes on a trailing newline in .bvp/.btq (quirk Q11);
    the host reader accepts it and reads the same boundary set."""
    mesh, prefix, _, _ = unpack("m_hll_o1_lc0_rk1", str(tmp_path))
    a = ss.Case.open(mesh, prefix).initial_state()
    for ext in ("bvp", "btq"):
        with open(prefix + "." + ext, "a") as f:
            f.write("\n")
    b = ss.Case.open(mesh, prefix).initial_state()
    assert np.array_equal(a, b)


def test_crlf_mesh_is_accepted(tmp_path):
    """CRLF line ends hang the reference reader (quirk Q12); the host reader
    parses by structure."""
    mesh, prefix, _, _ = unpack("m_hll_o1_lc0_rk1", str(tmp_path))
    with open(mesh, "rb") as f:
        data = f.read()
    with open(mesh + ".crlf", "wb") as f:
        f.write(data.replace(b"\n", b"\r\n"))
    a = ss.Case.open(mesh, prefix).raw()
    b = ss.Case.open(mesh + ".crlf", prefix).raw()
    assert np.array_equal(a["xyz"], b["xyz"]) and np.array_equal(a["face_v"], b["face_v"])


def _face_sections(text):
    """Spans of the face-zone sections "(13 (id first last bc type)( ... ))"
    of a mesh written by the package's generator (header zone 0 excluded)."""
    spans = []
    pos = 0
    while True:
        a = text.find("(13 (", pos)
        if a < 0:
            return spans
        head = text[a + 5:text.index(")", a)].split()
        b = text.index(")\n)\n", a) + 4 if head[0] != "0" else text.index("\n", a) + 1
        if head[0] != "0":
            spans.append((a, b))
        pos = b


def test_face_sections_in_any_order(tmp_path):
    """Face sections are placed at their declared indices: a file whose face
    zones are not in index order (as the reference's own .cas writer emits
    them) reads to the same faces."""
    mesh, prefix, _, _ = unpack("m_hll_o1_lc0_rk1", str(tmp_path))
    with open(mesh) as f:
        text = f.read()
    spans = _face_sections(text)
    assert len(spans) >= 3
    parts = [text[a:b] for a, b in spans]
    shuffled = text[:spans[0][0]] + "".join(parts[::-1]) + text[spans[-1][1]:]
    assert shuffled != text
    with open(mesh + ".shuffled", "w") as f:
        f.write(shuffled)
    a = ss.Case.open(mesh, prefix).raw()
    b = ss.Case.open(mesh + ".shuffled", prefix).raw()
    for k in ("xyz", "face_nv", "face_v", "face_c", "face_zone", "cell_zone"):
        assert np.array_equal(a[k], b[k]), k
    assert sorted(a["zone_id"].tolist()) == sorted(b["zone_id"].tolist())


def test_inconsistent_face_sections_raise(tmp_path):
    mesh, prefix, _, _ = unpack("m_hll_o1_lc0_rk1", str(tmp_path))
    with open(mesh) as f:
        text = f.read()
    spans = _face_sections(text)
    a, b = spans[1]
    with open(mesh + ".twice", "w") as f:
        f.write(text[:b] + text[a:b] + text[b:])
    with pytest.raises(ss.PsstError, match="listed twice"):
        ss.Case.open(mesh + ".twice", prefix)
    with open(mesh + ".gap", "w") as f:
        f.write(text[:a] + text[b:])
    with pytest.raises(ss.PsstError, match="without a face|outside"):
        ss.Case.open(mesh + ".gap", prefix)


def test_sweep_generator_and_draws():
    """splitmix64 first output for seed 0 is Vigna's published value; the
    instance draw is deterministic and inside its ranges."""
    from supersonicflowsolver_amd.sweep import _splitmix64, draw_instances
    _, z = _splitmix64(0)
    assert z == 0xE220A8397B1DCDAF
    a = draw_instances(64, seed=3)
    assert a == draw_instances(64, seed=3)
    assert a != draw_instances(64, seed=4)
    for d in a:
        assert 20000 * 0.97 <= d["nx"] * d["ny"] <= 400000 * 1.03
        assert 1.5 <= d["mach"] < 4.0 and 5.0 <= d["theta"] < 25.0
        assert d["nx"] == 3 * d["ny"]


def test_case_from_a_flattened_mesh_round_trip(tmp_path):
    """psst_case_from_mesh reads the settings files over a mesh the caller
    flattened (here: the library's own flattening of the reference's 80x240
    case, handed back with its zone table): same settings and initial state
    as psst_case_open; the writers that need the Fluent file are refused."""
    import ctypes as C

    import numpy as np

    import supersonicflowsolver_amd as ss
    from _fixtures import unpack
    from supersonicflowsolver_amd._lib import Zone
    mesh, prefix, _, _ = unpack("ref_wedge80x240", str(tmp_path))
    case = ss.Case.open(mesh, prefix)
    raw = case.raw()
    order = np.argsort(raw["zone_id"])
    keep = [int(i) for i in order if raw["zone_range"][i][1] > 0]
    names = [raw["zone_names"][i].encode() for i in keep]
    zones = (Zone * len(keep))(*[Zone(int(raw["zone_id"][i]), int(raw["zone_bc"][i]),
                                      int(raw["zone_range"][i][1] - raw["zone_range"][i][0] + 1), names[k])
                                 for k, i in enumerate(keep)])
    m = case.mesh()
    L = ss._lib.lib()
    h = C.c_void_p()
    assert L.psst_case_from_mesh(C.byref(m), len(keep), zones, prefix.encode(), C.byref(h)) == 0, \
        L.psst_case_last_error()
    c2 = ss.Case(h.value)
    try:
        assert (c2.n_cells, c2.n_ghosts, c2.n_faces) == (case.n_cells, case.n_ghosts, case.n_faces)
        assert np.array_equal(c2.initial_state(), case.initial_state())
        s1, s2 = case.settings(), c2.settings()
        for name, _ in ss._lib.Settings._fields_:
            if name == "bc":
                continue
            v1, v2 = getattr(s1, name), getattr(s2, name)
            assert (list(v1) == list(v2)) if name == "rk_coeff" else v1 == v2, name
        for i in range(s1.n_bc):
            assert bytes(s1.bc[i]) == bytes(s2.bc[i])
        m2 = c2.mesh()
        for fld in ("cell_volume", "face_area", "face_beta"):
            n = case.n_cells + case.n_ghosts if fld == "cell_volume" else case.n_faces
            assert np.array_equal(np.ctypeslib.as_array(getattr(m, fld), (n,)), np.ctypeslib.as_array(getattr(m2, fld), (n,)))
        with pytest.raises(ss.PsstError) as ei:
            c2.write_cas(str(tmp_path / "x.cas"))
        assert ei.value.status == 5
        c2.write_dat(c2.initial_state(), str(tmp_path / "x.dat"))
        case.write_dat(case.initial_state(), str(tmp_path / "y.dat"))
        assert open(tmp_path / "x.dat", "rb").read() == open(tmp_path / "y.dat", "rb").read()
    finally:
        c2.close()
