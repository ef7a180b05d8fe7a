"""The schedule of the reference-ordered time step on large meshes (host only).

Time step mode 1 updates the cells in place in the reference's loop order
(psstFiniteVolumeGrid.cpp:317-347), and a cell's time step reads the new
state of its earlier face-list neighbours (psstLocalTimeCalculation.cpp:
90-131).  The device sweeps that chain by chunks of cells, one workgroup
each, which wait for each other only where they share faces.
psst_debug_sweep_schedule builds the schedule a device context would use and
checks it by simulation: every earlier neighbour of a cell is finished
before the cell, every wait is for a level that its chunk publishes, and
every chunk finishes (no cycle of waits, which on the GPU would be a hang).
"""
import pytest

import supersonicflowsolver_amd as ss
from tests._fixtures import ADAPTED_CASES, THREE_D_CASES, unpack


def _wedge(tmp_path, nx, ny, numbering, jitter=0.1, nz=0, tets=0):
    p = ss.wedge_params(nx=nx, ny=ny, nz=nz, jitter=jitter, jitter_seed=4, numbering=numbering, perm_seed=2,
                        tets=tets, depth=0.5 if nz else 1.0)
    prefix = str(tmp_path / "c")
    walls = ("symm", "wedge", "wall", "up", "frontAndBack") if nz else ("symm", "wedge", "wall", "up")
    ss.write_case_files(prefix, dim=3 if nz else 2, time_mode=1, iterations=3, flux=3, order=2, grad=1, limiter=2,
                        wall_names=walls)
    return ss.Case.wedge(p, prefix)


@pytest.mark.parametrize("numbering", [1, 0, 2], ids=["rcm", "column", "random"])
@pytest.mark.parametrize("partition", [0, 1, 2], ids=["auto", "morton", "potential"])
def test_schedule_is_valid_on_2d_wedges(tmp_path, numbering, partition):
    case = _wedge(tmp_path, 160, 80, numbering)
    st = ss.sweep_schedule(case, 1024, partition)
    assert st["chunks"] >= 1 and st["entries"] >= st["levels"]
    # the makespan can never beat the dependency chain itself
    assert st["makespan_free"] >= st["levels"]


def test_levels_of_a_structured_numbering_are_its_anti_diagonals(tmp_path):
    """Column-major and RCM numberings of an nx x ny grid: a cell's earlier
    neighbours are its left and lower ones, so the levels are the
    anti-diagonals, nx + ny - 1 of them."""
    for numbering in (0, 1):
        case = _wedge(tmp_path, 96, 40, numbering)
        assert ss.sweep_schedule(case, 1024, 1)["levels"] == 96 + 40 - 1


def test_potential_boxes_follow_the_levels_on_a_jittered_grid(tmp_path):
    """Boxes of the per-axis potentials are rectangles of the grid, so no two
    chunks wait for each other: with waits costing four workgroup passes the
    critical path stays within a few per cent of the level count, whereas
    Morton ranges of jittered centroids have ragged borders that make
    neighbouring chunks wait for each other back and forth."""
    case = _wedge(tmp_path, 512, 256, 1)
    pot = ss.sweep_schedule(case, 4096, 2)
    mor = ss.sweep_schedule(case, 4096, 1)
    assert pot["levels"] == 512 + 256 - 1
    assert pot["makespan_wait4"] <= 1.1 * pot["levels"], pot
    assert pot["makespan_wait4"] < mor["makespan_wait4"], (pot, mor)
    assert ss.sweep_schedule(case, 4096, 0)["partition"] == 2
    # publications every few levels instead of every waited-for level
    assert pot["publishes"] < pot["entries"] / 2, pot


def test_random_numbering_keeps_morton_chunks(tmp_path):
    """A random numbering has few, wide levels; its potentials barely grow, so
    their boxes would hold the whole mesh.  The context keeps Morton ranges."""
    case = _wedge(tmp_path, 160, 80, 2)
    st = ss.sweep_schedule(case, 1024, 0)
    assert st["partition"] == 1 and st["chunks"] > 1


def test_one_cell_chunks(tmp_path):
    """The degenerate partition: every chunk one cell, every earlier neighbour
    a wait."""
    case = _wedge(tmp_path, 24, 12, 1)
    st = ss.sweep_schedule(case, 1, 1)
    assert st["chunks"] == case.n_cells and st["entries"] == case.n_cells


@pytest.mark.parametrize("tets", [0, 1], ids=["hexahedra", "tetrahedra"])
def test_schedule_is_valid_in_3d(tmp_path, tets):
    case = _wedge(tmp_path, 24, 12, 1, jitter=0.0, nz=8, tets=tets)
    for partition in (0, 1, 2):
        st = ss.sweep_schedule(case, 512, partition)
        assert st["makespan_free"] >= st["levels"]


@pytest.mark.parametrize("name", [ADAPTED_CASES[0], THREE_D_CASES[0]] if ADAPTED_CASES and THREE_D_CASES else [])
def test_schedule_is_valid_on_the_reference_meshes(tmp_path, name):
    """The reference's adapted (hanging-node, cells of up to eight faces) and
    3D meshes, in the reference's own numbering."""
    mesh, prefix, _, _ = unpack(name, str(tmp_path))
    case = ss.Case.open(mesh, prefix)
    for partition in (0, 1, 2):
        for k in (64, 1024):
            st = ss.sweep_schedule(case, k, partition)
            assert st["makespan_free"] >= st["levels"]


def test_bad_arguments(tmp_path):
    case = _wedge(tmp_path, 24, 12, 1)
    with pytest.raises(ss.PsstError):
        ss.sweep_schedule(case, 64, 3)
