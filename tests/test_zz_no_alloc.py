"""psst_run_noalloc (a run that allocates nothing; its workspace is a fixed
device array per GPU) against psst_run: the same residual rows, stopping
iteration and final state, bitwise in FP64, through the C-ABI.  Two calls in
a row, the second on a smaller mesh with a tolerance stop inside the call,
then the size limit.  Needs an MI355X."""
import ctypes as C

import numpy as np
import pytest

import supersonicflowsolver_amd as ss
from supersonicflowsolver_amd.sweep import make_case

pytestmark = pytest.mark.gpu


def _ctx(case):
    L = ss._lib.lib()
    m, st = case.mesh(), case.settings()
    h = C.c_void_p()
    assert L.psst_ctx_create(C.byref(m), C.byref(st), C.byref(h)) == 0, L.psst_last_error(None)
    prim = np.ascontiguousarray(case.initial_state())
    assert L.psst_set_state(h, prim.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return h


def _run(fn, h, case, n, tol):
    res = np.zeros((n, case.dim + 2))
    done = C.c_int32(-1)
    st = fn(h, n, tol, res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done))
    prim = np.zeros((case.dim + 3, case.n_cells + case.n_ghosts))
    assert ss._lib.lib().psst_get_state(h, prim.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return st, done.value, res, prim


def test_run_noalloc_matches_run(tmp_path):
    L = ss._lib.lib()
    big = make_case({"nx": 60, "ny": 20, "mach": 2.5, "theta": 15.0}, str(tmp_path), "big", jitter=0.1)
    small = make_case({"nx": 36, "ny": 12, "mach": 3.1, "theta": 9.0}, str(tmp_path), "small", jitter=0.1,
                      numbering=2)
    for case, n, tol in ((big, 40, 0.0), (small, 300, 2e-2)):
        ha, hb = _ctx(case), _ctx(case)
        try:
            sa, da, ra, pa = _run(L.psst_run, ha, case, n, tol)
            sb, db, rb, pb = _run(L.psst_run_noalloc, hb, case, n, tol)
            print("%d cells, %d iterations, tol %g: psst_run %d/%d, psst_run_noalloc %d/%d" %
                  (case.n_cells, n, tol, sa, da, sb, db), flush=True)
            assert sa == sb == 0
            assert da == db and (tol == 0.0 or 0 < da < n)
            assert np.array_equal(ra, rb)
            assert np.array_equal(pa, pb)
        finally:
            L.psst_ctx_destroy(ha)
            L.psst_ctx_destroy(hb)
    h = _ctx(small)
    try:
        res = np.zeros((1025, small.dim + 2))
        done = C.c_int32(-1)
        assert L.psst_run_noalloc(h, 1025, 0.0, res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done)) == 1
        assert b"1024" in L.psst_last_error(h)
    finally:
        L.psst_ctx_destroy(h)
