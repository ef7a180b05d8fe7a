"""ctypes bindings of ``libpsst.so`` (C-ABI in ``include/psst_gpu.h``).

The library is built in-tree by ``supersonicflowsolver_amd.build``.  Loading
fails loudly if it is missing: there is no Python or CPU fallback for the
solver path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsst.so")

PSST_OK = 0
STATUS_NAMES = {0: "PSST_OK", 1: "PSST_BAD_ARG", 2: "PSST_HIP_ERROR", 3: "PSST_NAN_DETECTED",
                4: "PSST_NO_CONVERGENCE", 5: "PSST_UNSUPPORTED", 6: "PSST_IO_ERROR"}


class MeshDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("n_cells", C.c_int32), ("n_ghosts", C.c_int32), ("n_faces", C.c_int32),
        ("cell_volume", C.POINTER(C.c_double)), ("cell_centre", C.POINTER(C.c_double)),
        ("face_normal", C.POINTER(C.c_double)), ("face_area", C.POINTER(C.c_double)),
        ("face_cells", C.POINTER(C.c_int32)), ("face_beta", C.POINTER(C.c_double)),
        ("visit_offsets", C.POINTER(C.c_int32)), ("visit_faces", C.POINTER(C.c_int32)),
        ("list_offsets", C.POINTER(C.c_int32)), ("list_faces", C.POINTER(C.c_int32)),
        ("ghost_face", C.POINTER(C.c_int32)), ("ghost_bc", C.POINTER(C.c_int32)),
        ("lsq_r", C.POINTER(C.c_double)), ("lsq_beta", C.POINTER(C.c_double)),
        ("min_volume", C.c_double), ("face_shift", C.POINTER(C.c_double)),
    ]


class BC(C.Structure):
    _fields_ = [("dynamic_kind", C.c_int32), ("thermal_kind", C.c_int32), ("velocity", C.c_double * 3),
                ("pressure", C.c_double), ("temperature", C.c_double)]


class Settings(C.Structure):
    _fields_ = [
        ("flux", C.c_int32), ("order", C.c_int32), ("grad", C.c_int32), ("limiter", C.c_int32),
        ("time_mode", C.c_int32), ("rk_levels", C.c_int32), ("rk_coeff", C.c_double * 8),
        ("cfl", C.c_double), ("gamma", C.c_double), ("rm", C.c_double), ("cp", C.c_double), ("cv", C.c_double),
        ("uniform_dt", C.c_double), ("precision", C.c_int32), ("dt_ordering", C.c_int32), ("fused", C.c_int32),
        ("n_bc", C.c_int32), ("bc", C.POINTER(BC)), ("venkat_k", C.c_double),
        ("tile_cells", C.c_int32), ("tile_threads", C.c_int32),
    ]


class Zone(C.Structure):
    _fields_ = [("id", C.c_int64), ("type", C.c_int32), ("count", C.c_int32), ("name", C.c_char_p)]


class WedgeParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("length", C.c_double), ("height", C.c_double), ("depth", C.c_double),
        ("ramp_start", C.c_double), ("ramp_length", C.c_double), ("ramp_deg", C.c_double),
        ("jitter", C.c_double), ("jitter_seed", C.c_uint64), ("numbering", C.c_int32),
        ("perm_seed", C.c_uint64), ("tets", C.c_int32), ("periodic_z", C.c_int32),
    ]


_lib = None

_SIGS = {
    "psst_ctx_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Settings), C.POINTER(C.c_void_p)]),
    "psst_set_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "psst_run_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int32)]),
    "psst_get_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_get_conservative": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_set_residual_reference": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_debug_gradients": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_debug_residual": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_debug_sweep_schedule": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
    "psst_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "psst_kernel_stats": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "psst_kernel_names": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "psst_ctx_destroy": (None, [C.c_void_p]),
    "psst_copy_calibration": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "psst_last_error": (C.c_char_p, [C.c_void_p]),
    "psst_case_open": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "psst_wedge_write_msh": (C.c_int, [C.POINTER(WedgeParams), C.c_char_p]),
    "psst_case_open_wedge": (C.c_int, [C.POINTER(WedgeParams), C.c_char_p, C.POINTER(C.c_void_p)]),
    "psst_case_mesh": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc)]),
    "psst_case_from_mesh": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.POINTER(Zone), C.c_char_p,
                                      C.POINTER(C.c_void_p)]),
    "psst_case_settings": (C.c_int, [C.c_void_p, C.POINTER(Settings)]),
    "psst_case_set_settings": (C.c_int, [C.c_void_p, C.POINTER(Settings)]),
    "psst_case_initial_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "psst_case_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "psst_case_raw_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "psst_case_raw": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int8), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_char_p,
                                C.c_int32]),
    "psst_case_write_sav": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_char_p]),
    "psst_case_write_dat": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_char_p]),
    "psst_case_write_cas": (C.c_int, [C.c_void_p, C.c_char_p]),
    "psst_case_solve": (C.c_int, [C.c_void_p]),
    "psst_case_close": (None, [C.c_void_p]),
    "psst_case_last_error": (C.c_char_p, []),
    "psst_set_precision": (C.c_int, [C.c_void_p, C.c_int32]),
    "psst_autotune": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Settings), C.POINTER(C.c_double), C.c_int32, C.c_int32,
                                C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32)]),
    "psst_batch_create": (C.c_int, [C.c_int32, C.POINTER(MeshDesc), C.POINTER(Settings), C.POINTER(C.c_void_p)]),
    "psst_batch_set_state": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "psst_batch_get_state": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "psst_batch_set_residual_reference": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "psst_batch_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "psst_batch_set_precision": (C.c_int, [C.c_void_p, C.c_int32]),
    "psst_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "psst_batch_set_enabled": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "psst_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "psst_batch_context": (C.c_void_p, [C.c_void_p]),
    "psst_batch_destroy": (None, [C.c_void_p]),
    "psst_batch_last_error": (C.c_char_p, [C.c_void_p]),
    # warm-start MLP (include/psst_mlp.h)
    "psst_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "psst_mlp_param_count": (C.c_int64, [C.c_void_p]),
    "psst_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "psst_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "psst_mlp_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)]),
    "psst_mlp_loss_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                                     C.POINTER(C.c_double), C.POINTER(C.c_float)]),
    "psst_mlp_set_data": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64]),
    "psst_mlp_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_uint64,
                                 C.POINTER(C.c_float)]),
    "psst_mlp_train_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "psst_mlp_destroy": (None, [C.c_void_p]),
    "psst_mlp_last_error": (C.c_char_p, [C.c_void_p]),
}


def exported_symbols():
    """Names of the C-ABI entry points the headers (psst_gpu.h, psst_mlp.h) declare."""
    return sorted(_SIGS)


def lib():
    """Loads ``libpsst.so`` once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libpsst.so is not built (run `python -m supersonicflowsolver_amd.build` or "
                "`python -c 'import __graft_entry__; __graft_entry__.build()'`)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class PsstError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status
