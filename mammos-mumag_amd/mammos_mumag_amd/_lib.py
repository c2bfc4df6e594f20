"""ctypes binding of libmumag_hip.so (declarations in include/mumag_hip.h).

Loading fails loudly: there is no CPU fallback for the hot path.
"""
import ctypes
import pathlib

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "_lib" / "libmumag_hip.so"

MM_OK = 0
MM_E_ARG = -1
MM_E_HIP = -2
MM_E_MESH = -3
MM_E_STATE = -4
MM_W_MAXITER = 1
MM_PREC_FP64 = 0
MM_PREC_MIXED = 1

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_uint8_p = ctypes.POINTER(ctypes.c_uint8)


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32), ("n_elems", ctypes.c_int32),
        ("coords", c_double_p), ("tets", c_int32_p), ("elem_mat", c_int32_p),
        ("n_mats", ctypes.c_int32),
        ("mat_Js", c_double_p), ("mat_A", c_double_p), ("mat_K1", c_double_p), ("mat_u", c_double_p),
        ("dirichlet", c_uint8_p), ("shell", c_uint8_p),
        ("shell_R", ctypes.c_double), ("shell_Rinf", ctypes.c_double),
        ("length_scale", ctypes.c_double),
        ("grid_n", ctypes.c_int32 * 3),
        ("grid_x", c_double_p), ("grid_y", c_double_p), ("grid_z", c_double_p),
        ("grid_mag0", ctypes.c_int32 * 3), ("grid_magn", ctypes.c_int32),
    ]


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("drop_tol", ctypes.c_double), ("pcg_batch", ctypes.c_int32),
                ("use_structured", ctypes.c_int32), ("inner_tol", ctypes.c_double),
                ("eval_precision", ctypes.c_int32)]


class Info(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_elems", ctypes.c_int64),
                ("n_mag_elems", ctypes.c_int64), ("n_mag_nodes", ctypes.c_int64),
                ("k_nnz", ctypes.c_int64), ("k_sell_entries", ctypes.c_int64),
                ("volume", ctypes.c_double), ("structured", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("device_bytes", ctypes.c_int64)]


class MinParams(ctypes.Structure):
    _fields_ = [("tol_fun", ctypes.c_double), ("tol_u", ctypes.c_double),
                ("precond_iter", ctypes.c_int32), ("hmag_on", ctypes.c_int32),
                ("iter_max", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("ncg_iters", ctypes.c_int64), ("fcalls", ctypes.c_int64),
                ("pcg_iters", ctypes.c_int64), ("pcg_calls", ctypes.c_int64),
                ("tcg_iters", ctypes.c_int64), ("c_applies", ctypes.c_int64),
                ("backtracks", ctypes.c_int64), ("restarts_g", ctypes.c_int64),
                ("restarts_z", ctypes.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class KernelTiming(ctypes.Structure):
    _fields_ = [("spmv_launches", ctypes.c_int64), ("spmv_ms", ctypes.c_double),
                ("spmv_bytes", ctypes.c_int64), ("update_launches", ctypes.c_int64),
                ("update_ms", ctypes.c_double), ("update_bytes", ctypes.c_int64),
                ("mg_smooth_launches", ctypes.c_int64), ("mg_smooth_ms", ctypes.c_double),
                ("mg_smooth_bytes", ctypes.c_int64), ("vcycles", ctypes.c_int64),
                ("ru_passes", ctypes.c_int64), ("corrections", ctypes.c_int64)]


VP = ctypes.c_void_p
D3 = ctypes.c_double * 3

# name -> (argtypes); every entry point returns int
SIGNATURES = {
    "mm_ctx_create": [ctypes.POINTER(MeshDesc), ctypes.POINTER(Options), ctypes.POINTER(VP)],
    "mm_ctx_destroy": [VP],
    "mm_ctx_info": [VP, ctypes.POINTER(Info)],
    "mm_get_meas_host": [VP, c_double_p],
    "mm_get_diag_host": [VP, c_double_p, c_double_p],
    "mm_external_eg": [VP, VP, ctypes.c_double, D3, c_double_p, VP],
    "mm_exani_eg": [VP, VP, c_double_p, VP],
    "mm_hmag_eg": [VP, VP, VP, ctypes.c_double, c_double_p, VP, c_int32_p],
    "mm_total_eg": [VP, VP, ctypes.c_double, D3, ctypes.c_double, ctypes.c_int32, VP, c_double_p,
                    VP, VP, c_int32_p],
    "mm_pcg": [VP, VP, VP, ctypes.c_double, ctypes.c_double, ctypes.c_int32, c_int32_p],
    "mm_precond": [VP, VP, VP, VP, ctypes.c_int32, VP, c_int32_p],
    "mm_update_m": [VP, VP, ctypes.c_double, VP, VP],
    "mm_dot": [VP, VP, VP, ctypes.c_int64, c_double_p],
    "mm_norm_inf": [VP, VP, ctypes.c_int64, c_double_p],
    "mm_dot_noalloc": [ctypes.c_int32, VP, VP, ctypes.c_int64, c_double_p],
    "mm_compute_mh": [VP, VP, D3, c_double_p],
    "mm_minimize": [VP, VP, ctypes.c_double, D3, ctypes.POINTER(MinParams), VP, VP, VP,
                    c_double_p, ctypes.POINTER(Stats)],
    "mm_apply_C": [VP, VP, VP],
    "mm_apply_K": [VP, VP, VP],
    "mm_divergence": [VP, VP, VP],
    "mm_gradient": [VP, VP, VP],
    "mm_timing_enable": [VP, ctypes.c_int32],
    "mm_timing_get": [VP, ctypes.POINTER(KernelTiming)],
    "mm_mlp_create": [VP, ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(VP)],
    "mm_mlp_destroy": [VP],
    "mm_mlp_set_params": [VP, c_float_p],
    "mm_mlp_get_params": [VP, c_float_p],
    "mm_mlp_features": [VP, VP, VP],
    "mm_mlp_forward": [VP, ctypes.c_int64, VP, VP],
    "mm_mlp_grads": [VP, ctypes.c_int64, VP, VP, c_float_p, c_double_p],
    "mm_mlp_train_step": [VP, ctypes.c_int64, VP, VP, ctypes.c_float, c_double_p],
    "mm_warm_start": [VP, VP, ctypes.c_double],
    "mm_multigrid": [VP, ctypes.c_int32],
    "mm_mg_apply": [VP, VP, VP],
    "mm_mg_levels": [VP, VP, VP],
    "mm_amg": [VP, ctypes.c_int32],
    "mm_amg_apply": [VP, VP, VP],
    "mm_amg_levels": [VP, VP, VP],
    "mm_warm_start_stats": [VP, ctypes.POINTER(ctypes.c_int64), c_double_p],
    "mm_warm_start_policy": [VP, ctypes.c_int32],
    "mm_secant_start": [VP, ctypes.c_int32],
    "mm_warm_start_report": [VP, VP, VP, VP, VP],
    "mm_lp_setup": [VP],
    "mm_lp_features": [VP, VP, VP],
    "mm_lp_pcg": [VP, VP, ctypes.c_double, VP, VP, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, c_int32_p,
                  c_int32_p],
}

MLP_IN = 32
MLP_PARAMS = 37377
WS_EVERY, WS_MEASURED = 0, 1

_lib = None


class MumagError(RuntimeError):
    pass


def lib_path():
    return _LIB_PATH


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise MumagError(f"{_LIB_PATH} is missing: build it with `python -m mammos_mumag_amd.build` "
                         "(there is no CPU fallback for the HIP hot path)")
    # torch-ROCm ships its own HIP/HSA runtime (sonames libamdhip64.so.7 and
    # libhsa-runtime64.so.1).  Importing torch first makes this library bind to that
    # same runtime instance, so that tensors handed over by data_ptr() and the
    # library's own allocations live in one HSA context.
    import torch  # noqa: F401
    lib = ctypes.CDLL(str(_LIB_PATH))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.mm_last_error.argtypes = []
    lib.mm_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(rc, what=""):
    """Raise on errors; return rc (0 or a warning code) otherwise."""
    if rc < 0:
        msg = load().mm_last_error().decode(errors="replace")
        raise MumagError(f"{what} failed ({rc}): {msg}")
    return rc


def exported_symbols():
    return list(SIGNATURES) + ["mm_last_error"]
