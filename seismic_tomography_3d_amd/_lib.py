"""ctypes binding of the C ABI in include/st_api.h (libst_hip.so).

The library is the product path: there is no CPU fallback.  If it is
missing, loading raises, naming the build command.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libst_hip.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_long_p = ctypes.POINTER(ctypes.c_long)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)

ST_STATUS = {
    0: "ST_OK", 1: "ST_ERR_ARG", 2: "ST_ERR_HIP", 3: "ST_ERR_NO_MODEL",
    4: "ST_ERR_DIVERGED", 5: "ST_ERR_NOT_UPPER_BOUND", 6: "ST_ERR_CAPACITY",
    7: "ST_ERR_NO_SURROGATE", 8: "ST_ERR_STATE",
}


class FimOpts(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("batch", ctypes.c_int),
                ("max_inner", ctypes.c_int), ("eps", ctypes.c_float),
                ("warm", ctypes.c_int), ("delta", ctypes.c_float),
                ("judge", ctypes.c_int), ("grid", ctypes.c_int),
                ("rounds_per_sync", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("band", ctypes.c_float), ("cold_radius", ctypes.c_float),
                ("warm_values", ctypes.c_int), ("sweep", ctypes.c_int),
                ("order", ctypes.c_int)]


class FimStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_longlong), ("visits", ctypes.c_longlong),
                ("inner_iters", ctypes.c_longlong), ("fallbacks", ctypes.c_longlong),
                ("batches", ctypes.c_longlong), ("deferred", ctypes.c_longlong),
                ("judge_max", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("launches", ctypes.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("st_fim_default_opts", None, [ctypes.POINTER(FimOpts)]),
    ("st_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("st_ctx_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("st_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("st_set_model", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("st_get_model", ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    ("st_fim_solve", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                    ctypes.c_int, c_float_p, c_float_p,
                                    ctypes.POINTER(FimOpts), ctypes.POINTER(FimStats)]),
    ("st_fim_solve_noalloc", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                            ctypes.c_int, c_float_p, ctypes.POINTER(FimOpts),
                                            ctypes.POINTER(FimStats)]),
    ("st_fim_solve_warm", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                         ctypes.c_int, c_float_p, c_float_p, c_float_p,
                                         ctypes.POINTER(FimOpts), ctypes.POINTER(FimStats)]),
    ("st_fsm_solve", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                    ctypes.c_int, c_float_p, c_float_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]),
    ("st_podvin_solve", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                       ctypes.c_int, c_float_p, c_float_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]),
    ("st_picks_volume", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, c_float_p, ctypes.c_int, c_float_p]),
    ("st_tomo_set_grid", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_int]),
    ("st_tomo_dims", ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p, c_int_p]),
    ("st_forward", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                  ctypes.c_int, c_float_p, c_float_p, ctypes.POINTER(FimOpts),
                                  ctypes.POINTER(FimStats)]),
    ("st_tomo_set_eikonal", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("st_forward_times", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                        ctypes.c_int, c_float_p, c_float_p, c_float_p,
                                        ctypes.POINTER(FimOpts), ctypes.POINTER(FimStats)]),
    ("st_csr_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("st_raytrace_volume", ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, c_float_p,
                                          ctypes.c_int, c_float_p]),
    ("st_csr_size", ctypes.c_int, [ctypes.c_void_p, c_ll_p, c_ll_p]),
    ("st_csr_copy", ctypes.c_int, [ctypes.c_void_p, c_ll_p, c_int_p, c_float_p]),
    ("st_residual_norm", ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]),
    ("st_optimize", ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_float, ctypes.c_int, c_float_p]),
    ("st_cgls_csr", ctypes.c_int, [ctypes.c_void_p, c_ll_p, c_int_p, c_float_p, ctypes.c_int,
                                   ctypes.c_int, c_float_p, ctypes.c_int, c_float_p]),
    ("st_optimize_update", ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                          ctypes.c_int, c_float_p, c_float_p, c_float_p]),
    ("st_model_update", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, c_float_p, c_float_p]),
    ("st_surrogate_init", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint]),
    ("st_surrogate_nparams", ctypes.c_int, [ctypes.c_void_p, c_ll_p]),
    ("st_surrogate_params", ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p]),
    ("st_surrogate_eval", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                         c_float_p, c_float_p, ctypes.POINTER(ctypes.c_double),
                                         c_float_p]),
    ("st_surrogate_make_data", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_float, ctypes.c_uint, ctypes.c_float]),
    ("st_surrogate_loss", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("st_surrogate_train", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_int, ctypes.c_uint, c_float_p]),
    ("st_surrogate_save", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("st_surrogate_load", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("st_surrogate_predict", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p]),
    ("st_nn_gemm", ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, c_float_p, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int]),
    ("st_acoustic_dampers", ctypes.c_int, [c_float_p, c_float_p, c_float_p, ctypes.c_float,
                                           ctypes.c_int]),
    ("st_ricker", ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                 c_float_p]),
    ("st_acoustic_set_model", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                             ctypes.c_float, ctypes.c_float]),
    ("st_acoustic_shots", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                         ctypes.c_int, c_float_p, ctypes.c_int, ctypes.c_int,
                                         c_float_p, c_float_p, ctypes.POINTER(ctypes.c_double)]),
    ("st_acoustic_picks", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, c_float_p,
                                         ctypes.c_int, c_float_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                         c_float_p, ctypes.POINTER(ctypes.c_double)]),
    ("st_auto_pick", ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, c_float_p]),
]

_lib = None


def load():
    """Load libst_hip.so (built by `python -m seismic_tomography_3d_amd.build`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m seismic_tomography_3d_amd.build`"
                           " (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_names():
    return [s[0] for s in SIGNATURES]
