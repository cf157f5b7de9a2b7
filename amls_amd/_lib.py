"""ctypes binding of libamls_hip.so (declared in include/amls_hip.h).

The library is built in-tree by `amls_amd.build`; importing this module does
not build anything and raises if the library is missing, so no caller can
silently fall back to another implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libamls_hip.so"

AMLS_OK = 0
STATUS_NAMES = {0: "AMLS_OK", 1: "AMLS_E_ARG", 2: "AMLS_E_NOT_SPD", 3: "AMLS_E_NOCONV",
                4: "AMLS_E_HIP", 5: "AMLS_E_OOM", 6: "AMLS_E_STATE", 7: "AMLS_E_INTERNAL"}

# every symbol declared in include/amls_hip.h
EXPORTS = [
    "amls_options_default", "amls_hip_create", "amls_hip_destroy", "amls_hip_last_error",
    "amls_hip_set_problem", "amls_hip_tree_info", "amls_hip_tree_get", "amls_hip_ge",
    "amls_hip_spectral_prepare", "amls_hip_spectral", "amls_hip_project",
    "amls_hip_reduced_solve", "amls_hip_backtransform", "amls_hip_subspace_iter",
    "amls_hip_solve", "amls_hip_solve_numeric", "amls_hip_get_node_factor", "amls_hip_get_local",
    "amls_hip_get_reduced", "amls_hip_gemm", "amls_hip_ldlt_solve", "amls_hip_syevx_top",
    "amls_hip_profile", "amls_hip_kernel_stats", "amls_hip_adjoint",
    "amls_hip_set_problem_batch", "amls_hip_batch_info", "amls_hip_solve_batch",
    "amls_hip_autotune", "amls_hip_set_tuning", "amls_hip_set_sil_mode", "amls_hip_get_sil",
    "amls_hip_ldlt_solve_noalloc", "amls_hip_gemm_gather", "amls_hip_ldlt_factor",
    "amls_hip_ldlt_factor_solve", "amls_hip_ldlt_factor_get", "amls_hip_ldlt_factor_free",
    "amls_hip_dense_eig", "amls_hip_set_rr_mode", "amls_hip_get_rr_info",
    "amls_hip_setup_timing", "amls_hip_trace", "amls_hip_set_partition",
    "amls_hip_store_open", "amls_hip_store_close", "amls_hip_store_last_error",
    "amls_hip_store_append", "amls_hip_store_read_sum", "amls_hip_store_shape",
    "amls_hip_store_remove", "amls_hip_store_reset", "amls_hip_store_disk_bytes",
    "amls_hip_spectral_pinned", "amls_hip_truncation_margin",
]


class AmlsOptionsC(C.Structure):
    _fields_ = [
        ("nLevels", C.c_int32),
        ("leafSize", C.c_int64),
        ("thresholdFactor", C.c_double),
        ("fixedKeep", C.c_int64),
        ("nev", C.c_int64),
        ("denseProjectedMax", C.c_int64),
        ("shift", C.c_double),
        ("twoPass", C.c_int32),
        ("numThreads", C.c_int32),
        ("recoverVectors", C.c_int32),
        ("verbosity", C.c_int32),
        ("precision", C.c_int32),
        ("siSteps", C.c_int32),
        ("siVectors", C.c_int64),
        ("silMode", C.c_int32),
    ]


class AmlsError(RuntimeError):
    """Raised for any non-OK status; `.status` holds the code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None

P = C.c_void_p
I32, I64, D = C.c_int32, C.c_int64, C.c_double
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)
PP = C.POINTER(C.c_void_p)      # array of per-problem pointers


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m amls_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "amls_options_default": ([C.POINTER(AmlsOptionsC)], C.c_int),
        "amls_hip_create": ([I32, C.c_uint32, C.POINTER(P)], C.c_int),
        "amls_hip_destroy": ([P], C.c_int),
        "amls_hip_last_error": ([P], C.c_char_p),
        "amls_hip_set_problem": ([P, I64, PI64, PI64, PD, PI64, PI64, PD, PD, I32, I64, D], C.c_int),
        "amls_hip_tree_info": ([P, PI64, PI64, PI64], C.c_int),
        "amls_hip_tree_get": ([P, PI64, PI64, PI64, PI64, PI64, PI64, PI64], C.c_int),
        "amls_hip_ge": ([P, I32], C.c_int),
        "amls_hip_spectral_prepare": ([P], C.c_int),
        "amls_hip_spectral": ([P, I32, D, PI64], C.c_int),
        "amls_hip_project": ([P, I32, PI64], C.c_int),
        "amls_hip_reduced_solve": ([P, I64, I32, PD], C.c_int),
        "amls_hip_backtransform": ([P, PD], C.c_int),
        "amls_hip_subspace_iter": ([P, I32, I64, I64, PD, PD], C.c_int),
        "amls_hip_solve": ([P, I64, PI64, PI64, PD, PI64, PI64, PD, PD,
                            C.POINTER(AmlsOptionsC), PD, PD, PI64, PD], C.c_int),
        "amls_hip_solve_numeric": ([P, C.POINTER(AmlsOptionsC), PD, PD, PI64, PD], C.c_int),
        "amls_hip_profile": ([P, I32], C.c_int),
        "amls_hip_autotune": ([P, PI64, I64, PI64], C.c_int),
        "amls_hip_set_tuning": ([P, PI64, I64], C.c_int),
        "amls_hip_adjoint": ([P, I64, I32, PI64, PD, PD, I64, PD, PD, PD, PD], C.c_int),
        "amls_hip_set_problem_batch": ([P, I64, PI64, PP, PP, PP, PP, PP, PP, PP, I32, I64, D],
                                       C.c_int),
        "amls_hip_batch_info": ([P, PI64, PI64, PI64], C.c_int),
        "amls_hip_solve_batch": ([P, I64, PI64, PP, PP, PP, PP, PP, PP, PP,
                                  C.POINTER(AmlsOptionsC), PD, PD, PI64, PD], C.c_int),
        "amls_hip_kernel_stats": ([P, I32, C.c_char_p, I32, PI64, PD, PD, PD], C.c_int),
        "amls_hip_get_node_factor": ([P, I64, PD, PD], C.c_int),
        "amls_hip_get_local": ([P, I64, PD, PD], C.c_int),
        "amls_hip_get_reduced": ([P, PD, PI64], C.c_int),
        "amls_hip_gemm": ([P, I32, I32, I32, I64, I64, I64, D, PD, I64, PD, I64, D, PD, I64], C.c_int),
        "amls_hip_ldlt_solve": ([P, I64, PD, I64, PD, PD, PD], C.c_int),
        "amls_hip_syevx_top": ([P, I64, PD, I64, PD, PD], C.c_int),
        "amls_hip_set_sil_mode": ([P, I32], C.c_int),
        "amls_hip_ldlt_solve_noalloc": ([P, I64, P, I64, P], C.c_int),
        "amls_hip_gemm_gather": ([P, I32, I32, I32, I64, I64, I64, D, PD, I64, I64, PI32, PI32,
                                  PD, I64, I64, PI32, PI32, D, PD, I64], C.c_int),
        "amls_hip_ldlt_factor": ([P, I64, PI64, PI64, PD, I32, C.POINTER(P)], C.c_int),
        "amls_hip_ldlt_factor_solve": ([P, P, I64, PD, PD], C.c_int),
        "amls_hip_ldlt_factor_get": ([P, P, PD, PI64], C.c_int),
        "amls_hip_ldlt_factor_free": ([P], C.c_int),
        "amls_hip_dense_eig": ([P, I64, PD, PD, I32, D, D, I64, I64, PI64, PD, PD], C.c_int),
        "amls_hip_get_sil": ([P, C.POINTER(C.c_int32), PI64], C.c_int),
        "amls_hip_set_rr_mode": ([P, I32], C.c_int),
        "amls_hip_get_rr_info": ([P, C.POINTER(C.c_int32), PD], C.c_int),
        "amls_hip_setup_timing": ([P, PD], C.c_int),
        "amls_hip_trace": ([P, C.c_char_p, PI64], C.c_int),
        "amls_hip_set_partition": ([P, I32], C.c_int),
        # block pointers as void*: host arrays or device addresses
        "amls_hip_store_open": ([C.c_char_p, I32, I64, C.POINTER(P)], C.c_int),
        "amls_hip_store_close": ([P], C.c_int),
        "amls_hip_store_last_error": ([P], C.c_char_p),
        "amls_hip_store_append": ([P, I64, P, I64, I64, I64], C.c_int),
        "amls_hip_store_read_sum": ([P, I64, P, I64, I64, I64, I32, PI32], C.c_int),
        "amls_hip_store_shape": ([P, I64, PI64, PI64, PI64], C.c_int),
        "amls_hip_store_remove": ([P, I64], C.c_int),
        "amls_hip_store_reset": ([P], C.c_int),
        "amls_hip_store_disk_bytes": ([P, PD], C.c_int),
        "amls_hip_spectral_pinned": ([P, PI64, PI64], C.c_int),
        "amls_hip_truncation_margin": ([P, PD, PI64], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(PD)


def i32ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(PI32)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(PI64)
