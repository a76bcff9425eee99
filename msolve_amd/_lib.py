"""ctypes binding of include/msolve_amd.h (msolve_amd/_lib/libmsolve_amd.so).

There is no fallback: if the HIP library is missing the import of any
compute entry point raises, and every call that needs a GPU returns
MSA_ERR_NO_DEVICE, which is raised as MsaError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmsolve_amd.so")

MSA_OK, MSA_NOT_CONVERGED, MSA_BREAKDOWN, MSA_NEG_DETJ, MSA_NAN_INPUT = 0, 1, 2, 3, 4
MSA_ERR_ARG, MSA_ERR_HIP, MSA_ERR_NO_DEVICE, MSA_ERR_OOM, MSA_ERR_STATE = -1, -2, -3, -4, -5
MSA_FP64, MSA_MIXED = 0, 1
MSA_REFINE_RESTART, MSA_REFINE_RELIABLE = 0, 1
MSA_SOLVER_PCG, MSA_SOLVER_DIRECT = 0, 1
MSA_PRECOND_JACOBI, MSA_PRECOND_MULTIGRID, MSA_PRECOND_FIRST_SAMPLE = 0, 1, 2
TUNE_CANDIDATES = 8
STATUS_NAMES = {0: "OK", 1: "NOT_CONVERGED", 2: "BREAKDOWN", 3: "NEG_DETJ", 4: "NAN_INPUT"}

# every symbol include/msolve_amd.h declares
EXPORTS = [
    "msa_last_error", "msa_version", "CALCH8GAUSSMATRICES", "CALCH8K",
    "CALCH8STRAINS", "CALCH8FORCES", "CALCH8MLUMPED", "msa_batch_element_results",
    "msa_ctx_create", "msa_ctx_destroy", "msa_ctx_stream", "msa_ctx_synchronize",
    "msa_enumerate_dofs", "msa_mesh_create", "msa_mesh_info", "msa_mesh_destroy",
    "msa_ke_packed", "msa_ke_packed_noalloc", "msa_batch_create", "msa_batch_destroy",
    "msa_batch_bytes",
    "msa_batch_set_modulus", "msa_batch_set_kl", "msa_batch_set_xi", "msa_batch_eval_kl",
    "msa_batch_set_rhs", "msa_batch_assemble", "msa_batch_solve", "msa_batch_run",
    "msa_batch_get_solution", "msa_batch_get_stats", "msa_batch_get_modulus",
    "msa_batch_get_matrix_csr", "msa_batch_set_matrix_csr", "msa_batch_prepare_preconditioner",
    "msa_batch_set_zem", "msa_batch_get_zem",
    "msa_batch_set_persistent", "msa_batch_persistent_info",
    "msa_batch_set_smoothing_interval", "msa_batch_get_smoothing_interval", "msa_batch_mg_lambda",
    "msa_batch_spmv", "msa_batch_set_kernel_timing",
    "msa_batch_get_kernel_timing", "msa_batch_spmv_bytes", "msa_autotune_tile",
    "msa_batch_set_refinement", "msa_batch_set_integration",
    "msa_batch_set_dynamics", "msa_batch_set_probe", "msa_batch_newmark",
    "msa_batch_get_history", "msa_batch_get_dynamic_state",
    "msa_batch_factorize", "msa_batch_direct_solve", "msa_batch_set_solver",
    "msa_batch_direct_info", "msa_batch_set_preconditioner", "msa_batch_mg_info",
    "msa_batch_mg_apply", "msa_batch_graph_info", "msa_batch_layout_info", "msa_batch_set_symmetric_product",
    "msa_batch_set_matrix_free_product", "msa_batch_product_form",
    "msa_autotune_schedule",
    "msa_autotune_candidates", "msa_batch_set_schedule", "msa_batch_set_matrix_cores",
    "msa_batch_time_spmv",
]


class MsaError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"msolve_amd error {rc}: {msg}")
        self.rc = rc


_lib = None

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_llp = C.POINTER(C.c_longlong)


def lib():
    """Load the HIP library (raises FileNotFoundError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m msolve_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.msa_last_error.restype = C.c_char_p
    L.msa_last_error.argtypes = []
    L.msa_version.restype = _i
    L.msa_ctx_stream.restype = _vp
    L.msa_ctx_stream.argtypes = [_vp]
    L.msa_batch_bytes.restype = C.c_longlong
    L.msa_batch_bytes.argtypes = [_vp]
    L.CALCH8GAUSSMATRICES.restype = None
    L.CALCH8GAUSSMATRICES.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.CALCH8K.restype = None
    L.CALCH8K.argtypes = [_ip, _dp, _dp, _dp, _dp]
    L.CALCH8STRAINS.restype = None
    L.CALCH8STRAINS.argtypes = [_ip, _dp, _dp, _dp]
    L.CALCH8FORCES.restype = None
    L.CALCH8FORCES.argtypes = [_ip, _dp, _dp, _dp, _dp]
    L.CALCH8MLUMPED.restype = None
    L.CALCH8MLUMPED.argtypes = [_ip, _dp, _dp, _dp]
    sig = {
        "msa_ctx_create": [_i, C.POINTER(_vp)],
        "msa_ctx_destroy": [_vp],
        "msa_ctx_synchronize": [_vp],
        "msa_enumerate_dofs": [_i, _vp, _ip, _ip],
        "msa_mesh_create": [_vp, _i, _dp, _i, _ip, _ip, _i, C.POINTER(_vp)],
        "msa_mesh_info": [_vp, _llp],
        "msa_mesh_destroy": [_vp],
        "msa_ke_packed": [_vp, _i, _dp, _dp, _d, _dp, _ip],
        "msa_ke_packed_noalloc": [_vp, _i, _dp, _dp, _d, _dp, _ip],
        "msa_batch_create": [_vp, _i, C.POINTER(_vp), _ip, _i, _i, C.POINTER(_vp)],
        "msa_batch_destroy": [_vp],
        "msa_batch_set_modulus": [_vp, _i, _dp],
        "msa_batch_set_kl": [_vp, _i, _i, _i, _dp, _dp, _ip, _dp, _dp, _d, _d],
        "msa_batch_set_xi": [_vp, _i, _dp],
        "msa_batch_eval_kl": [_vp],
        "msa_batch_set_rhs": [_vp, _i, _dp],
        "msa_batch_assemble": [_vp, _d],
        "msa_batch_solve": [_vp, _d, _i, _d, _i, _i],
        "msa_batch_run": [_vp, _d, _d, _i, _d, _i, _i],
        "msa_batch_get_solution": [_vp, _i, _i, _dp],
        "msa_batch_get_stats": [_vp, _i, _ip, _ip, _dp, _ip],
        "msa_batch_get_modulus": [_vp, _i, _i, _dp],
        "msa_batch_element_results": [_vp, _i, _i, _dp, _dp, _dp],
        "msa_batch_get_matrix_csr": [_vp, _i, _i, _llp, _ip, _dp],
        "msa_batch_set_matrix_csr": [_vp, _i, _i, _llp, _ip, _dp],
        "msa_batch_prepare_preconditioner": [_vp],
        "msa_batch_set_zem": [_vp, _d, _i],
        "msa_batch_set_persistent": [_vp, _i],
        "msa_batch_set_smoothing_interval": [_vp, _i, _d],
        "msa_batch_get_smoothing_interval": [_vp, _ip, _dp],
        "msa_batch_mg_lambda": [_vp, _i, _i, _i, _dp],
        "msa_batch_persistent_info": [_vp, _llp],
        "msa_batch_get_zem": [_vp, _i, _i, _i, _ip, _ip, _dp],
        "msa_batch_spmv": [_vp, _i, _i, _dp, _dp],
        "msa_batch_set_kernel_timing": [_vp, _i],
        "msa_batch_get_kernel_timing": [_vp, _dp, _llp],
        "msa_batch_spmv_bytes": [_vp, _dp],
        "msa_autotune_tile": [_vp, _vp, _i, _i, _ip, _dp],
        "msa_batch_set_refinement": [_vp, _i, _i, _d],
        "msa_batch_set_integration": [_vp, _i],
        "msa_batch_set_dynamics": [_vp, _d, _d, _d, _d, _d, _d],
        "msa_batch_set_probe": [_vp, _i, _i],
        "msa_batch_newmark": [_vp, _i, _dp, _i, _d, _i, _i, _i],
        "msa_batch_get_history": [_vp, _i, _i, _dp],
        "msa_batch_get_dynamic_state": [_vp, _i, _i, _dp, _dp, _dp, _llp],
        "msa_batch_factorize": [_vp],
        "msa_batch_direct_solve": [_vp],
        "msa_batch_set_solver": [_vp, _i],
        "msa_batch_direct_info": [_vp, _i, _llp],
        "msa_batch_set_preconditioner": [_vp, _i, _i, _d],
        "msa_batch_mg_info": [_vp, _i, _llp],
        "msa_batch_mg_apply": [_vp, _i, _i, _dp, _dp],
        "msa_batch_graph_info": [_vp, _llp],
        "msa_batch_layout_info": [_vp, _i, _llp],
        "msa_batch_set_symmetric_product": [_vp, _i],
        "msa_batch_set_matrix_free_product": [_vp, _i],
        "msa_batch_product_form": [_vp, _i, _ip],
        "msa_autotune_schedule": [_vp, _i, C.POINTER(_vp), _i, _i, _i, _ip, _ip, _dp],
        "msa_autotune_candidates": [_ip, _ip],
        "msa_batch_set_schedule": [_vp, _i],
        "msa_batch_set_matrix_cores": [_vp, _i],
        "msa_batch_time_spmv": [_vp, _i, _dp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.restype = _i
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != MSA_OK:
        raise MsaError(rc, lib().msa_last_error().decode(errors="replace"))
    return rc


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
