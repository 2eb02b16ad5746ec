"""ctypes loader for libsolvix_hip.so (the HIP path). There is no fallback:
if the library is missing or fails to load, every entry point raises."""
from __future__ import annotations

import ctypes as C
import os

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsolvix_hip.so")
_lib = None

# (name, restype, argtypes) for every entry point of include/svx.h
_vp = C.c_void_p
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_errp = C.POINTER(_abi.SvxError)

SIGNATURES = [
    ("svx_ctx_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("svx_ctx_destroy", None, [_vp]),
    ("svx_version", C.c_char_p, []),
    ("svx_device_name", C.c_int, [_vp, C.c_char_p, C.c_size_t]),
    ("svx_last_error", C.c_char_p, []),
    ("svx_set_tuning", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
    ("svx_set_operator", C.c_int, [_vp, C.c_int]),
    ("svx_system_operator", C.c_int, [_vp]),
    ("svx_mf_plan_check", C.c_int, [C.c_int, _i64, _dp, _i64, C.POINTER(C.c_uint8), _i64p, _i32p, _i64p, _dp]),
    ("svx_blk_plan_check", C.c_int, [C.c_int, _i64, _dp, _i64, _i64p, _i32p, C.c_int32, _i64p, _dp]),
    ("svx_model_reaction_count", _i64, [C.POINTER(_abi.SvxModelDesc)]),
    ("svx_solve_model", C.c_int, [_vp, C.POINTER(_abi.SvxModelDesc), C.c_int, C.c_int,
                                  C.POINTER(_abi.SvxSolution), _errp]),
    ("svx_mesh_create", C.c_int, [_vp, C.c_int, _i64, _dp, _i64, C.POINTER(C.c_uint8), _i64p, _i32p,
                                  C.POINTER(_vp), _errp]),
    ("svx_mesh_destroy", None, [_vp]),
    ("svx_mesh_rows", _i64, [_vp]),
    ("svx_mesh_nnz", _i64, [_vp]),
    ("svx_mesh_pattern", C.c_int, [_vp, _i32p, _i32p]),
    ("svx_assemble", C.c_int, [_vp, _dp, _dp, C.c_double, _i64, _i64p, _dp, C.POINTER(_vp), _errp]),
    ("svx_system_destroy", None, [_vp]),
    ("svx_system_values", C.c_int, [_vp, _dp, _dp]),
    ("svx_system_set_rhs", C.c_int, [_vp, _dp]),
    ("svx_system_solution", C.c_int, [_vp, _dp]),
    ("svx_apply_dirichlet", C.c_int, [_vp, _i64, _i64p]),
    ("svx_solve", C.c_int, [_vp, C.c_double, _i64, C.c_int, _dp, _i64p, _dp, _errp]),
    ("svx_system_stats", C.c_int, [_vp, C.POINTER(_abi.SvxSolveStats)]),
    ("svx_recover", C.c_int, [_vp, _dp, _dp, _i64, _i64p, _dp, _dp]),
    ("svx_csr_create", C.c_int, [_vp, _i64, _i32p, _i32p, _dp, C.POINTER(_vp)]),
    ("svx_csr_destroy", None, [_vp]),
    ("svx_csr_spmv", C.c_int, [_vp, _dp, _dp, C.c_int]),
    ("svx_csr_pcg", C.c_int, [_vp, _dp, C.c_double, _i64, C.c_int, _dp, _i64p, _dp, _errp]),
    ("svx_csr_pcg_noalloc", C.c_int, [_vp, _dp, C.c_double, _i64, _dp, _i64p, _dp, _errp]),
    ("svx_system_time_spmv", C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    ("svx_system_time_phases", C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    ("svx_time_spmv_rotation", C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, _dp, _dp]),
    ("svx_batch_create", C.c_int, [_vp, C.POINTER(_vp)]),
    ("svx_batch_destroy", None, [_vp]),
    ("svx_batch_add", C.c_int, [_vp, _vp, C.c_double, _i64, _i64p]),
    ("svx_batch_size", _i64, [_vp]),
    ("svx_batch_solve", C.c_int, [_vp, C.c_int, _errp]),
    ("svx_batch_result", C.c_int, [_vp, _i64, _i64p, _dp, _i32p]),
    ("svx_batch_time_spmv", C.c_int, [_vp, C.c_int, _dp, _dp]),
    ("svx_batch_prepare", C.c_int, [_vp, C.c_int]),
    ("svx_system_prepare_solver", C.c_int, [_vp, C.c_int]),
    ("svx_batch_autotune", C.c_int, [_vp, C.c_int, C.c_int, _i32p, _i32p, _i32p, _dp]),
    ("svx_set_tile_order", C.c_int, [_vp, C.c_int]),
    ("svx_set_load_hint", C.c_int, [_vp, C.c_int]),
    ("svx_set_step_kernel_tiles", C.c_int, [_vp, C.c_int32]),
    ("svx_set_multi_group", C.c_int, [_vp, C.c_int32]),
    ("svx_set_batch_streams", C.c_int, [_vp, C.c_int32]),
    ("svx_multi_spmv_timing", C.c_int, [_vp, C.c_int32, C.c_int32, _dp, _dp]),
    ("svx_server_create", C.c_int, [_vp, C.POINTER(_abi.SvxServerConfig), C.POINTER(_vp)]),
    ("svx_server_destroy", None, [_vp]),
    ("svx_server_submit", C.c_int, [_vp, C.POINTER(_abi.SvxModelDesc), _i64p]),
    ("svx_server_submit_deck", C.c_int, [_vp, C.c_char_p, C.c_size_t, _i64p]),
    ("svx_deck_parse", C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(_vp), _errp]),
    ("svx_deck_model", C.POINTER(_abi.SvxModelDesc), [_vp]),
    ("svx_deck_destroy", None, [_vp]),
    ("svx_server_wait", C.c_int, [_vp, _i64, C.c_double, C.POINTER(_abi.SvxSolution),
                                  C.POINTER(_abi.SvxRequestInfo), _errp]),
    ("svx_server_stats", C.c_int, [_vp, C.POINTER(_abi.SvxServerSummary)]),
    ("svx_solve_multi", C.c_int, [_vp, C.c_int32, _dp, C.c_double, _i64, C.c_int, _dp, _i64p, _dp, _i32p, _dp,
                                  _errp]),
]


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m solvix_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            if not hasattr(L, name):
                continue
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names of every function declared in include/svx.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "svx.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(svx_\w+)\s*\(", text, flags=re.M)))
