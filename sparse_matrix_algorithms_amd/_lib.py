"""ctypes binding of ``libsmalgo.so`` (the C ABI declared in ``include/smalgo.h``).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsmalgo.so"

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_CONFIG = -3
ERR_INTERNAL = -4
ERR_CALLBACK = -5

WINDOWS = {"step": 0, "delta": 1, "smooth": 2}
PRECISIONS = {"fp64": 0, "fp32": 1, "mixed": 1, "fp32_refine64": 1}

MATBLOCK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_double),
                          C.c_int64, C.POINTER(C.c_double), C.c_int64)


class PolfedParams(C.Structure):
    _fields_ = [
        ("s", C.c_int), ("has_bounds", C.c_int), ("smin", C.c_double), ("smax", C.c_double),
        ("sigma", C.c_double), ("emin", C.c_double), ("emax", C.c_double), ("tol", C.c_double),
        ("overextract", C.c_int), ("refine", C.c_int), ("window_type", C.c_int),
        ("jackson", C.c_int), ("mmax", C.c_int), ("rtol", C.c_double),
        ("reorth_passes", C.c_int), ("extra_keep", C.c_int), ("seed", C.c_uint64),
        ("chunk", C.c_int), ("show_progress", C.c_int), ("precision", C.c_int),
        ("refine_guard", C.c_int), ("refine_degree", C.c_int), ("refine_maxiter", C.c_int),
        ("refine_tol", C.c_double), ("refine_guard_degree", C.c_int),
    ]


class SliceParams(C.Structure):
    _fields_ = [
        ("emin", C.c_double), ("emax", C.c_double), ("sigma", C.c_double), ("nev", C.c_int),
        ("s", C.c_int), ("degree", C.c_int), ("mmax", C.c_int),
    ]


class ChebParams(C.Structure):
    _fields_ = [
        ("has_bounds", C.c_int), ("smin", C.c_double), ("smax", C.c_double),
        ("sigma", C.c_double), ("m0", C.c_int), ("degree", C.c_int), ("maxiter", C.c_int),
        ("tol", C.c_double), ("res_gate", C.c_double), ("jackson", C.c_int), ("debug", C.c_int),
        ("two_hit", C.c_int), ("seed", C.c_uint64), ("window_type", C.c_int),
    ]


# (name, restype, argtypes) of every symbol in include/smalgo.h
_P = C.c_void_p
_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)
_FP = C.POINTER(C.c_float)
SIGNATURES = [
    ("smalgo_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("smalgo_ctx_destroy", None, [_P]),
    ("smalgo_last_error", C.c_char_p, [_P]),
    ("smalgo_version", C.c_char_p, []),
    ("smalgo_julia_float_string", C.c_int, [C.c_double, C.c_char_p, C.c_int64]),
    ("smalgo_csr_upload", C.c_int, [_P, C.c_int64, _I64P, _I64P, _DP, C.c_int, C.POINTER(_P)]),
    ("smalgo_mat_free", None, [_P]),
    ("smalgo_mat_n", C.c_int64, [_P]),
    ("smalgo_mat_nnz", C.c_int64, [_P]),
    ("smalgo_mat_is_symmetric", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("smalgo_gershgorin_bounds", C.c_int, [_P, _DP, _DP]),
    ("smalgo_jackson_kernel", C.c_int, [C.c_int, _DP]),
    ("smalgo_cheb_delta_coeffs", C.c_int, [C.c_double, C.c_int, C.c_int, _DP]),
    ("smalgo_cheb_window_coeffs", C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, _DP]),
    ("smalgo_cheb_smooth_window_coeffs", C.c_int,
     [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, _DP]),
    ("smalgo_build_coeffs", C.c_int,
     [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _DP]),
    ("smalgo_apply_poly_block", C.c_int,
     [_P, _P, C.c_int64, _DP, C.c_int64, _DP, C.c_int, C.c_double, C.c_double, C.c_int, _DP,
      C.c_int64]),
    ("smalgo_apply_poly_block_noalloc", C.c_int,
     [_P, _P, C.c_int64, _P, C.c_int64, _DP, C.c_int, C.c_double, C.c_double, _P, C.c_int64]),
    ("smalgo_apply_poly_block_mf", C.c_int,
     [_P, C.c_int64, MATBLOCK_FN, _P, C.c_int64, _DP, C.c_int64, _DP, C.c_int, C.c_double,
      C.c_double, _DP, C.c_int64]),
    ("smalgo_orth_against", C.c_int,
     [_P, C.c_int64, C.c_int64, _DP, C.c_int64, C.c_int64, _DP, C.c_int64, C.c_int]),
    ("smalgo_block_qr", C.c_int,
     [_P, C.c_int64, C.c_int64, _DP, C.c_int64, C.c_double, C.c_int, _DP, C.c_int64, _DP,
      C.c_int64, C.POINTER(C.c_int)]),
    ("smalgo_build_X", C.c_int,
     [_P, C.c_int64, C.c_int64, C.c_int64, _DP, C.c_int64, _DP, C.c_int64, _DP, C.c_int64]),
    ("smalgo_rayleigh", C.c_int,
     [_P, _P, MATBLOCK_FN, _P, C.c_int64, C.c_int64, _DP, C.c_int64, _DP, _DP]),
    ("smalgo_residuals", C.c_int,
     [_P, _P, MATBLOCK_FN, _P, C.c_int64, C.c_int64, _DP, C.c_int64, _DP, _DP]),
    ("smalgo_refine_original_rr", C.c_int,
     [_P, _P, MATBLOCK_FN, _P, C.c_int64, C.c_int64, _DP, C.c_int64, C.POINTER(_P)]),
    ("smalgo_kpm_moments", C.c_int,
     [_P, _P, C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_double, _DP]),
    ("smalgo_copy_bench", C.c_int, [_P, C.c_int64, C.c_int, _DP]),
    ("smalgo_mlp_create", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)]),
    ("smalgo_mlp_free", None, [_P]),
    ("smalgo_mlp_num_params", C.c_int64, [_P]),
    ("smalgo_mlp_get_params", C.c_int, [_P, _FP]),
    ("smalgo_mlp_set_params", C.c_int, [_P, _FP]),
    ("smalgo_mlp_predict", C.c_int, [_P, _FP, C.c_int, _FP]),
    ("smalgo_mlp_gradient", C.c_int, [_P, _FP, _FP, C.c_int, _FP, _FP]),
    ("smalgo_mlp_train", C.c_int, [_P, _FP, _FP, C.c_int, C.c_int, C.c_float, _FP, _FP]),
    ("smalgo_polfed_default_params", None, [C.POINTER(PolfedParams)]),
    ("smalgo_polfed", C.c_int,
     [_P, _P, MATBLOCK_FN, _P, C.c_int64, C.c_int, C.c_int, C.POINTER(PolfedParams), _DP,
      C.POINTER(_P)]),
    ("smalgo_subspace_rr", C.c_int,
     [_P, _P, MATBLOCK_FN, _P, C.c_int64, C.c_int64, _DP, C.c_int64, _DP, C.c_int64,
      C.POINTER(_P)]),
    ("smalgo_csr_save", C.c_int,
     [C.c_char_p, C.c_int64, _I64P, _I64P, _DP, C.c_int, C.c_int]),
    ("smalgo_csr_info", C.c_int, [C.c_char_p, _I64P, _I64P, C.POINTER(C.c_int)]),
    ("smalgo_csr_load", C.c_int, [C.c_char_p, _I64P, _I64P, _DP]),
    ("smalgo_csr_upload_file", C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    ("smalgo_eig_save", C.c_int, [C.c_char_p, C.c_int64, _DP, _DP]),
    ("smalgo_eig_load", C.c_int, [C.c_char_p, _I64P, C.POINTER(C.c_int), _DP, _DP]),
    ("smalgo_io_last_error", C.c_char_p, []),
    ("smalgo_polfed_slices", C.c_int,
     [_P, _P, C.c_int, C.POINTER(SliceParams), C.POINTER(PolfedParams), C.c_int,
      C.POINTER(_P)]),
    ("smalgo_apply_poly_slices", C.c_int,
     [_P, _P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _DP, C.c_double, C.c_double,
      C.c_int, C.c_int, C.c_int, C.c_int, _DP, C.c_int64, _DP, C.c_int64, _DP]),
    ("smalgo_cheb_default_params", None, [C.POINTER(ChebParams)]),
    ("smalgo_chebyshev", C.c_int,
     [_P, _P, MATBLOCK_FN, _P, C.c_int64, C.c_double, C.c_double, C.POINTER(ChebParams), _DP,
      C.POINTER(_P)]),
    ("smalgo_result_count", C.c_int64, [_P]),
    ("smalgo_result_n", C.c_int64, [_P]),
    ("smalgo_result_copy", C.c_int, [_P, _DP, _DP, C.c_int64, _DP]),
    ("smalgo_result_stat", C.c_int, [_P, C.c_char_p, _DP]),
    ("smalgo_result_progress_line", C.c_char_p, [_P]),
    ("smalgo_result_free", None, [_P]),
]

_lib = None


def load() -> C.CDLL:
    """Load the HIP library once; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; "
                          "g.build()'` (hipcc, gfx950) first; there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
