"""MI355X-native solve-free interior eigensolvers (POLFED, Chebyshev subspace
iteration) behind a C ABI (``include/smalgo.h``, ``libsmalgo.so``).

The Python functions mirror the reference Julia API
(``polfed.jl``, ``chebyshev_subspace_iteration.jl``); all numerical work runs in
hand-written HIP kernels for gfx950.
"""
from .solvers import (block_qr_factor, build_X, orth_against, orth_block_polfed, rayleigh,
                      refine_original_rr, residuals)
from .solvers import (Context, DeviceMatrix, MatrixFreeRealSymOp, SmalgoError, apply_poly_block,
                      build_coeffs, cheb_delta_coeffs, cheb_smooth_window_coeffs,
                      cheb_window_coeffs, chebyshev_poly_real_symm_sparse,
                      chebyshev_poly_real_symm_sparse_mf, default_context, gershgorin_bounds,
                      jackson_kernel, polfed, polfed_mf)
from .solvers import apply_poly_slices, polfed_slices, subspace_rr

__all__ = [
    "Context", "DeviceMatrix", "MatrixFreeRealSymOp", "SmalgoError", "apply_poly_block",
    "build_coeffs", "cheb_delta_coeffs", "cheb_smooth_window_coeffs", "cheb_window_coeffs",
    "chebyshev_poly_real_symm_sparse", "chebyshev_poly_real_symm_sparse_mf", "default_context",
    "gershgorin_bounds", "jackson_kernel", "polfed", "polfed_mf", "block_qr_factor", "build_X",
    "orth_against", "orth_block_polfed", "rayleigh", "refine_original_rr", "residuals",
    "apply_poly_slices", "polfed_slices", "subspace_rr",
]
