"""CPU oracle for the solve-free interior eigensolver path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library and the
host-side mirror in ``sparse_matrix_algorithms_amd``) imports, calls or links
anything under ``oracle/``.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg use it, and there only as the checker (or
as the timed CPU baseline), never as the thing measured or shipped.

Modules
-------
``oracle.polfed_ref``
    FP64 NumPy/SciPy restatement of ``polfed.jl`` (coefficients, Gershgorin
    bounds, block Clenshaw filter, truncated QR, block-Lanczos basis build,
    reduced-matrix assembly, extraction, refinement, ``polfed``).
``oracle.cheb_ref``
    FP64 restatement of ``chebyshev_subspace_iteration.jl`` (both the sparse
    and the ``matvec!`` drivers, ``thin_orth!``, ``residuals!``).
``oracle.lap1d``
    The reference's own Lap1D known-answer helpers (closed-form spectrum,
    count windows, classification of results).
``oracle.clenshaw_c``
    ctypes binding to ``oracle/clenshaw.c`` (OpenMP FP64 CSR Clenshaw filter),
    used as the fast CPU leg of the baseline and to cross-check the NumPy
    filter at sizes where NumPy is slow.

Pinning status of every function is listed in ``oracle/PINNING.md``.
"""
