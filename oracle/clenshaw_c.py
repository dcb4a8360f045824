"""ctypes binding of ``oracle/clenshaw.c`` (test infrastructure only)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from .build_c import OUT

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not OUT.exists():
            raise ImportError(f"{OUT} missing: run oracle/build_c.py")
        _lib = C.CDLL(str(OUT))
        _lib.oracle_clenshaw.restype = C.c_int
    return _lib


def apply_poly_block(A, Q: np.ndarray, coeffs, smin: float, smax: float, nthreads: int = 0):
    """FP64 ``Y = p(Â) Q`` on the host cores (column-major, OpenMP over columns)."""
    lib = _load()
    A = sp.csr_matrix(A)
    A.sort_indices()
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    idx = np.ascontiguousarray(A.indices, dtype=np.int64)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    Qf = np.asfortranarray(Q, dtype=np.float64)
    n, b = Qf.shape
    co = np.ascontiguousarray(coeffs, dtype=np.float64)
    Y = np.zeros((n, b), order="F")
    work = np.empty(3 * n * b)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    lib.oracle_clenshaw(C.c_int64(n), C.c_int64(b), P(ptr, C.c_int64), P(idx, C.c_int64),
                        P(val, C.c_double), P(Qf, C.c_double), P(co, C.c_double),
                        C.c_int(len(co) - 1), C.c_double(smin), C.c_double(smax),
                        P(Y, C.c_double), P(work, C.c_double), C.c_int(nthreads))
    return Y


def apply_poly_block_rowmajor(A, Q: np.ndarray, coeffs, smin: float, smax: float,
                              nthreads: int = 0):
    """The same filter with the block stored row-major and OpenMP over rows
    (one matrix pass per step for all columns: the fast form for wide blocks)."""
    lib = _load()
    lib.oracle_clenshaw_rm.restype = C.c_int
    A = sp.csr_matrix(A)
    A.sort_indices()
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    idx = np.ascontiguousarray(A.indices, dtype=np.int64)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    Qc = np.ascontiguousarray(Q, dtype=np.float64)
    n, b = Qc.shape
    co = np.ascontiguousarray(coeffs, dtype=np.float64)
    Y = np.zeros((n, b))
    work = np.empty(3 * n * b)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    lib.oracle_clenshaw_rm(C.c_int64(n), C.c_int64(b), P(ptr, C.c_int64), P(idx, C.c_int64),
                           P(val, C.c_double), P(Qc, C.c_double), P(co, C.c_double),
                           C.c_int(len(co) - 1), C.c_double(smin), C.c_double(smax),
                           P(Y, C.c_double), P(work, C.c_double), C.c_int(nthreads))
    return Y
