"""FP64 CPU restatement of ``chebyshev_subspace_iteration.jl`` (oracle; test
infrastructure only).

Both public drivers are restated: the sparse form (``:667-757``, always runs
``maxiter`` iterations) and the ``matvec!`` form (``:771-848``, ``two_hit``
early exit).  Filter, bounds and coefficients are shared with
:mod:`oracle.polfed_ref` because the reference's copies are identical
(``chebyshev_subspace_iteration.jl:376-489,554-635`` vs ``polfed.jl:275-388,464-545``).
"""
from __future__ import annotations

import math
import warnings
from typing import Optional

import numpy as np
import scipy.sparse as sp

from .polfed_ref import (MatrixFreeRealSymOp, apply_A_block, apply_poly_block,
                         cheb_delta_coeffs, cheb_smooth_window_coeffs, cheb_window_coeffs,
                         gershgorin_bounds)


def thin_orth(Q: np.ndarray) -> np.ndarray:
    """Thin Householder QR, keep all ``min(n,m)`` columns (``:223-231``)."""
    Qf, _ = np.linalg.qr(Q, mode="reduced")
    return np.ascontiguousarray(Qf)


def residuals(A, X: np.ndarray, lam: np.ndarray) -> np.ndarray:
    """``res_j = ‖A x_j − λ_j x_j‖₂`` (``:311-344``)."""
    AX = apply_A_block(A, X)
    R = AX - X * lam[None, :]
    return np.sqrt(np.einsum("ij,ij->j", R, R))


def _coeffs(window_type, a, b, sh, degree, jackson):
    """Window dispatch of ``:692-700``; an unknown type is an error here
    (the reference only logs ``@error`` and then fails on an undefined name)."""
    if window_type == "step":
        return cheb_window_coeffs(a, b, degree, jackson)
    if window_type == "delta":
        return cheb_delta_coeffs(sh, degree, jackson)
    if window_type == "smooth":
        return cheb_smooth_window_coeffs(a, b, degree, jackson)
    raise ValueError("Unknown window type, possible are :step, :delta, and :smooth")


def _iterate(A, Q, coeffs, smin_, smax_, emin, emax, maxiter, tol, res_gate, two_hit,
             early_exit, log):
    m0 = Q.shape[1]
    lam = np.zeros(m0)
    X = np.zeros_like(Q)
    res = np.zeros(m0)
    prev_ok = False
    it_done = 0
    for it in range(1, maxiter + 1):
        it_done = it
        Y = apply_poly_block(A, Q, coeffs, smin_, smax_)
        Q = thin_orth(Y)
        R = apply_A_block(A, Q)
        Aq = Q.T @ R
        lam, Xq = np.linalg.eigh(Aq, UPLO="U")
        X = Q @ Xq
        res = residuals(A, X, lam)
        inside = (lam >= emin) & (lam <= emax) & (res < res_gate)
        maxres = float(res[inside].max()) if inside.any() else math.inf
        log.append(dict(it=it, inside=int(inside.sum()), maxres=maxres))
        if early_exit:
            ok = bool(inside.any()) and maxres < tol
            if two_hit:
                if ok and (it == 1 or prev_ok):
                    break
                prev_ok = ok
            elif ok:
                break
    sel = np.nonzero((lam >= emin) & (lam <= emax) & (res < max(res_gate, 10 * tol)))[0]
    n = Q.shape[0]
    if len(sel) == 0:
        return np.zeros(0), np.zeros((n, 0)), np.zeros(0), it_done
    p = np.argsort(lam[sel], kind="stable")
    sel = sel[p]
    return lam[sel], X[:, sel], res[sel], it_done


def chebyshev_poly_real_symm_sparse(Ain, emin: float, emax: float, smin=None, smax=None,
                                    sigma: float = 0.0, m0: int = 200, degree: int = 80,
                                    maxiter: int = 8, tol: float = 1e-10, res_gate: float = 1e-6,
                                    jackson: bool = True, two_hit: bool = True,
                                    window_type: str = "step", Q0: Optional[np.ndarray] = None,
                                    return_info: bool = False):
    """Sparse-matrix driver (``:667-757``): validation, bounds, map, coefficients,
    ``maxiter`` iterations without early exit, selection by
    ``λ∈[emin,emax] & res < max(res_gate, 10·tol)``, sorted by λ."""
    A = sp.csr_matrix(Ain)
    if (A != A.T).nnz != 0:
        raise ValueError("Chebyshev filtered solver requires a real symmetric matrix")
    if not emin < emax:
        raise ValueError("need emin<emax")
    if not m0 > 0:
        raise ValueError("m0 must be positive")
    if not degree >= 0:
        raise ValueError("degree must be nonnegative")
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("A must be square")
    m0 = min(m0, n)
    if smin is None or smax is None:
        smin_, smax_ = gershgorin_bounds(A)
    else:
        smin_, smax_ = smin, smax
    if not (emin >= smin_ and emax <= smax_):
        warnings.warn("target interval not contained in supplied spectral bounds; "
                      "filter quality may degrade")
    c = 0.5 * (smin_ + smax_)
    d = 0.5 * (smax_ - smin_)
    a = (emin - c) / d
    b = (emax - c) / d
    if not (-1.0 <= a < b <= 1.0):
        raise ValueError(f"mapped interval [{a},{b}] not inside [-1,1]; provide better spectral bounds")
    sh = (sigma - c) / d
    coeffs = _coeffs(window_type, a, b, sh, degree, jackson)
    Q = (np.random.default_rng(1).standard_normal((n, m0)) if Q0 is None
         else np.array(Q0[:, :m0], dtype=np.float64))
    log = []
    lam, X, res, its = _iterate(A, Q, coeffs, smin_, smax_, emin, emax, maxiter, tol, res_gate,
                                two_hit, early_exit=False, log=log)
    if return_info:
        return lam, X, res, dict(iterations=its, log=log)
    return lam, X, res


def chebyshev_poly_real_symm_sparse_mf(matvec, n: int, emin: float, emax: float, matblock=None,
                                       smin=None, smax=None, sigma: float = 0.0, m0: int = 200,
                                       degree: int = 80, maxiter: int = 8, tol: float = 1e-10,
                                       res_gate: float = 1e-6, jackson: bool = True,
                                       two_hit: bool = True, window_type: str = "step",
                                       Q0: Optional[np.ndarray] = None, return_info: bool = False):
    """``matvec!`` driver (``:771-848``): bounds are required (the reference's
    Gershgorin call has no matrix-free method, ``:775-777``), ``two_hit``
    early exit (``:827-836``)."""
    Aop = MatrixFreeRealSymOp(n, matvec, matblock)
    m0 = min(m0, n)
    if smin is None or smax is None:
        raise ValueError("matrix-free form requires smin and smax")
    c = 0.5 * (smin + smax)
    d = 0.5 * (smax - smin)
    a = (emin - c) / d
    b = (emax - c) / d
    sh = (sigma - c) / d
    if not (-1.0 <= a < b <= 1.0):
        raise ValueError(f"mapped interval [{a},{b}] not inside [-1,1]")
    coeffs = _coeffs(window_type, a, b, sh, degree, jackson)
    Q = (np.random.default_rng(1).standard_normal((n, m0)) if Q0 is None
         else np.array(Q0[:, :m0], dtype=np.float64))
    log = []
    lam, X, res, its = _iterate(Aop, Q, coeffs, smin, smax, emin, emax, maxiter, tol, res_gate,
                                two_hit, early_exit=True, log=log)
    if return_info:
        return lam, X, res, dict(iterations=its, log=log)
    return lam, X, res
