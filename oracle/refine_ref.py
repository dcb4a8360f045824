"""Test infrastructure only (no product path imports this): NumPy restatement
of the residual form of the refinement filter used by the mixed-precision
POLFED path.  No reference counterpart (the reference is FP64 throughout);
the filter it reproduces is apply_poly_block! (polfed.jl:464-501).

For a Chebyshev series p(x) = Σ c_k T_k(x) and a point a,
    p(x) - p(a) = (x - a) q(x),   q = Σ_{k<K} q_k T_k,
so p(Â) z = p(â) z + q(Â) (Â - â) z for any vector z.  With â the Rayleigh
quotient of z the second term acts on a small residual, and computing it in
FP32 keeps the FP64 accuracy of p(Â) z.
"""
from __future__ import annotations

import numpy as np


def cheb_divide(c: np.ndarray, a: float):
    """Coefficients q (degree K-1) with p(x) - p(a) = (x - a) q(x), and p(a),
    by the backward recurrence of Clenshaw's algorithm at x = a."""
    c = np.asarray(c, dtype=np.float64)
    K = len(c) - 1
    q = np.zeros(K + 2)
    for m in range(K, 1, -1):
        q[m - 1] = 2.0 * c[m] + 2.0 * a * q[m] - q[m + 1]
    if K >= 1:
        q[0] = c[1] + a * q[1] - 0.5 * q[2]
    pa = c[0] + a * q[0] - 0.5 * q[1]
    return q[:max(K, 0)], pa


def clenshaw_percol(apply_hat, R: np.ndarray, Q: np.ndarray, dtype=np.float64) -> np.ndarray:
    """Y[:, i] = Σ_k Q[k, i] T_k(Â) R[:, i] by Clenshaw in `dtype` (per-column
    coefficients Q of shape (K+1, cols)); apply_hat(X) = Â X."""
    R = np.asarray(R, dtype=dtype)
    Q = np.asarray(Q, dtype=dtype)
    K = Q.shape[0] - 1
    b1 = np.zeros_like(R)
    b2 = np.zeros_like(R)
    for k in range(K, 0, -1):
        t = (Q[k] * R + 2 * apply_hat(b1) - b2).astype(dtype)
        b2, b1 = b1, t
    return (Q[0] * R + apply_hat(b1) - b2).astype(dtype)


def residual_form_filter(A, Z: np.ndarray, theta: np.ndarray, c: np.ndarray, smin: float,
                         smax: float, dtype=np.float32) -> np.ndarray:
    """p(Â) Z as p(â_i) z_i + q_i(Â) r̂_i, the q part evaluated in `dtype`."""
    c0 = 0.5 * (smin + smax)
    d = 0.5 * (smax - smin)
    Z = np.asarray(Z, dtype=np.float64)
    a = np.clip((np.asarray(theta) - c0) / d, -1.0, 1.0)
    qs, pas = zip(*(cheb_divide(c, ai) for ai in a))
    Y = Z * np.array(pas)
    if len(c) <= 1:
        return Y
    Qm = np.array(qs).T
    R = (A @ Z - Z * np.asarray(theta)) / d

    def apply_hat(X):
        Xd = X.astype(np.float64)
        return ((A @ Xd - c0 * Xd) / d).astype(X.dtype)

    return Y + clenshaw_percol(apply_hat, R, Qm, dtype).astype(np.float64)
