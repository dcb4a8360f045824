"""Pins for the oracle's block Clenshaw filter (polfed.jl:464-501): the
forward three-term recurrence, the dense spectral form p(Â)=V p(Θ) Vᵀ, the
Lap1D sine modes in closed form, and the C restatement."""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npc

from oracle import polfed_ref as P
from sparse_matrix_algorithms_amd.lattices import anderson3d, gaussian_block, lap1d


def _forward(A, Q, c, smin, smax):
    cc, d = 0.5 * (smin + smax), 0.5 * (smax - smin)
    Ah = lambda X: (A @ X - cc * X) / d  # noqa: E731
    T0, T1 = Q, Ah(Q)
    Y = c[0] * T0 + (c[1] * T1 if len(c) > 1 else 0)
    for k in range(2, len(c)):
        T0, T1 = T1, 2 * Ah(T1) - T0
        Y = Y + c[k] * T1
    return Y


@pytest.mark.parametrize("K", [0, 1, 2, 5, 60])
def test_clenshaw_equals_forward_recurrence(K):
    A = anderson3d(4, 10.0, 1)
    smin, smax = P.gershgorin_bounds(A)
    Q = gaussian_block(2, A.shape[0], 5)
    c = P.build_coeffs(K, smin, smax, 0.3, -1.0, 2.0, "step")
    Y = P.apply_poly_block(A, Q, c, smin, smax)
    assert np.max(np.abs(Y - _forward(A, Q, c, smin, smax))) <= 1e-12 * max(1, np.abs(Y).max())


def test_clenshaw_equals_spectral_form():
    A = anderson3d(5, 16.5, 8)
    smin, smax = P.gershgorin_bounds(A)
    th, V = np.linalg.eigh(A.toarray())
    x = (th - 0.5 * (smin + smax)) / (0.5 * (smax - smin))
    c = P.build_coeffs(80, smin, smax, 0.0, -0.5, 0.5, "delta")
    Q = gaussian_block(3, A.shape[0], 4)
    Y = P.apply_poly_block(A, Q, c, smin, smax)
    ref = V @ (npc.chebval(x, c)[:, None] * (V.T @ Q))
    assert np.max(np.abs(Y - ref)) <= 1e-11


def test_clenshaw_lap1d_sine_modes():
    n = 200
    A = lap1d(n)
    j = np.array([3, 50, 100, 151])
    i = np.arange(1, n + 1)
    Q = np.sin(np.outer(i, j) * math.pi / (n + 1))
    lam = 2 - 2 * np.cos(j * math.pi / (n + 1))
    c = P.build_coeffs(40, 0.0, 4.0, 1.0, 0.8, 1.6, "smooth")
    Y = P.apply_poly_block(A, Q, c, 0.0, 4.0)
    p = npc.chebval((lam - 2.0) / 2.0, c)
    assert np.max(np.abs(Y - Q * p[None, :])) <= 1e-12


def test_c_restatement_matches_numpy():
    from oracle import clenshaw_c
    A = anderson3d(6, 16.5, 5)
    smin, smax = P.gershgorin_bounds(A)
    Q = gaussian_block(4, A.shape[0], 7)
    c = P.build_coeffs(50, smin, smax, 0.0, -0.4, 0.4, "step")
    Y = clenshaw_c.apply_poly_block(A, Q, c, smin, smax, nthreads=2)
    Yo = P.apply_poly_block(A, Q, c, smin, smax)
    assert np.max(np.abs(Y - Yo)) <= 1e-13 * np.abs(Yo).max()


def test_c_rowmajor_restatement_matches_numpy():
    from oracle import clenshaw_c
    A = anderson3d(6, 16.5, 5)
    smin, smax = P.gershgorin_bounds(A)
    Q = gaussian_block(4, A.shape[0], 7)
    for K in (0, 1, 2, 50):
        c = P.build_coeffs(K, smin, smax, 0.0, -0.4, 0.4, "step")
        Y = clenshaw_c.apply_poly_block_rowmajor(A, Q, c, smin, smax, nthreads=2)
        Yo = P.apply_poly_block(A, Q, c, smin, smax)
        assert np.max(np.abs(Y - Yo)) <= 1e-13 * np.abs(Yo).max(), K
