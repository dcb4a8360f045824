"""Pins for the residual form of the refinement filter (oracle/refine_ref.py):
the Chebyshev division identity against numpy's Chebyshev evaluation, and the
FP32 residual-form filter against the FP64 oracle filter (apply_poly_block!,
polfed.jl:464-501)."""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npc

from oracle import polfed_ref as P
from oracle import refine_ref as RF
from sparse_matrix_algorithms_amd.lattices import anderson3d


@pytest.mark.parametrize("K", [0, 1, 2, 7, 60, 400])
@pytest.mark.parametrize("a", [-1.0, -0.73, 0.0, 0.31, 1.0])
def test_cheb_divide_identity(K, a):
    rng = np.random.default_rng(K + 7)
    c = rng.standard_normal(K + 1) / (1 + np.arange(K + 1))
    q, pa = RF.cheb_divide(c, a)
    assert len(q) == K
    assert pa == pytest.approx(npc.chebval(a, c), abs=1e-12 * max(1.0, np.abs(c).sum()))
    x = np.linspace(-1, 1, 301)
    lhs = npc.chebval(x, c) - pa
    rhs = (x - a) * (npc.chebval(x, q) if K else 0.0)
    assert np.max(np.abs(lhs - rhs)) <= 1e-11 * max(1.0, np.abs(c).sum())


def test_cheb_divide_linear_closed_form():
    # p = c0 + c1 x + c2 (2x^2 - 1): (p(x) - p(a)) / (x - a) = c1 + 2 c2 (x + a)
    c = np.array([0.3, -1.2, 0.7])
    a = 0.4
    q, pa = RF.cheb_divide(c, a)
    assert np.allclose(q, [c[1] + 2 * c[2] * a, 2 * c[2]], atol=1e-15)
    assert pa == pytest.approx(c[0] + c[1] * a + c[2] * (2 * a * a - 1), abs=1e-15)


def test_residual_form_fp32_matches_fp64_filter():
    """Near-eigenvectors (FP32-sized perturbations): the FP32 residual form
    equals the FP64 filter to ~1e-13, while filtering the vectors themselves
    in FP32 is only good to ~1e-6."""
    A = anderson3d(10, 16.5, 5).tocsr()
    n = A.shape[0]
    ev, V = np.linalg.eigh(A.toarray())
    smin, smax = P.gershgorin_bounds(A)
    nev = 30
    i0 = np.searchsorted(ev, 0.0) - nev // 2
    emin, emax = 0.5 * (ev[i0 - 1] + ev[i0]), 0.5 * (ev[i0 + nev - 1] + ev[i0 + nev])
    K = math.ceil(8 / ((emax - emin) / (smax - smin)))
    co = P.build_coeffs(K, smin, smax, 0.0, emin, emax, "step", True)
    rng = np.random.default_rng(1)
    Z = V[:, i0 - 5:i0 + nev + 5] + 1e-7 * rng.standard_normal((n, nev + 10)) / math.sqrt(n)
    th = np.einsum("ij,ij->j", Z, A @ Z) / np.einsum("ij,ij->j", Z, Z)
    ref = P.apply_poly_block(A, Z, co, smin, smax)
    got = RF.residual_form_filter(A, Z, th, co, smin, smax, np.float32)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-12 * scale
    # the direct FP32 filter of Z for contrast
    c0, d = 0.5 * (smin + smax), 0.5 * (smax - smin)
    direct = RF.clenshaw_percol(
        lambda X: ((A @ X.astype(np.float64) - c0 * X.astype(np.float64)) / d).astype(X.dtype),
        Z, np.repeat(co[:, None], Z.shape[1], axis=1), np.float32).astype(np.float64)
    assert np.abs(direct - ref).max() > 1e-9 * scale
