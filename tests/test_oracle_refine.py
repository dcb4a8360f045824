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


def _rr(A, Y):
    Q, _ = np.linalg.qr(Y)
    H = Q.T @ (A @ Q)
    w, U = np.linalg.eigh(0.5 * (H + H.T))
    X = Q @ U
    return w, X, np.linalg.norm(A @ X - X * w, axis=0)


@pytest.mark.parametrize("guard_scale", [1.0, 0.1, 0.02, 0.0])
def test_low_degree_guards_keep_the_refined_pairs_exact(guard_scale):
    """The design claim behind the guard-degree default of the mixed path:
    the guards only widen the Rayleigh–Ritz subspace by directions just
    outside the window, so filtering them at a tenth or a fiftieth of the
    candidates' degree (or not at all: degree 0) leaves the refined pairs of the window at
    FP64 accuracy.  Candidates and guards carry FP32-sized noise spread over
    the whole spectrum plus a 1e-4 admixture of their neighbours, as Ritz
    vectors of an FP32-built basis do; the candidates get the full-degree
    residual-form filter in FP32."""
    A = anderson3d(10, 16.5, 11).tocsr()
    n = A.shape[0]
    ev, V = np.linalg.eigh(A.toarray())
    smin, smax = P.gershgorin_bounds(A)
    anorm = max(abs(smin), abs(smax))
    nev = 24
    i0 = np.searchsorted(ev, 0.0) - nev // 2
    emin, emax = 0.5 * (ev[i0 - 1] + ev[i0]), 0.5 * (ev[i0 + nev - 1] + ev[i0 + nev])
    K = math.ceil(8 / ((emax - emin) / (smax - smin)))
    co = P.build_coeffs(K, smin, smax, 0.0, emin, emax, "step", True)
    rng = np.random.default_rng(3)
    g = nev  # guards: the next Ritz vectors outside the window, half on each side
    idx_c = np.arange(i0, i0 + nev)
    idx_g = np.r_[np.arange(i0 - g // 2, i0), np.arange(i0 + nev, i0 + nev + g // 2)]

    def noisy(idx):
        near = V[:, np.clip(idx + 1, 0, n - 1)] - V[:, np.clip(idx - 1, 0, n - 1)]
        Z = V[:, idx] + 1e-4 * near + 1e-7 * rng.standard_normal((n, len(idx))) / math.sqrt(n)
        return Z / np.linalg.norm(Z, axis=0)

    Zc, Zg = noisy(idx_c), noisy(idx_g)
    rq = lambda Z: np.einsum("ij,ij->j", Z, A @ Z)  # noqa: E731
    Yc = RF.residual_form_filter(A, Zc, rq(Zc), co, smin, smax, np.float32)
    Kg = int(round(guard_scale * K))
    cg = P.build_coeffs(Kg, smin, smax, 0.0, emin, emax, "step", True)
    Yg = RF.residual_form_filter(A, Zg, rq(Zg), cg, smin, smax, np.float32)
    w, X, res = _rr(A, np.hstack([Yc, Yg]))
    inside = (w >= emin) & (w <= emax)
    assert inside.sum() == nev
    assert np.max(np.abs(w[inside] - ev[idx_c])) <= 1e-12 * anorm
    assert res[inside].max() <= 1e-10 * anorm
    # controls: without the guards, or without filtering the candidates, the
    # Rayleigh–Ritz pairs stay at the noise level (~1e-8·|A| here)
    _, _, res0 = _rr(A, Yc)
    assert res0.max() > 100 * res[inside].max()
    w1, _, res1 = _rr(A, np.hstack([Zc, Yg]))
    assert res1[(w1 >= emin) & (w1 <= emax)].max() > 100 * res[inside].max()
