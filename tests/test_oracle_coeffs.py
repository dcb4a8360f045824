"""Pins for the oracle's bounds and filter coefficients (polfed.jl:254-414)
against independent mathematics: Weiße's closed-form Jackson kernel, numerical
quadrature of Chebyshev projections, numpy's Chebyshev evaluation."""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npc
from scipy import integrate

from oracle import polfed_ref as P


def _weisse_jackson(K):
    # Weiße et al., RMP 78, 275 (2006), Eq. (71) with N = K+1 moments
    N = K + 1
    k = np.arange(K + 1)
    q = math.pi / (N + 1)
    return ((N - k + 1) * np.cos(q * k) + np.sin(q * k) / math.tan(q)) / (N + 1)


@pytest.mark.parametrize("K", [1, 2, 5, 32, 400])
def test_jackson_closed_form(K):
    g = P.jackson_kernel(K)
    assert abs(g[0] - 1.0) <= 1e-15
    assert np.max(np.abs(g - _weisse_jackson(K))) <= 1e-14


def test_jackson_degenerate():
    assert list(P.jackson_kernel(0)) == [1.0]
    assert list(P.jackson_kernel(-3)) == [1.0]


def _proj(f, k):
    """(2-δ_k0)/π ∫_{-1}^{1} f(x) T_k(x) / sqrt(1-x²) dx, in θ to remove the weight."""
    val, _ = integrate.quad(lambda t: f(math.cos(t)) * math.cos(k * t), 0, math.pi, limit=400,
                            epsabs=1e-13, epsrel=1e-13)
    return (1.0 if k == 0 else 2.0) / math.pi * val


@pytest.mark.parametrize("a,b", [(-0.3, 0.5), (0.5, -0.3), (-2.0, 0.1), (0.2, 7.0)])
def test_window_coeffs_are_chebyshev_projection(a, b):
    K = 12
    c = P.cheb_window_coeffs(a, b, K, jackson=False)
    lo, hi = max(-1, min(1, min(a, b))), max(-1, min(1, max(a, b)))
    for k in range(K + 1):
        # split at the discontinuities for quad accuracy
        t1, t2 = math.acos(hi), math.acos(lo)
        val, _ = integrate.quad(lambda t: math.cos(k * t), t1, t2, epsabs=1e-14)
        ref = (1.0 if k == 0 else 2.0) / math.pi * val
        assert abs(c[k] - ref) <= 1e-13


@pytest.mark.parametrize("sh", [-0.7, 0.0, 0.35, 1.5])
def test_delta_coeffs_are_chebyshev_values(sh):
    K = 20
    c = P.cheb_delta_coeffs(sh, K, jackson=False)
    x = min(max(sh, -1), 1)
    for k in range(K + 1):
        Tk = npc.chebval(x, [0] * k + [1])
        assert abs(c[k] - (1 if k == 0 else 2) / math.pi * Tk) <= 1e-13


def test_jackson_damping_is_applied():
    K = 30
    assert np.allclose(P.cheb_window_coeffs(-0.2, 0.3, K),
                       P.cheb_window_coeffs(-0.2, 0.3, K, jackson=False) * P.jackson_kernel(K))
    assert np.allclose(P.cheb_delta_coeffs(0.1, K),
                       P.cheb_delta_coeffs(0.1, K, jackson=False) * P.jackson_kernel(K))


def test_smooth_window_matches_quadrature():
    a, b, K = -0.2, 0.4, 10
    c = P.cheb_smooth_window_coeffs(a, b, K, jackson=False)
    f = lambda x: 0.5 * (1 + math.tanh((x - a) / 0.02)) * 0.5 * (1 - math.tanh((x - b) / 0.02))  # noqa
    for k in range(K + 1):
        assert abs(c[k] - _proj(f, k)) <= 1e-6


def test_window_series_approximates_indicator():
    c = P.cheb_window_coeffs(-0.3, 0.4, 2000, jackson=True)
    x = np.array([-0.9, -0.5, -0.2, 0.0, 0.3, 0.6, 0.95])
    y = npc.chebval(x, c)
    want = ((x >= -0.3) & (x <= 0.4)).astype(float)
    assert np.max(np.abs(y - want)) <= 1e-2


def test_build_coeffs_mapping_and_errors():
    smin, smax = -3.0, 5.0
    c0, d = 1.0, 4.0
    np.testing.assert_allclose(P.build_coeffs(9, smin, smax, 0.5, -1.0, 2.0, "step"),
                               P.cheb_window_coeffs((-1 - c0) / d, (2 - c0) / d, 9))
    np.testing.assert_allclose(P.build_coeffs(9, smin, smax, 0.5, -1.0, 2.0, "delta"),
                               P.cheb_delta_coeffs((0.5 - c0) / d, 9))
    with pytest.raises(ValueError, match="Unknown window type"):
        P.build_coeffs(9, smin, smax, 0.5, -1.0, 2.0, "box")


def test_gershgorin_brute_force():
    rng = np.random.default_rng(3)
    M = rng.standard_normal((30, 30)) * (rng.random((30, 30)) < 0.2)
    M = M + M.T + np.diag(rng.standard_normal(30))
    import scipy.sparse as sp
    lo, hi = P.gershgorin_bounds(sp.csr_matrix(M))
    d = np.diag(M)
    r = np.abs(M).sum(axis=1) - np.abs(d)
    assert lo == pytest.approx(np.min(d - r), abs=1e-14)
    assert hi == pytest.approx(np.max(d + r), abs=1e-14)
    ev = np.linalg.eigvalsh(M)
    assert lo <= ev[0] and ev[-1] <= hi
