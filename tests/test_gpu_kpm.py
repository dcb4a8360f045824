"""KPM moments on the GPU against exact Chebyshev traces of the same probes."""
import numpy as np
import pytest

from sparse_matrix_algorithms_amd.lattices import anderson3d, gaussian_block

pytestmark = pytest.mark.gpu


def _exact_probe_moments(A, nmom, R, seed, smin, smax):
    n = A.shape[0]
    r = gaussian_block(seed, n, R)
    c, d = 0.5 * (smin + smax), 0.5 * (smax - smin)
    T0, T1 = r, (A @ r - c * r) / d
    mu = [np.sum(r * T0), np.sum(r * T1)]
    for _ in range(2, nmom):
        T0, T1 = T1, 2 * (A @ T1 - c * T1) / d - T0
        mu.append(np.sum(r * T1))
    return np.array(mu) / (R * n)


@pytest.mark.parametrize("nmom,R", [(2, 1), (7, 3), (64, 8), (129, 16)])
def test_kpm_moments_match_probe_traces(ctx, nmom, R):
    from sparse_matrix_algorithms_amd import kpm
    A = anderson3d(6, 12.0, 4)
    mu, smin, smax = kpm.kpm_moments(A, nmom, R, seed=11, ctx=ctx)
    ref = _exact_probe_moments(A, nmom, R, 11, smin, smax)
    assert np.max(np.abs(mu - ref)) <= 1e-10


def test_kpm_count_vs_dense(ctx):
    """Stochastic count over a wide interval within 5% of the dense count."""
    from sparse_matrix_algorithms_amd import kpm
    A = anderson3d(12, 16.5, 2)
    ev = np.linalg.eigh(A.toarray())[0]
    mu, smin, smax = kpm.kpm_moments(A, 256, 32, seed=5, ctx=ctx)
    est = kpm.count(mu, A.shape[0], -1.0, 1.0, smin, smax)
    exact = np.count_nonzero((ev >= -1) & (ev <= 1))
    assert abs(est - exact) <= 0.05 * exact, (est, exact)
