"""The Rayleigh–Ritz tail shared by the reference's subspace iterations
(chebyshev_subspace_iteration.jl:733-748; feast.jl:540-553 and
zoloterov.jl:1698-1706, real symmetric case) through smalgo_subspace_rr,
against NumPy's QR + eigh, and inside a FEAST-style loop whose shifted
solves stay on the host."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import polfed_ref as P
from sparse_matrix_algorithms_amd.lattices import anderson3d, gaussian_block

pytestmark = pytest.mark.gpu


def _np_rr(A, Y):
    Q, _ = np.linalg.qr(Y)
    Aq = Q.T @ (A @ Q)
    w, V = np.linalg.eigh(0.5 * (Aq + Aq.T))
    X = Q @ V
    return w, X, np.linalg.norm(A @ X - X * w, axis=0)


@pytest.mark.parametrize("k", [1, 7, 48])
def test_subspace_rr_matches_numpy(ctx, k):
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(8, 16.5, 2)
    smin, smax = P.gershgorin_bounds(A)
    # a filtered block: columns of very different norms, as after a window filter
    co = P.build_coeffs(60, smin, smax, 0.0, -0.5, 0.5, "step", True)
    Y = P.apply_poly_block(A, gaussian_block(3, A.shape[0], k), co, smin, smax)
    lam, X, res, Q = sm.subspace_rr(A, Y, return_Q=True, ctx=ctx)
    w, Xn, rn = _np_rr(A, Y)
    anorm = max(abs(smin), abs(smax))
    assert np.all(np.diff(lam) >= 0)
    assert np.max(np.abs(lam - w)) <= 1e-12 * anorm
    assert np.max(np.abs(res - rn)) <= 1e-10 * anorm
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) <= 1e-12
    # same span as Y, vectors equal up to sign
    assert np.linalg.norm(Y - Q @ (Q.T @ Y)) <= 1e-10 * np.linalg.norm(Y)
    d = np.abs(np.einsum("ij,ij->j", X, Xn))
    assert np.all(np.abs(d - 1) <= 1e-8)


def test_feast_style_loop_with_host_solves(ctx):
    """Contour-integration subspace iteration (the structure of feast.jl's
    real-symmetric loop): shifted sparse LU solves on the host, the
    Rayleigh–Ritz tail on the GPU; converges to the dense-eigh eigenvalues."""
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(6, 16.5, 4).tocsc()
    n = A.shape[0]
    ev = np.linalg.eigh(A.toarray())[0]
    emin, emax = -1.0, 1.0
    exact = ev[(ev >= emin) & (ev <= emax)]
    m0 = 2 * len(exact) + 4
    # Gauss-Legendre nodes on the upper half circle around [emin, emax]; for a
    # real symmetric A the lower half contributes the complex conjugate, so the
    # spectral projector is a real part (the scale does not matter here)
    c, r = 0.5 * (emin + emax), 0.5 * (emax - emin)
    x, wgt = np.polynomial.legendre.leggauss(8)
    theta = 0.5 * np.pi * (1.0 - x)
    z = c + r * np.exp(1j * theta)
    wz = wgt * r * np.exp(1j * theta)
    eye = sp.identity(n, format="csc", dtype=complex)
    lus = [spla.splu((z_k * eye - A).tocsc()) for z_k in z]
    Q = gaussian_block(5, n, m0)
    for it in range(4):
        Y = np.zeros((n, m0))
        for lu, w_k in zip(lus, wz):
            Y += np.real(w_k * lu.solve(Q.astype(complex)))
        lam, X, res, Q = sm.subspace_rr(A, Y, return_Q=True, ctx=ctx)
    inside = (lam >= emin) & (lam <= emax) & (res < 1e-8)
    anorm = np.abs(ev).max()
    assert inside.sum() == len(exact)
    assert np.max(np.abs(np.sort(lam[inside]) - exact)) <= 1e-10 * anorm


def test_large_result_download_paths(ctx):
    """Results of 128 MiB and more come back through pinned, pipelined chunks
    and host threads; the copy must equal the plain one column for column, for
    a packed and for a padded leading dimension (n=110,592, 180 columns =
    159 MB, three 64 MiB chunks)."""
    import ctypes as C
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd import _lib as L
    A = anderson3d(48, 16.5, 48)
    n, k = A.shape[0], 180
    dm = sm.DeviceMatrix(A, ctx)
    Y = np.asfortranarray(gaussian_block(9, n, k))
    h = C.c_void_p()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ctx.check(ctx.lib.smalgo_subspace_rr(ctx.h, dm.h, L.MATBLOCK_FN(), None, n, k, dp(Y), n, None,
                                         n, C.byref(h)))
    try:
        assert ctx.lib.smalgo_result_count(h) == k
        lam, res = np.zeros(k), np.zeros(k)
        X = np.zeros((n, k), order="F")
        ctx.check(ctx.lib.smalgo_result_copy(h, dp(lam), dp(X), n, dp(res)))
        ldx = n + 5
        Xp = np.full((ldx, k), np.nan, order="F")
        ctx.check(ctx.lib.smalgo_result_copy(h, None, dp(Xp), ldx, None))
        assert np.array_equal(Xp[:n], X)
        assert np.isnan(Xp[n:]).all()  # padding rows untouched
    finally:
        ctx.lib.smalgo_result_free(h)
        dm.free()
    # the vectors are the Rayleigh–Ritz vectors of span(Y): orthonormal, with
    # the reported residuals
    assert np.max(np.abs(X.T @ X - np.eye(k))) <= 1e-10
    R = A @ X - X * lam[None, :]
    np.testing.assert_allclose(np.linalg.norm(R, axis=0), res, rtol=1e-6, atol=1e-12)
