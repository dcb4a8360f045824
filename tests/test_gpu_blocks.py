"""GPU parity of the reference's block kernels, one C-ABI call each:
orth_against! (polfed.jl:605-609), block_qr_factor / orth_block_polfed!
(:644-685, :561-594), build_X! (:767-776), rayleigh! (:791-830),
residuals! (chebyshev_subspace_iteration.jl:311-344), refine_original_rr
(polfed.jl:837-880)."""
import numpy as np
import pytest

from oracle import cheb_ref as CR
from oracle import polfed_ref as P
from sparse_matrix_algorithms_amd.lattices import anderson3d, gaussian_block, lap1d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,b,m", [(200, 8, 16), (1000, 64, 64), (3000, 17, 130), (500, 100, 3)])
def test_orth_against_fp64(ctx, n, b, m):
    import sparse_matrix_algorithms_amd as sm
    W = gaussian_block(1, n, b)
    Q = np.linalg.qr(gaussian_block(2, n, m))[0]
    got = sm.orth_against(W, Q, ctx=ctx)
    ref = P.orth_against(W.copy(), Q)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(W))


def test_orth_against_fp32(ctx):
    import sparse_matrix_algorithms_amd as sm
    W = gaussian_block(1, 4000, 64)
    Q = np.linalg.qr(gaussian_block(2, 4000, 128))[0]
    got = sm.orth_against(W, Q, precision="fp32", ctx=ctx)
    ref = P.orth_against(W.copy(), Q)
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(W))


# b <= 64 and 65..128 take the one-workgroup Cholesky + inverse kernel at its
# two thread-grid sizes; b = 130 takes rocSOLVER's potrf + trtri
@pytest.mark.parametrize("n,b", [(100, 1), (500, 8), (2000, 64), (2000, 65), (1500, 128),
                                 (3000, 130)])
def test_block_qr_full_rank(ctx, n, b):
    import sparse_matrix_algorithms_amd as sm
    W = gaussian_block(4, n, b) @ np.diag(np.linspace(1, 50, b))
    Q, R = sm.block_qr_factor(W, ctx=ctx)
    assert Q.shape == (n, b)
    assert np.max(np.abs(Q.T @ Q - np.eye(b))) <= 1e-13
    assert np.max(np.abs(Q @ R - W)) <= 1e-12 * np.max(np.abs(W))
    assert np.allclose(np.tril(R, -1), 0)
    # same R as Householder up to column signs of Q (rows of R)
    Qo, Ro = P.block_qr_factor(W.copy())
    sgn = np.sign(np.diag(R)) * np.sign(np.diag(Ro))
    assert np.max(np.abs(sgn[:, None] * R - Ro)) <= 1e-10 * np.max(np.abs(Ro))


def test_block_qr_rank_rule(ctx):
    """Rank by count of |R_ii| > rtol·max (polfed.jl:679-684): a block whose
    last two columns repeat earlier ones keeps 6 of 8 directions."""
    import sparse_matrix_algorithms_amd as sm
    W = gaussian_block(7, 400, 8)
    W[:, 6] = W[:, 0]
    W[:, 7] = 2 * W[:, 1] - W[:, 2]
    Q, R = sm.block_qr_factor(W, rtol=1e-6, ctx=ctx)
    Qo, Ro = P.block_qr_factor(W.copy(), rtol=1e-6)
    assert Q.shape[1] == Qo.shape[1] == 6
    assert np.max(np.abs(Q.T @ Q - np.eye(6))) <= 1e-10


def test_block_qr_zero_block(ctx):
    import sparse_matrix_algorithms_amd as sm
    Q, R = sm.block_qr_factor(np.zeros((50, 4)), ctx=ctx)
    assert Q.shape == (50, 0) and R.shape == (0, 4)


def test_block_qr_fp32(ctx):
    import sparse_matrix_algorithms_amd as sm
    W = gaussian_block(4, 5000, 64)
    Q, R = sm.block_qr_factor(W, precision="fp32", ctx=ctx)
    assert np.max(np.abs(Q.T @ Q - np.eye(64))) <= 1e-5


@pytest.mark.parametrize("n,M,k", [(300, 24, 5), (2000, 200, 70)])
def test_build_X(ctx, n, M, k):
    import sparse_matrix_algorithms_amd as sm
    V = gaussian_block(1, n, M)
    U = gaussian_block(2, M, k)
    X = sm.build_X(V, U, ctx=ctx)
    Qh = [V[:, i:i + 8] for i in range(0, M, 8)]
    ref = P.build_X(Qh, U)
    assert np.max(np.abs(X - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_rayleigh_and_residuals(ctx):
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(7, 10.0, 5)
    X = np.linalg.qr(gaussian_block(3, A.shape[0], 20))[0]
    lam, res = sm.rayleigh(A, X, ctx=ctx)
    lo, ro = P.rayleigh(A, X)
    assert np.max(np.abs(lam - lo)) <= 1e-12
    assert np.max(np.abs(res - ro)) <= 1e-12
    r2 = sm.residuals(A, X, lam + 0.5, ctx=ctx)
    assert np.max(np.abs(r2 - CR.residuals(A, X, lam + 0.5))) <= 1e-12


def test_refine_original_rr(ctx):
    import sparse_matrix_algorithms_amd as sm
    A = lap1d(400)
    ev, U = np.linalg.eigh(A.toarray())
    X = U[:, 100:110] + 1e-4 * gaussian_block(1, 400, 10)
    lam, Y, res = sm.refine_original_rr(A, X, ctx=ctx)
    lo, Yo, ro = P.refine_original_rr(A, X)
    assert np.max(np.abs(lam - lo)) <= 1e-12
    assert np.max(np.abs(res - ro)) <= 1e-10
    assert np.max(np.abs(np.abs(np.sum(Y * Yo, axis=0)) - 1)) <= 1e-10
