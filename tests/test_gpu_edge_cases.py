"""Edge cases against the oracle's reading of the reference: an interval with
no eigenvalue returns empty results (polfed.jl:1156,
chebyshev_subspace_iteration.jl:754), a rank-deficient start block keeps its
first r columns (polfed.jl:561-594), degree 0 runs, and the boundary's
argument checks."""
import math

import numpy as np
import pytest

from oracle import polfed_ref as P
from sparse_matrix_algorithms_amd.lattices import (anderson3d, gaussian_block, lap1d,
                                                   lap1d_exact_eigs)

pytestmark = pytest.mark.gpu


def _gap_window():
    n = 1000
    lex = lap1d_exact_eigs(n)
    k = 500
    g = lex[k + 1] - lex[k]
    return n, lex[k] + 0.25 * g, lex[k + 1] - 0.25 * g


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_polfed_interval_without_eigenvalues(ctx, precision):
    import sparse_matrix_algorithms_amd as sm
    n, emin, emax = _gap_window()
    kw = dict(s=8, smin=0.0, smax=4.0, emin=emin, emax=emax, sigma=0.5 * (emin + emax), mmax=4)
    Q1 = gaussian_block(1, n, 8)
    lam, X, res = sm.polfed(lap1d(n), 10, 200, Q1=Q1, precision=precision, ctx=ctx, **kw)
    lo, _, _ = P.polfed(lap1d(n), 10, 200, Q0=Q1, **kw)
    assert len(lo) == 0
    assert len(lam) == 0 and X.shape == (n, 0) and len(res) == 0


def test_chebyshev_interval_without_eigenvalues(ctx):
    import sparse_matrix_algorithms_amd as sm
    n, emin, emax = _gap_window()
    lam, X, res = sm.chebyshev_poly_real_symm_sparse(lap1d(n), emin, emax, smin=0.0, smax=4.0,
                                                     m0=16, degree=100, maxiter=3, ctx=ctx)
    assert len(lam) == 0 and X.shape == (n, 0) and len(res) == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_polfed_rank_deficient_start_block(ctx, precision):
    """A start block of rank 5 in 8 columns: orth_block_polfed! keeps the first
    5 columns (polfed.jl:561-594), every later block has width 5, and the run
    still converges to the dense eigenvalues, as the oracle does."""
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(10, 16.5, 21)
    n = A.shape[0]
    ev = np.linalg.eigh(A.toarray())[0]
    i0 = np.searchsorted(ev, 0.0) - 10
    emin, emax = 0.5 * (ev[i0 - 1] + ev[i0]), 0.5 * (ev[i0 + 19] + ev[i0 + 20])
    exact = ev[i0:i0 + 20]
    smin, smax = P.gershgorin_bounds(A)
    anorm = max(abs(smin), abs(smax))
    Q1 = gaussian_block(4, n, 8)
    Q1[:, 5] = Q1[:, 0] + 2.0 * Q1[:, 1]
    Q1[:, 6] = Q1[:, 2] - Q1[:, 3]
    Q1[:, 7] = 0.5 * Q1[:, 4]
    kw = dict(s=8, smin=smin, smax=smax, emin=emin, emax=emax, mmax=16)
    lam, X, res, st = sm.polfed(A, 20, 700, Q1=Q1, precision=precision, return_stats=True,
                                ctx=ctx, **kw)
    lo, _, ro = P.polfed(A, 20, 700, Q0=Q1, **kw)
    assert len(lo) == 20 and np.max(np.abs(np.sort(lo) - exact)) <= 1e-10 * anorm
    assert st["basis_dim"] == 5 * st["block_steps"]
    assert len(lam) == 20
    assert np.max(np.abs(np.sort(lam) - exact)) <= 1e-10 * anorm
    assert res.max() <= 1e-10 * anorm


def _check_residual_order_fallback(lam, res, lo, ro, anorm, tol):
    """polfed.jl:1148-1157: when the nev-th smallest residual is not below tol,
    the best pairs come back in residual order, not λ order."""
    assert len(lam) == len(lo) >= 2
    assert np.sort(ro)[min(len(ro), 10) - 1] >= tol     # the fallback branch is taken
    assert np.all(np.diff(ro) >= 0) and np.all(np.diff(res) >= 0)
    assert not np.all(np.diff(lam) >= 0)                 # not λ order
    # residuals differ by >= 6% between neighbours, so the order is unambiguous at
    # the 1e-10·‖A‖ agreement of the two FP64 implementations
    np.testing.assert_allclose(res, ro, rtol=0, atol=1e-10 * anorm)
    np.testing.assert_allclose(lam, lo, rtol=0, atol=1e-10 * anorm)


def test_polfed_degree_zero_returns_residual_order(ctx):
    """Degree 0 (p = c_0): after Q_1 the filtered block is Q_1 itself, so the
    later blocks orthonormalise rounding noise (kept by the relative rank rule
    in both implementations, which noise survives depends on rounding); the
    returned pairs are unconverged and come back in residual order, as many
    as the oracle returns."""
    import sparse_matrix_algorithms_amd as sm
    B = anderson3d(4, 16.5, 3)
    n = B.shape[0]
    smin, smax = P.gershgorin_bounds(B)
    anorm = max(abs(smin), abs(smax))
    Q1 = gaussian_block(2, n, 8)
    kw = dict(s=8, smin=smin, smax=smax, emin=-2.0, emax=2.0, mmax=4, tol=1e-10 * anorm)
    lam, X, res = sm.polfed(B, 5, 0, Q1=Q1, ctx=ctx, **kw)
    lo, _, ro = P.polfed(B, 5, 0, Q0=Q1, **kw)
    assert len(lam) == len(lo) == 5
    assert np.all(np.isfinite(lam)) and np.all(np.isfinite(res))
    assert res[4] >= kw["tol"] and np.all(np.diff(res) >= 0) and np.all(np.diff(ro) >= 0)


def test_polfed_unconverged_run_returns_residual_order_like_oracle(ctx):
    """Three block steps at degree 400 leave all 12 wanted pairs at residual
    0.1-5: the reference's fallback returns the best nev by residual, in
    residual order; the GPU returns the oracle's pairs in the oracle's order."""
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(8, 16.5, 7)
    n = A.shape[0]
    ev = np.linalg.eigh(A.toarray())[0]
    i0 = np.searchsorted(ev, 0.0)
    emin, emax = 0.5 * (ev[i0 - 7] + ev[i0 - 6]), 0.5 * (ev[i0 + 5] + ev[i0 + 6])
    smin, smax = P.gershgorin_bounds(A)
    anorm = max(abs(smin), abs(smax))
    Q1 = gaussian_block(2, n, 8)
    kw = dict(s=8, smin=smin, smax=smax, emin=emin, emax=emax, mmax=3, tol=1e-10 * anorm)
    lam, X, res = sm.polfed(A, 12, 400, Q1=Q1, ctx=ctx, **kw)
    lo, _, ro = P.polfed(A, 12, 400, Q0=Q1, **kw)
    _check_residual_order_fallback(lam, res, lo, ro, anorm, kw["tol"])
    assert len(lam) == 12


def test_polfed_slices_with_an_empty_slice(ctx):
    import sparse_matrix_algorithms_amd as sm
    n, gmin, gmax = _gap_window()
    lex = lap1d_exact_eigs(n)
    i0 = 200
    emin, emax = 0.5 * (lex[i0 - 1] + lex[i0]), 0.5 * (lex[i0 + 19] + lex[i0 + 20])
    slices = [dict(emin=gmin, emax=gmax, nev=6, s=8, degree=200, mmax=4),
              dict(emin=emin, emax=emax, nev=20, s=8, degree=300, mmax=8)]
    outs = sm.polfed_slices(lap1d(n), slices, smin=0.0, smax=4.0, ctx=ctx)
    assert len(outs[0][0]) == 0
    lam, _, res, _ = outs[1]
    assert len(lam) == 20
    assert np.max(np.abs(np.sort(lam) - lex[i0:i0 + 20])) <= 1e-10 * 4


def _upload_raw(ctx, n, ptr, idx, val, base):
    import ctypes as C
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    h = C.c_void_p()
    rc = ctx.lib.smalgo_csr_upload(ctx.h, n, ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                   idx.ctypes.data_as(C.POINTER(C.c_int64)),
                                   val.ctypes.data_as(C.POINTER(C.c_double)), base, C.byref(h))
    return rc, h


def test_csr_upload_rejects_malformed_input(ctx):
    """The boundary checks its CSR arguments and reports SMALGO_ERR_ARG with a
    message, leaving no handle behind."""
    from sparse_matrix_algorithms_amd import _lib as L
    cases = [
        (3, [0, 1, 2, 3], [0, 1, 3], [1.0, 1.0, 1.0], 0, "column index out of range"),
        (3, [0, 2, 1, 3], [0, 1, 2], [1.0, 1.0, 1.0], 0, "nondecreasing"),
        # row 0 claims 5 entries of a 2-entry matrix: rejected before idx/val are read
        (2, [0, 5, 2], [0, 1], [1.0, 1.0], 0, "nondecreasing"),
        (3, [0, 1, 2, 3], [0, 1, 2], [1.0, 1.0, 1.0], 2, "index_base"),
        (3, [1, 2, 3, 4], [0, 2, 3], [1.0, 1.0, 1.0], 1, "column index out of range"),
        (-1, [0], [0], [0.0], 0, "n out of range"),
    ]
    for n, ptr, idx, val, base, msg in cases:
        rc, h = _upload_raw(ctx, n, ptr, idx, val, base)
        assert rc == L.ERR_ARG, (n, ptr, idx, base)
        assert msg in ctx.lib.smalgo_last_error(ctx.h).decode()
        assert not h.value
    # a 1-based CSC of a symmetric matrix (Julia's arrays) is accepted as is
    import ctypes as C
    rc, h = _upload_raw(ctx, 3, [1, 3, 5, 6], [1, 2, 1, 2, 3], [2.0, -1.0, -1.0, 2.0, 5.0], 1)
    assert rc == L.OK and h.value
    s = C.c_int()
    assert ctx.lib.smalgo_mat_is_symmetric(h, C.byref(s)) == L.OK
    assert s.value == 1
    ctx.lib.smalgo_mat_free(h)
    # the same pattern with one value changed is not symmetric
    rc, h = _upload_raw(ctx, 3, [1, 3, 5, 6], [1, 2, 1, 2, 3], [2.0, -1.0, -1.5, 2.0, 5.0], 1)
    assert rc == L.OK
    assert ctx.lib.smalgo_mat_is_symmetric(h, C.byref(s)) == L.OK
    assert s.value == 0
    ctx.lib.smalgo_mat_free(h)


def test_chebyshev_rejects_nonsymmetric_matrix(ctx):
    """chebyshev_subspace_iteration.jl:669: issymmetric(Ain) or ArgumentError."""
    import scipy.sparse as sp
    import sparse_matrix_algorithms_amd as sm
    A = sp.csr_matrix(np.array([[2.0, 1.0, 0.0], [0.0, 2.0, 1.0], [0.0, 0.0, 2.0]]))
    with pytest.raises(ValueError, match="symmetric"):
        sm.chebyshev_poly_real_symm_sparse(A, 1.0, 3.0, m0=2, degree=4, ctx=ctx)


@pytest.mark.parametrize("s", [1, 3, 5, 12])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_polfed_block_sizes_off_the_vector_width(ctx, s, precision):
    """Block sizes that are not a multiple of the 4-element row padding (the
    reference takes any s >= 1, polfed.jl:1128): the same eigenvalues as dense
    eigh and as the oracle from the same start block."""
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(8, 16.5, 40)
    n = A.shape[0]
    ev = np.linalg.eigh(A.toarray())[0]
    i0 = np.searchsorted(ev, 0.0)
    emin, emax = 0.5 * (ev[i0 - 7] + ev[i0 - 6]), 0.5 * (ev[i0 + 5] + ev[i0 + 6])
    exact = ev[(ev >= emin) & (ev <= emax)]
    smin, smax = P.gershgorin_bounds(A)
    anorm = max(abs(smin), abs(smax))
    nev = len(exact)
    K = math.ceil(8 / (0.5 * (emax - emin) / (0.5 * (smax - smin))))
    kw = dict(s=s, smin=smin, smax=smax, emin=emin, emax=emax, mmax=math.ceil(3 * nev / s) + 4,
              tol=1e-10 * anorm)
    Q1 = gaussian_block(2, n, s)
    lam, X, res = sm.polfed(A, nev, K, Q1=Q1, precision=precision, ctx=ctx, **kw)
    lo, _, ro = P.polfed(A, nev, K, Q0=Q1, **kw)
    assert len(lam) == len(lo) == nev == 12, (len(lam), len(lo))
    assert np.max(np.abs(np.sort(lam) - exact)) <= 1e-10 * anorm
    assert np.max(np.abs(np.sort(lam) - np.sort(lo))) <= 1e-10 * anorm
    assert res.max() <= 1e-10 * anorm
    assert np.max(np.abs(X.T @ X - np.eye(nev))) <= 1e-8


def test_chebyshev_m0_beyond_n(ctx):
    """m0 > n: the reference keeps min(n, m0) columns (thin QR of an n x m0
    block, chebyshev_subspace_iteration.jl:733-734), so the subspace is the
    whole space and the window's eigenpairs come out exactly."""
    import sparse_matrix_algorithms_amd as sm
    from oracle import cheb_ref as CR
    n = 40
    A = lap1d(n)
    lex = lap1d_exact_eigs(n)
    emin, emax = 0.5 * (lex[15] + lex[16]), 0.5 * (lex[21] + lex[22])
    kw = dict(smin=0.0, smax=4.0, m0=60, degree=20, maxiter=2)
    lam, X, res = sm.chebyshev_poly_real_symm_sparse(A, emin, emax, ctx=ctx, **kw)
    lo, _, ro = CR.chebyshev_poly_real_symm_sparse(A, emin, emax, **kw)
    ref = lex[(lex >= emin) & (lex <= emax)]
    assert len(lam) == len(lo) == len(ref) == 6
    assert np.max(np.abs(lam - ref)) <= 1e-12 and np.max(np.abs(lo - ref)) <= 1e-12
    assert res.max() <= 1e-10
