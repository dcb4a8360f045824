"""Chebyshev filtered subspace iteration on the GPU
(chebyshev_subspace_iteration.jl:667-848) against the FP64 oracle."""
import numpy as np
import pytest

from oracle import cheb_ref as CR
from oracle import polfed_ref as P
from sparse_matrix_algorithms_amd.lattices import (anderson3d, gaussian_block, lap1d,
                                                   lap1d_exact_eigs)

pytestmark = pytest.mark.gpu


def test_cheb_sparse_lap1d_matches_oracle(ctx):
    import sparse_matrix_algorithms_amd as sm
    n = 1000
    A = lap1d(n)
    lex = lap1d_exact_eigs(n)
    emin, emax = 0.5 * (lex[479] + lex[480]), 0.5 * (lex[519] + lex[520])
    Q0 = gaussian_block(3, n, 80)
    kw = dict(smin=0.0, smax=4.0, sigma=0.5 * (emin + emax), m0=80, degree=64, maxiter=6,
              tol=1e-8, window_type="delta")
    lam, X, res = sm.chebyshev_poly_real_symm_sparse(A, emin, emax, Q0=Q0, ctx=ctx, **kw)
    lo, Xo, ro = CR.chebyshev_poly_real_symm_sparse(A, emin, emax, Q0=Q0, **kw)
    assert len(lam) == len(lo) > 0
    assert np.max(np.abs(lam - lo)) <= 1e-10
    ref = lex[(lex >= emin) & (lex <= emax)]
    assert len(lam) == len(ref)
    assert np.max(np.abs(lam - ref)) <= 1e-10


def test_cheb_mf_two_hit_iterations(ctx):
    import sparse_matrix_algorithms_amd as sm
    n = 600
    A = lap1d(n)
    lex = lap1d_exact_eigs(n)
    emin, emax = 0.5 * (lex[289] + lex[290]), 0.5 * (lex[309] + lex[310])
    Q0 = gaussian_block(5, n, 40)

    def matblock(Y, X):
        Y[:] = A @ X

    kw = dict(smin=0.0, smax=4.0, m0=40, degree=150, maxiter=10, tol=1e-9)
    lam, X, res, st = sm.chebyshev_poly_real_symm_sparse_mf(None, n, emin, emax,
                                                            matblock=matblock, Q0=Q0,
                                                            return_stats=True, ctx=ctx, **kw)
    lo, Xo, ro, info = CR.chebyshev_poly_real_symm_sparse_mf(
        lambda y, x: None, n, emin, emax, matblock=lambda Y, Xb: Y.__setitem__(slice(None), A @ Xb),
        Q0=Q0, return_info=True, **kw)
    assert st["iterations"] == info["iterations"] == 6   # two_hit exit before maxiter
    ref = lex[(lex >= emin) & (lex <= emax)]
    assert len(lam) == len(lo) == len(ref)
    assert np.max(np.abs(lam - lo)) <= 1e-10
    assert np.max(np.abs(lam - ref)) <= 1e-10


def test_cheb_anderson_dense(ctx):
    import sparse_matrix_algorithms_amd as sm
    A = anderson3d(8, 16.5, 9)
    ev = np.linalg.eigh(A.toarray())[0]
    emin, emax = -0.6, 0.6
    ref = ev[(ev >= emin) & (ev <= emax)]
    lam, X, res = sm.chebyshev_poly_real_symm_sparse(A, emin, emax, m0=2 * len(ref) + 10,
                                                     degree=200, maxiter=6, ctx=ctx)
    assert len(lam) == len(ref)
    assert np.max(np.abs(lam - ref)) <= 1e-9


def test_cheb_argument_errors(ctx):
    import sparse_matrix_algorithms_amd as sm
    import scipy.sparse as sp
    A = lap1d(50)
    with pytest.raises(ValueError, match="emin<emax"):
        sm.chebyshev_poly_real_symm_sparse(A, 1.0, 0.5, ctx=ctx)
    with pytest.raises(ValueError, match="m0 must be positive"):
        sm.chebyshev_poly_real_symm_sparse(A, 0.5, 1.0, m0=0, ctx=ctx)
    with pytest.raises(ValueError, match="symmetric"):
        sm.chebyshev_poly_real_symm_sparse(sp.csr_matrix(np.triu(A.toarray())), 0.5, 1.0, ctx=ctx)
    with pytest.raises(ValueError, match="not inside"):
        sm.chebyshev_poly_real_symm_sparse(A, 0.5, 9.0, smin=0.0, smax=4.0, ctx=ctx)


def test_cheb_debug_lines(ctx, capfd):
    """debug=true prints the reference's per-iteration lines: the sparse form
    `it=… inside=… maxres=…` (chebyshev_subspace_iteration.jl:748), the matrix-
    free form the same plus r90 = quantile(res[inside], 0.9) and a `converged`
    line at the two-hit exit (:826, :830).  On the last iteration the inside
    pairs are the returned ones (res_gate >= 10 tol), so the fields follow from
    the returned residuals."""
    import re

    import sparse_matrix_algorithms_amd as sm
    n = 600
    A = lap1d(n)
    lex = lap1d_exact_eigs(n)
    emin, emax = 0.5 * (lex[289] + lex[290]), 0.5 * (lex[309] + lex[310])
    Q0 = gaussian_block(5, n, 40)
    kw = dict(smin=0.0, smax=4.0, m0=40, degree=150, maxiter=10, tol=1e-9, debug=True, Q0=Q0)

    def matblock(Y, X):
        Y[:] = A @ X

    capfd.readouterr()
    lam, X, res, st = sm.chebyshev_poly_real_symm_sparse_mf(None, n, emin, emax,
                                                            matblock=matblock,
                                                            return_stats=True, ctx=ctx, **kw)
    out = capfd.readouterr().out.strip().splitlines()
    its = [l for l in out if l.startswith("it=")]
    assert len(its) == st["iterations"] == 6 and out[-1] == "converged"
    pat = re.compile(r"it=(\d+) inside=(\d+) maxres=(\S+) r90=(\S+)$")
    assert all(pat.match(l) for l in its), its
    it, inside, maxres, r90 = pat.match(its[-1]).groups()
    assert (int(it), int(inside)) == (6, len(res))
    assert maxres == f"{res.max():.3e}" and r90 == f"{np.quantile(res, 0.9):.3e}"
    # the sparse form: maxiter iterations, no r90, no converged line
    lam2, X2, res2 = sm.chebyshev_poly_real_symm_sparse(A, emin, emax, ctx=ctx, **kw)
    out = capfd.readouterr().out.strip().splitlines()
    pat2 = re.compile(r"it=(\d+) inside=(\d+) maxres=(\S+)$")
    assert len(out) == 10 and all(pat2.match(l) for l in out), out
    it, inside, maxres = pat2.match(out[-1]).groups()
    assert (int(it), int(inside), maxres) == (10, len(res2), f"{res2.max():.3e}")
