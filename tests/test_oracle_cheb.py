"""Pins for the oracle's Chebyshev subspace iteration
(chebyshev_subspace_iteration.jl:667-848) on the reference's own Lap1D
count test: matched / legit extra / spurious classification (:918-979)."""
import numpy as np
import pytest

from oracle import cheb_ref as CR
from oracle.lap1d import classify_local_extras, interval_by_count
from sparse_matrix_algorithms_amd.lattices import anderson3d, gaussian_block, lap1d, lap1d_exact_eigs


def test_lap1d_count_classification():
    n, nev = 2000, 40
    A = lap1d(n)
    lall = lap1d_exact_eigs(n)
    emin, emax, sigma, lref, _ = interval_by_count(n, nev)
    m0 = int(np.ceil(2.0 * nev))
    lam, X, res = CR.chebyshev_poly_real_symm_sparse(A, emin, emax, smin=0.0, smax=4.0,
                                                     sigma=sigma, m0=m0, degree=128, maxiter=8,
                                                     tol=1e-8, window_type="delta",
                                                     Q0=gaussian_block(1, n, m0))
    cls = classify_local_extras(np.sort(lref), np.sort(lam), np.sort(lall), tol=1e-8)
    assert cls["nmatched_target"] == nev
    assert cls["nspurious_extra"] == 0


def test_sparse_form_runs_maxiter_mf_form_stops():
    n = 600
    A = lap1d(n)
    lex = lap1d_exact_eigs(n)
    emin, emax = 0.5 * (lex[289] + lex[290]), 0.5 * (lex[309] + lex[310])
    Q0 = gaussian_block(5, n, 40)
    kw = dict(smin=0.0, smax=4.0, m0=40, degree=150, maxiter=10, tol=1e-9, Q0=Q0,
              return_info=True)
    *_, info_sp = CR.chebyshev_poly_real_symm_sparse(A, emin, emax, **kw)
    *_, info_mf = CR.chebyshev_poly_real_symm_sparse_mf(
        lambda y, x: None, n, emin, emax, matblock=lambda Y, X: Y.__setitem__(slice(None), A @ X),
        **kw)
    assert info_sp["iterations"] == 10
    assert info_mf["iterations"] == 6


def test_argument_errors():
    A = lap1d(30)
    with pytest.raises(ValueError, match="emin<emax"):
        CR.chebyshev_poly_real_symm_sparse(A, 1.0, 0.0)
    with pytest.raises(ValueError, match="not inside"):
        CR.chebyshev_poly_real_symm_sparse(A, 0.5, 9.0, smin=0.0, smax=4.0)
    with pytest.raises(ValueError, match="Unknown window"):
        CR.chebyshev_poly_real_symm_sparse(A, 0.5, 1.0, window_type="box")
    # the @warn of :683 is always followed by the ArgumentError of :689 (an
    # interval outside the bounds maps outside [-1,1])
    with pytest.warns(UserWarning, match="not contained"):
        with pytest.raises(ValueError, match="not inside"):
            CR.chebyshev_poly_real_symm_sparse(A, 0.5, 1.0, smin=0.6, smax=4.0)


def test_anderson_dense_golden_c1_scaled():
    """Chebyshev FP64 run as on the L=20 benchmark cube, on a small Anderson cube vs dense eigh."""
    A = anderson3d(8, 16.5, 20)
    ev = np.linalg.eigh(A.toarray())[0]
    emin, emax = -0.5, 0.5
    ref = ev[(ev >= emin) & (ev <= emax)]
    lam, X, res = CR.chebyshev_poly_real_symm_sparse(A, emin, emax, m0=2 * len(ref) + 10,
                                                     degree=250, maxiter=6,
                                                     Q0=gaussian_block(1, A.shape[0],
                                                                       2 * len(ref) + 10))
    assert len(lam) == len(ref)
    assert np.max(np.abs(lam - ref)) <= 1e-9
