"""The XXZ workload (random-field XXZ chain, L=20, Sz=0, n=184,756, h=5,
seed 20, ~1000 pairs at the centre of the spectrum, block 128) solved at its
full size with its mixed path (FP32 basis, FP64 residual-form refinement,
plain degree x0.7) against an independent FP64 fixture
(golden/make_xxz20_golden.py): the oracle restatement of polfed in FP64 on a
wider interval, each eigenvalue certified by its residual.  The fixture
interval holds the 960 converged oracle eigenvalues nearest tr(H)/n with at
least 40 more on each side.

Each returned pair is checked on the host as well: FP64 residual recomputed
from the returned vector, and orthonormality.  Correctness is the same
eigenvalue set as the fixture, |Δλ| <= 1e-10·‖A‖ and relative residual <=
1e-10 (polfed.jl:1128-1158 solve, :892-984 extraction)."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "xxz_L20_h5_seed20_centre.json"
pytestmark = pytest.mark.gpu


def test_xxz20_polfed_mixed_against_oracle_fixture(ctx):
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd import kpm
    from sparse_matrix_algorithms_amd.lattices import xxz_sz0
    g = json.loads(GOLDEN.read_text())
    A = xxz_sz0(g["L"], g["h"], g["seed"])
    n = A.shape[0]
    d = A.diagonal()
    assert (n, A.nnz) == (g["n"], g["nnz"])
    assert np.isclose(d.sum(), g["diag_sum"], rtol=1e-12)
    assert np.isclose((d ** 2).sum(), g["diag_sq_sum"], rtol=1e-12)
    dm = sm.DeviceMatrix(A, ctx)
    try:
        smin, smax = dm.gershgorin_bounds()
        assert np.isclose(smin, g["smin"], rtol=1e-14) and np.isclose(smax, g["smax"], rtol=1e-14)
        anorm = max(abs(smin), abs(smax))
        emin, emax, nev, s = g["emin"], g["emax"], 1000, 128
        # the workload's mixed parameters: plain degree x0.7, mmax = ceil(2.8 nev / s)
        K, mm = kpm.plain_parameters(n, nev, emin, emax, smin, smax, s, precision="fp32")
        lam, X, res, st = sm.polfed(dm, nev, K, s=s, smin=smin, smax=smax, sigma=g["centre"],
                                    emin=emin, emax=emax, tol=1e-10 * anorm, mmax=mm,
                                    precision="fp32", ctx=ctx, return_stats=True)
        print(f"xxz L=20 [{emin:.6f},{emax:.6f}] K={K} mmax={mm}: {len(lam)} pairs in "
              f"{st['t_total']:.1f}s ({st['line']})", flush=True)
        exact = np.array(g["eigenvalues"])
        assert len(exact) == g["count"]
        assert len(lam) == len(exact), (len(lam), len(exact))
        dl = np.max(np.abs(np.sort(lam) - exact))
        assert dl <= 1e-10 * anorm, dl
        assert res.max() <= 1e-10 * anorm, res.max() / anorm
        R = A @ X - X * lam[None, :]
        rn = np.linalg.norm(R, axis=0)
        assert rn.max() <= 1.01e-10 * anorm, rn.max() / anorm
        assert np.allclose(rn, res, rtol=1e-3, atol=1e-14 * anorm)
        del R
        G = X.T @ X
        assert np.max(np.abs(G - np.eye(len(lam)))) <= 1e-8
        print(f"xxz L=20: {len(lam)} pairs, max |Δλ| {dl:.1e}, max rel res "
              f"{res.max() / anorm:.1e}", flush=True)
    finally:
        dm.free()
