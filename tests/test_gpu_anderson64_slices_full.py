"""The batched-slice workload (Anderson L=64, n=262,144) solved at its full
size against an independent eigenvalue fixture: four of its sixteen slices,
one polfed_slices run, mixed precision with learned parameters (as the
workload runs them), against shift-invert Lanczos eigenvalues whose count
around each slice is fixed exactly by Sylvester inertia
(golden/make_anderson64_slices_golden.py).

Each returned pair is also checked on the host: FP64 residual recomputed from
the returned vector, and orthonormality.  Correctness is the same eigenvalue
set as the fixture, |Δλ| <= 1e-10·‖A‖, relative residual <= 1e-10
(polfed.jl:1128-1158 solve, :892-984 extraction)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _host_checks(A, lam, X, res, anorm, tag):
    R = A @ X - X * lam[None, :]
    rn = np.linalg.norm(R, axis=0)
    assert rn.max() <= 1.01e-10 * anorm, (tag, rn.max() / anorm)
    assert np.allclose(rn, res, rtol=1e-3, atol=1e-14 * anorm), tag
    G = X.T @ X
    assert np.max(np.abs(G - np.eye(len(lam)))) <= 1e-8, tag


def _slice_width(nev):
    # the batched workload's block width: the power of two nearest 2.8 nev / 20 in [8, 64]
    return int(min(64, max(8, 2 ** round(math.log2(2.8 * nev / 20)))))


def test_anderson64_slices_mixed_learned_against_inertia_fixture(ctx):
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd import kpm
    from sparse_matrix_algorithms_amd.lattices import anderson3d
    from sparse_matrix_algorithms_amd.learned import MLP, polfed_slices_learned
    g = json.loads((GOLDEN / "anderson_L64_W16.5_seed64_slices.json").read_text())
    A = anderson3d(g["L"], g["W"], g["seed"])
    n = A.shape[0]
    assert n == g["n"] and np.isclose(A.diagonal().sum(), g["diag_sum"], rtol=1e-12)
    dm = sm.DeviceMatrix(A, ctx)
    try:
        smin, smax = dm.gershgorin_bounds()
        assert (smin, smax) == (g["smin"], g["smax"])
        anorm = max(abs(smin), abs(smax))
        mu, _, _ = kpm.kpm_moments(dm, 256, 16, 7, smin, smax, ctx=ctx)
        model = MLP.load(ctx=ctx)
        base = [dict(emin=sl["emin"], emax=sl["emax"], sigma=sl["centre"], nev=sl["nev"],
                     s=_slice_width(sl["nev"])) for sl in g["slices"]]
        outs, slices, info = polfed_slices_learned(dm, base, mu, model, smin, smax,
                                                   precision="fp32", ctx=ctx)
        for sl, d in zip(g["slices"], slices):
            print(f"slice {sl['slice']}: [{sl['emin']:.6f},{sl['emax']:.6f}] nev={sl['nev']} "
                  f"s={d['s']} K={d['degree']} mmax={d['mmax']} ({d['source']})", flush=True)
        print(f"batches {info['seconds']} s, retried {info['retried']}", flush=True)
        for sl, d, (lam, X, res, st) in zip(g["slices"], slices, outs):
            ev = np.array(sl["eigenvalues"])
            exact = ev[(ev >= sl["emin"]) & (ev <= sl["emax"])]
            assert len(exact) == sl["count"]
            assert len(lam) == len(exact), (sl["slice"], len(lam), len(exact))
            assert np.max(np.abs(np.sort(lam) - exact)) <= 1e-10 * anorm, sl["slice"]
            assert res.max() <= 1e-10 * anorm, sl["slice"]
            _host_checks(A, lam, X, res, anorm, f"slice {sl['slice']}")
            print(f"slice {sl['slice']}: {len(lam)} pairs, max |Δλ| "
                  f"{np.max(np.abs(np.sort(lam) - exact)):.1e}, max rel res "
                  f"{res.max() / anorm:.1e}, {st['t_total']:.1f}s batch", flush=True)
    finally:
        dm.free()
