"""Self-consistency of the eigenvalue fixture of the XXZ workload (random-field
XXZ chain L=20, Sz=0: the oracle's FP64 polfed on a wider interval,
golden/make_xxz20_golden.py), and a check that the generator still builds the
matrix it was made from."""
import json
from pathlib import Path

import numpy as np

from oracle.polfed_ref import gershgorin_bounds
from sparse_matrix_algorithms_amd.lattices import xxz_sz0

GOLDEN = Path(__file__).resolve().parent / "golden" / "xxz_L20_h5_seed20_centre.json"


def test_xxz20_fixture():
    g = json.loads(GOLDEN.read_text())
    A = xxz_sz0(g["L"], g["h"], g["seed"])
    d = A.diagonal()
    assert (A.shape[0], A.nnz) == (g["n"], g["nnz"]) == (184_756, g["nnz"])
    assert np.isclose(d.sum(), g["diag_sum"], rtol=1e-13, atol=1e-9)
    assert np.isclose((d ** 2).sum(), g["diag_sq_sum"], rtol=1e-13)
    assert gershgorin_bounds(A) == (g["smin"], g["smax"])
    assert np.isclose(g["centre"], d.sum() / A.shape[0], rtol=1e-14)
    anorm = max(abs(g["smin"]), abs(g["smax"]))
    ev = np.array(g["eigenvalues"])
    assert len(ev) == g["count"] == 960
    assert np.all(np.diff(ev) > 0)
    # the fixture interval holds exactly these, strictly inside the oracle's window
    a, b = g["oracle_interval"]
    assert a < g["emin"] < ev[0] and ev[-1] < g["emax"] < b
    assert g["emin"] <= g["centre"] <= g["emax"]
    # at least 40 more converged oracle eigenvalues on each side of the fixture interval
    assert g["oracle_converged"] >= g["count"] + 80
    # each eigenvalue certified by its residual (|λ - λ_true| <= ||r|| for unit v)
    assert g["max_residual"] <= 1e-10 * anorm
