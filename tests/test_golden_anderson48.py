"""Self-consistency of the L=48 Anderson eigenvalue fixture near E=0
(``golden/make_anderson48_golden.py``: shift-invert Lanczos, with the count
fixed by Sylvester inertia), and a check that the generator still builds the
matrix the fixture was made from."""
import json
from pathlib import Path

import numpy as np

from oracle.polfed_ref import gershgorin_bounds
from sparse_matrix_algorithms_amd.lattices import anderson3d

GOLDEN = Path(__file__).resolve().parent / "golden" / "anderson_L48_W16.5_seed48_centre.json"


def test_fixture_matches_generator():
    g = json.loads(GOLDEN.read_text())
    A = anderson3d(g["L"], g["W"], g["seed"])
    d = A.diagonal()
    assert A.shape[0] == g["n"] == 48 ** 3
    assert np.isclose(d.sum(), g["diag_sum"], rtol=1e-13, atol=1e-9)
    assert np.isclose((d ** 2).sum(), g["diag_sq_sum"], rtol=1e-13)
    smin, smax = gershgorin_bounds(A)
    assert (smin, smax) == (g["smin"], g["smax"])


def test_fixture_is_consistent():
    g = json.loads(GOLDEN.read_text())
    ev = np.array(g["eigenvalues"])
    anorm = max(abs(g["smin"]), abs(g["smax"]))
    assert len(ev) == g["count"] == g["negative_pivots_hi"] - g["negative_pivots_lo"]
    assert np.all(np.diff(ev) > 0)
    assert g["s_lo"] < ev[0] and ev[-1] < g["s_hi"]
    assert g["s_lo"] < 0 < g["s_hi"]
    # every pair certified: |λ - λ_true| <= residual of a unit vector
    assert g["max_residual"] <= 1e-10 * anorm
    assert g["backward_error"] <= 1e-10
    # the benchmark interval (about ±0.0394, ~494 pairs) lies well inside
    assert g["s_hi"] > 0.04 and g["s_lo"] < -0.04 and g["count"] > 494
