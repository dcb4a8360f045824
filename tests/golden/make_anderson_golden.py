"""Write dense-FP64 eigenvalue fixtures for small Anderson workloads.

    python tests/golden/make_anderson_golden.py

Chebyshev workload: 3D Anderson L=20 (n=8000), W=16.5, disorder seed 20, all
eigenvalues in [-0.2, 0.2] from numpy.linalg.eigvalsh (LAPACK syevd) of the
dense matrix.  Also records the Gershgorin bounds and a checksum of the
onsite energies, so a change in the generator is caught.
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.polfed_ref import gershgorin_bounds  # noqa: E402
from sparse_matrix_algorithms_amd.lattices import anderson3d  # noqa: E402


def make(L=20, W=16.5, seed=20, emin=-0.2, emax=0.2):
    A = anderson3d(L, W, seed)
    ev = np.linalg.eigvalsh(A.toarray())
    sel = ev[(ev >= emin) & (ev <= emax)]
    # the eigenvalues just outside, to check the window edges are not ambiguous
    below = float(ev[ev < emin][-1])
    above = float(ev[ev > emax][0])
    smin, smax = gershgorin_bounds(A)
    return dict(L=L, W=W, seed=seed, n=A.shape[0], emin=emin, emax=emax, smin=smin, smax=smax,
                eigenvalues=[float(x) for x in sel], nearest_below=below, nearest_above=above,
                diag_sum=float(A.diagonal().sum()), diag_sq_sum=float((A.diagonal() ** 2).sum()),
                lowest=float(ev[0]), highest=float(ev[-1]))


if __name__ == "__main__":
    out = Path(__file__).resolve().parent / "anderson_L20_W16.5_seed20.json"
    out.write_text(json.dumps(make(), indent=1))
    print(f"wrote {out}")
