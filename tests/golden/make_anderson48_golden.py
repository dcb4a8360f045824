"""Eigenvalues of the L=48 Anderson workload near E=0, by a method independent
of the filter solvers under test.

    python tests/golden/make_anderson48_golden.py      # ~15 min, ~6 GB

Matrix: 3D Anderson L=48 (n=110,592), W=16.5, disorder seed 48, the matrix the
POLFED benchmark solves.

1. Shift-invert Lanczos (ARPACK through ``scipy.sparse.linalg.eigsh``,
   sigma=0, k=600).  The shifted solves use a sparse LU (SuperLU) of A under a
   geometric nested-dissection ordering of the periodic lattice, with diagonal
   pivots.  Every returned pair is checked by its residual: for unit v and
   symmetric A, |λ - λ_true| <= ||Av - λv||.
2. Sylvester inertia.  With diagonal pivots and no row exchanges the LU of the
   symmetric A - σI is L·D·Lᵀ, so the number of negative pivots is the number
   of eigenvalues below σ.  Two factorizations at shifts s_lo < 0 < s_hi
   (midpoints between neighbouring eigenvalues found in step 1) give the exact
   count in (s_lo, s_hi).  It must equal the number of step-1 eigenvalues
   there, so no eigenvalue in the window was missed.

The fixture keeps the eigenvalues in (s_lo, s_hi), the shifts, the counts and
the largest residual.  Any interval inside (s_lo, s_hi) can be checked against
it for completeness and values.
"""
import json
import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.polfed_ref import gershgorin_bounds  # noqa: E402
from sparse_matrix_algorithms_amd.lattices import anderson3d  # noqa: E402


def nested_dissection(L: int) -> np.ndarray:
    """Nested-dissection order of the periodic L³ lattice (site x + L·y + L²·z).

    A box is a range per axis: the whole periodic axis, or an interval of
    coordinates (mod L).  The longest axis is cut: a periodic axis by two
    planes (at 0 and L/2), an interval by its middle plane.  Both halves are
    ordered first, then the separator planes.  Boxes of <= 64 sites are
    ordered as they come.
    """
    out = []

    def coords(r):
        return np.arange(L) if r[0] == "p" else np.arange(r[1], r[2]) % L

    def length(r):
        return L if r[0] == "p" else r[2] - r[1]

    def sites(rs):
        X, Y, Z = np.meshgrid(*(coords(r) for r in rs), indexing="ij")
        return (X + L * Y + L * L * Z).ravel()

    def rec(rs):
        lens = [length(r) for r in rs]
        if min(lens) <= 0:
            return
        if np.prod(lens) <= 64:
            out.append(sites(rs))
            return
        d = int(np.argmax(lens))
        r = rs[d]
        if r[0] == "p":
            h = L // 2
            parts, seps = [("i", 1, h), ("i", h + 1, L)], [("i", 0, 1), ("i", h, h + 1)]
        else:
            m = (r[1] + r[2]) // 2
            parts, seps = [("i", r[1], m), ("i", m + 1, r[2])], [("i", m, m + 1)]
        for piece in parts + seps:
            rr = list(rs)
            rr[d] = piece
            if piece in parts:
                rec(rr)
            else:
                out.append(sites(rr))

    rec([("p", L)] * 3)
    perm = np.concatenate(out)
    assert len(perm) == L ** 3 and len(np.unique(perm)) == L ** 3
    return perm


def ldlt(Ap: sp.spmatrix, sigma: float):
    """LU of Ap - sigma·I with diagonal pivots only (= L·D·Lᵀ for symmetric Ap)."""
    M = (Ap - sigma * sp.identity(Ap.shape[0], format="csc")).tocsc()
    lu = sla.splu(M, permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(SymmetricMode=True))
    n = Ap.shape[0]
    if not (np.array_equal(lu.perm_r, np.arange(n)) and np.array_equal(lu.perm_c, np.arange(n))):
        raise RuntimeError("SuperLU exchanged rows or columns: the pivots are not D of LDLᵀ")
    # backward error of one solve, as a check on pivot growth
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n)
    x = lu.solve(b)
    berr = np.linalg.norm(M @ x - b) / (np.linalg.norm(b) + np.abs(M).sum(axis=1).max() *
                                         np.linalg.norm(x))
    return lu, int(np.sum(lu.U.diagonal() < 0)), float(berr)


def make(L=48, W=16.5, seed=48, k=600):
    A = anderson3d(L, W, seed).tocsr()
    n = A.shape[0]
    p = nested_dissection(L)
    Ap = A[p][:, p].tocsc()
    t = time.time()
    lu0, _, berr0 = ldlt(Ap, 0.0)
    print(f"LDLt at 0: {time.time() - t:.0f}s, backward error {berr0:.1e}", flush=True)
    op = sla.LinearOperator((n, n), matvec=lu0.solve, dtype=np.float64)
    v0 = np.random.default_rng(1).standard_normal(n)
    t = time.time()
    lam, V = sla.eigsh(Ap, k=k, sigma=0.0, which="LM", OPinv=op, v0=v0, tol=1e-14)
    print(f"eigsh: {time.time() - t:.0f}s", flush=True)
    del lu0
    order = np.argsort(lam)
    lam, V = lam[order], V[:, order]
    V /= np.linalg.norm(V, axis=0)
    res = np.linalg.norm(Ap @ V - V * lam, axis=0)
    # shifts: midpoints between neighbouring eigenvalues near 95% of the radius
    r = np.max(np.abs(lam))
    pos, neg = lam[lam > 0], lam[lam < 0][::-1]
    ip = int(np.searchsorted(pos, 0.95 * r))
    ineg = int(np.searchsorted(-neg, 0.95 * r))
    s_hi = 0.5 * (pos[ip - 1] + pos[ip])
    s_lo = 0.5 * (neg[ineg - 1] + neg[ineg])
    counts = {}
    for name, s in (("lo", s_lo), ("hi", s_hi)):
        t = time.time()
        _, nneg, berr = ldlt(Ap, s)
        counts[name] = (nneg, berr)
        print(f"LDLt at {s:+.6f}: {nneg} negative pivots, backward error {berr:.1e}, "
              f"{time.time() - t:.0f}s", flush=True)
    inertia = counts["hi"][0] - counts["lo"][0]
    sel = (lam > s_lo) & (lam < s_hi)
    if inertia != int(sel.sum()):
        raise RuntimeError(f"inertia count {inertia} != {int(sel.sum())} Lanczos eigenvalues")
    smin, smax = gershgorin_bounds(A)
    d = A.diagonal()
    return dict(L=L, W=W, seed=seed, n=n, smin=smin, smax=smax, s_lo=float(s_lo),
                s_hi=float(s_hi), count=inertia,
                negative_pivots_lo=counts["lo"][0], negative_pivots_hi=counts["hi"][0],
                backward_error=max(berr0, counts["lo"][1], counts["hi"][1]),
                max_residual=float(res[sel].max()),
                eigenvalues=[float(x) for x in lam[sel]],
                diag_sum=float(d.sum()), diag_sq_sum=float((d ** 2).sum()))


if __name__ == "__main__":
    out = Path(__file__).resolve().parent / "anderson_L48_W16.5_seed48_centre.json"
    g = make()
    out.write_text(json.dumps(g, indent=1))
    print(f"wrote {out}: {g['count']} eigenvalues in ({g['s_lo']:.6f}, {g['s_hi']:.6f}), "
          f"max residual {g['max_residual']:.1e}")
