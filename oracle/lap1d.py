"""The reference's Lap1D known-answer helpers (test infrastructure only).

``interval_by_count``      chebyshev_subspace_iteration.jl:980-992
``match_stats_sorted``     chebyshev_subspace_iteration.jl:918-941
``classify_*``             chebyshev_subspace_iteration.jl:942-979
``target_window``          polfed.jl:1223-1233
"""
from __future__ import annotations

import numpy as np

from sparse_matrix_algorithms_amd.lattices import lap1d_exact_eigs


def interval_by_count(n: int, nev_target: int, center_frac: float = 0.5, pad: int = 2):
    lam = lap1d_exact_eigs(n)
    ic = min(max(int(np.round(center_frac * n)), 1), n)
    h = nev_target // 2
    i1 = max(1, ic - h)
    i2 = min(n, i1 + nev_target - 1)
    i1 = max(1, i2 - nev_target + 1)
    emin = lam[max(1, i1 - pad) - 1]
    emax = lam[min(n, i2 + pad) - 1]
    sigma = 0.5 * (lam[i1 - 1] + lam[i2 - 1])
    return emin, emax, sigma, lam[i1 - 1:i2].copy(), (i1, i2)


def target_window(n: int, nev: int):
    lam = lap1d_exact_eigs(n)
    mid = min(max(int(np.round(n / 2)), 1), n)
    i1 = min(max(mid - nev // 2, 1), n)
    i2 = min(max(i1 + nev - 1, 1), n)
    i1 = min(max(i2 - nev + 1, 1), n)
    emin, emax = lam[i1 - 1], lam[i2 - 1]
    return lam, i1, i2, emin, emax, 0.5 * (emin + emax)


def match_stats_sorted(ref, got, tol=1e-8):
    nref, ngot = len(ref), len(got)
    ref_hit = np.zeros(nref, bool)
    got_hit = np.zeros(ngot, bool)
    i = j = 0
    while i < nref and j < ngot:
        d = got[j] - ref[i]
        if abs(d) <= tol:
            ref_hit[i] = got_hit[j] = True
            i += 1
            j += 1
        elif d < 0:
            j += 1
        else:
            i += 1
    return dict(nmatched_ref=int(ref_hit.sum()), nmissed=int((~ref_hit).sum()),
                extra_idx=np.nonzero(~got_hit)[0])


def classify_against_full_spectrum(got, allv, tol=1e-8):
    legit = np.zeros(len(got), bool)
    i = j = 0
    while i < len(got) and j < len(allv):
        d = got[i] - allv[j]
        if abs(d) <= tol:
            legit[i] = True
            i += 1
            j += 1
        elif d < 0:
            i += 1
        else:
            j += 1
    return legit


def classify_local_extras(ref, got, allv, tol=1e-8):
    loc = match_stats_sorted(ref, got, tol)
    legit = classify_against_full_spectrum(got, allv, tol)
    extra = loc["extra_idx"]
    return dict(nmatched_target=loc["nmatched_ref"], nmissed_target=loc["nmissed"],
                nextra=len(extra), nlegit_extra=int(legit[extra].sum()),
                nspurious_extra=int((~legit[extra]).sum()))
