"""Seeded sparse test matrices: 3D Anderson, random-field XXZ (Sz=0), 1D Laplacian.

All generators return ``scipy.sparse.csr_matrix`` (float64, sorted int32
column indices, no duplicates).  Disorder is drawn from the counter-based
``splitmix64`` stream below, so a (seed, site) pair always gives the same value
on every platform; ``u = (z >> 11) · 2⁻⁵³`` maps each 64-bit output to [0, 1).

The reference holds only the 1D Laplacian (``polfed.jl:1192-1195``,
``chebyshev_subspace_iteration.jl:875-879``); the Anderson and XXZ models are
the workloads the eigensolver is measured on.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Counter-based splitmix64: element ``i`` is ``mix(seed + (offset+i+1)·γ)``."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """``count`` doubles in [0,1) from :func:`splitmix64`."""
    z = splitmix64(seed, count, offset)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian_block(seed: int, n: int, b: int) -> np.ndarray:
    """``n×b`` standard normals by Box–Muller on two splitmix64 streams
    (column-major fill order: element ``(i,j)`` uses counter ``j·n+i``)."""
    cnt = n * b
    u1 = uniform01(seed, cnt)
    u2 = uniform01(seed ^ 0x5DEECE66D, cnt)
    g = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)
    return g.reshape(b, n).T.copy()


def anderson3d(L: int, W: float, seed: int, t: float = -1.0) -> sp.csr_matrix:
    """3D Anderson model on a periodic ``L³`` lattice.

    Site ``i = x + L·y + L²·z``; hopping ``t`` (default −1) to the six nearest
    neighbours; onsite energies ``ε_i = W·(u_i − ½)`` i.i.d. uniform in
    ``[−W/2, W/2]`` with ``u_i`` from :func:`uniform01` (``seed``).
    """
    n = L ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    y = (idx // L) % L
    z = idx // (L * L)
    eps = W * (uniform01(seed, n) - 0.5)
    rows = [idx]
    cols = [idx]
    vals = [eps]
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        j = ((x + dx) % L) + L * ((y + dy) % L) + L * L * ((z + dz) % L)
        rows.append(idx)
        cols.append(j)
        vals.append(np.full(n, t))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def anderson3d_clean_spectrum(L: int, t: float = -1.0) -> np.ndarray:
    """Closed-form spectrum of the W=0 periodic model: ``2t·Σ_d cos(2πk_d/L)``."""
    c = 2.0 * t * np.cos(2.0 * math.pi * np.arange(L) / L)
    return np.sort((c[:, None, None] + c[None, :, None] + c[None, None, :]).ravel())


def xxz_sz0(L: int, h: float, seed: int, J: float = 1.0, Delta: float = 1.0) -> sp.csr_matrix:
    """Random-field XXZ spin-½ chain, periodic, in the ``S^z_total = 0`` sector.

    ``H = J Σ_i (S^x_i S^x_{i+1} + S^y_i S^y_{i+1} + Δ S^z_i S^z_{i+1}) + Σ_i h_i S^z_i``
    with ``S = σ/2`` and ``h_i = h·(2u_i − 1)`` uniform in ``[−h, h]``.  Basis:
    all ``L``-bit words with ``L/2`` set bits in ascending order (bit ``i`` set =
    spin up at site ``i``).
    """
    if L % 2:
        raise ValueError("Sz=0 sector needs even L")
    hs = h * (2.0 * uniform01(seed, L) - 1.0)
    words = np.arange(1 << L, dtype=np.int64)
    pop = np.zeros_like(words)
    for i in range(L):
        pop += (words >> i) & 1
    states = words[pop == L // 2]
    n = len(states)
    spin = np.stack([((states >> i) & 1) * 1.0 - 0.5 for i in range(L)])   # (L, n)
    diag = np.zeros(n)
    for i in range(L):
        diag += J * Delta * spin[i] * spin[(i + 1) % L] + hs[i] * spin[i]
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    vals = [diag]
    for i in range(L):
        j = (i + 1) % L
        mask = (1 << i) | (1 << j)
        bi = (states >> i) & 1
        bj = (states >> j) & 1
        sel = np.nonzero(bi != bj)[0]
        flipped = states[sel] ^ mask
        rows.append(sel)
        cols.append(np.searchsorted(states, flipped))
        vals.append(np.full(len(sel), 0.5 * J))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def lap1d(n: int) -> sp.csr_matrix:
    """``tridiag(−1, 2, −1)`` (``polfed.jl:1192-1195``)."""
    e = np.ones(n)
    return sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")


def lap1d_exact_eigs(n: int) -> np.ndarray:
    """``2 − 2cos(jπ/(n+1))``, j=1..n (``polfed.jl:1220-1222``)."""
    j = np.arange(1, n + 1)
    return 2 - 2 * np.cos(j * math.pi / (n + 1))
