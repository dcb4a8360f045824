"""FP64 CPU restatement of ``polfed.jl`` (oracle; test infrastructure only).

Every function follows the reference control flow, selection rules and
defaults; each docstring cites the ``polfed.jl`` lines it restates.  Dense
linear algebra uses NumPy's LAPACK (Householder ``geqrf``/``orgqr`` QR, ``syevd``
symmetric eigensolver), the same LAPACK families the reference reaches through
Julia's ``qr!``/``eigen``.  Sparse products use SciPy CSR.

Operators are either a SciPy sparse matrix or a :class:`MatrixFreeRealSymOp`.
Blocks are NumPy arrays of shape ``(n, b)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp


# ----------------------------------------------------------------------------
# operator abstraction                                   polfed.jl:148-219
# ----------------------------------------------------------------------------
@dataclass
class MatrixFreeRealSymOp:
    """Matrix-free symmetric operator (``polfed.jl:148-165``).

    ``matvec(y, x)`` writes ``A x`` into ``y``; ``matblock(Y, X)`` (optional)
    writes ``A X`` into ``Y``.
    """

    n: int
    matvec: Callable
    matblock: Optional[Callable] = None


def apply_A_block(A, X: np.ndarray) -> np.ndarray:
    """``Y = A X`` (``polfed.jl:196-199`` sparse, ``209-219`` matrix-free)."""
    if isinstance(A, MatrixFreeRealSymOp):
        Y = np.zeros_like(X)
        if A.matblock is not None:
            A.matblock(Y, X)
        else:
            for j in range(X.shape[1]):
                y = np.zeros(X.shape[0])
                A.matvec(y, np.ascontiguousarray(X[:, j]))
                Y[:, j] = y
        return Y
    return np.asarray(A @ X)


def op_size(A) -> int:
    return A.n if isinstance(A, MatrixFreeRealSymOp) else A.shape[0]


# ----------------------------------------------------------------------------
# spectral bounds and filter coefficients               polfed.jl:254-414
# ----------------------------------------------------------------------------
def gershgorin_bounds(A) -> tuple[float, float]:
    """``[min(a_ii - r_i), max(a_ii + r_i)]``, ``r_i = sum_{j!=i} |a_ij|``
    (``polfed.jl:254-266``; same rule as ``chebyshev_subspace_iteration.jl:351-367``)."""
    C = sp.coo_matrix(A)
    n = C.shape[0]
    d = np.zeros(n)
    r = np.zeros(n)
    on = C.row == C.col
    np.add.at(d, C.row[on], C.data[on])
    np.add.at(r, C.row[~on], np.abs(C.data[~on]))
    return float(np.min(d - r)), float(np.max(d + r))


def jackson_kernel(K: int) -> np.ndarray:
    """Jackson damping ``g_0..g_K`` (``polfed.jl:275-284``)."""
    if K <= 0:
        return np.array([1.0])
    a = math.pi / (K + 2)
    sa = math.sin(a)
    g = np.zeros(K + 1)
    for k in range(K + 1):
        g[k] = math.sin((k + 1) * a) / ((K + 2) * sa) + (1 - (k + 1) / (K + 2)) * math.cos(k * a)
    return g


def cheb_delta_coeffs(sh: float, K: int, jackson: bool = True) -> np.ndarray:
    """Delta-peak coefficients at ``sh`` (``polfed.jl:299-313``)."""
    c = np.zeros(K + 1)
    th = math.acos(min(max(sh, -1.0), 1.0))
    c[0] = 1 / math.pi
    for k in range(1, K + 1):
        c[k] = 2 * math.cos(k * th) / math.pi
    if jackson:
        c *= jackson_kernel(K)
    return c


def cheb_window_coeffs(a: float, b: float, K: int, jackson: bool = True) -> np.ndarray:
    """Indicator of ``[a,b]`` (``polfed.jl:329-347``)."""
    aa = max(-1.0, min(1.0, a))
    bb = max(-1.0, min(1.0, b))
    if aa > bb:
        aa, bb = bb, aa
    al = math.acos(bb)
    be = math.acos(aa)
    c = np.zeros(K + 1)
    c[0] = (be - al) / math.pi
    for k in range(1, K + 1):
        c[k] = 2 * (math.sin(k * be) - math.sin(k * al)) / (k * math.pi)
    if jackson:
        c *= jackson_kernel(K)
    return c


def cheb_smooth_window_coeffs(a: float, b: float, K: int, jackson: bool = True,
                              nsamp: int = 4096, delta: float = 0.02) -> np.ndarray:
    """tanh-ramped window, trapezoid in theta (``polfed.jl:364-388``)."""
    aa = max(-1.0, min(1.0, a))
    bb = max(-1.0, min(1.0, b))
    if aa > bb:
        aa, bb = bb, aa
    w = math.pi / (nsamp - 1)
    th = np.arange(nsamp) * w
    xi = np.cos(th)
    fi = 0.5 * (1 + np.tanh((xi - aa) / delta)) * 0.5 * (1 - np.tanh((xi - bb) / delta))
    wt = np.ones(nsamp)
    wt[0] = wt[-1] = 0.5
    c = np.zeros(K + 1)
    for k in range(K + 1):
        s = float(np.sum(wt * fi * np.cos(k * th)))
        c[k] = (1 / math.pi if k == 0 else 2 / math.pi) * w * s
    if jackson:
        c *= jackson_kernel(K)
    return c


def build_coeffs(degree: int, smin: float, smax: float, sigma: float, emin: float,
                 emax: float, window_type: str = "step", jackson: bool = True) -> np.ndarray:
    """Map to ``[-1,1]`` and dispatch on the window (``polfed.jl:404-414``)."""
    c = 0.5 * (smin + smax)
    d = 0.5 * (smax - smin)
    sh = (sigma - c) / d
    a = (emin - c) / d
    b = (emax - c) / d
    if window_type == "step":
        return cheb_window_coeffs(a, b, degree, jackson)
    if window_type == "delta":
        return cheb_delta_coeffs(sh, degree, jackson)
    if window_type == "smooth":
        return cheb_smooth_window_coeffs(a, b, degree, jackson)
    raise ValueError(f"Unknown window type: {window_type}")


# ----------------------------------------------------------------------------
# block Clenshaw filter                                  polfed.jl:464-545
# ----------------------------------------------------------------------------
def apply_poly_block(A, Q: np.ndarray, coeffs: np.ndarray, smin: float, smax: float) -> np.ndarray:
    """``Y = p(Â) Q`` with ``Â=(A-cI)/d`` by block Clenshaw (``polfed.jl:464-501``).

    ``B_{K+1}=B_{K+2}=0``; ``B_k = c_k Q + α A B_{k+1} + β B_{k+1} - B_{k+2}``
    with the reference's scalars ``invd=2/(smax-smin)``, ``α=2·invd``,
    ``β=-2c·invd``; closing ``Y = c_0 Q + invd·A B_1 - c·invd·B_1 - B_2``.
    """
    K = len(coeffs) - 1
    c = 0.5 * (smin + smax)
    invd = 2.0 / (smax - smin)
    al = 2.0 * invd
    be = -2.0 * c * invd
    B1 = np.zeros_like(Q)
    B2 = np.zeros_like(Q)
    for k in range(K + 1, 1, -1):          # reference 1-based k = K+1 .. 2
        AX = apply_A_block(A, B1)
        ck = coeffs[k - 1]
        TMP = ck * Q + al * AX + be * B1 - B2
        B1, B2 = TMP, B1
    AX = apply_A_block(A, B1)
    c0 = coeffs[0]
    return c0 * Q + invd * AX - c * invd * B1 - B2


# ----------------------------------------------------------------------------
# orthogonalisation                                      polfed.jl:561-685
# ----------------------------------------------------------------------------
def _rank_from_R(R: np.ndarray, rtol: float) -> int:
    k = min(R.shape)
    if k == 0:
        return 0
    dr = np.abs(np.diag(R[:k, :k]))
    r0 = dr.max()
    if r0 == 0:
        return 0
    return int(np.count_nonzero(dr > rtol * r0))


def orth_block_polfed(Q: np.ndarray, rtol: float = 1e-12) -> np.ndarray:
    """QR, keep the FIRST ``r = #{|R_ii| > rtol·max|R_ii|}`` columns (``polfed.jl:561-594``)."""
    Qf, R = np.linalg.qr(Q, mode="reduced")
    r = _rank_from_R(R, rtol)
    return np.ascontiguousarray(Qf[:, :r])


def orth_against(W: np.ndarray, Q: np.ndarray) -> np.ndarray:
    """``W -= Q (Qᵀ W)`` (``polfed.jl:605-609``)."""
    if Q.shape[1] == 0:
        return W
    W -= Q @ (Q.T @ W)
    return W


def reorth_full(W: np.ndarray, Qhist: list, Qcur: np.ndarray, passes: int = 1) -> np.ndarray:
    """Against every stored block, then once more against ``Qcur`` (``polfed.jl:623-631``)."""
    for _ in range(passes):
        for Q in Qhist:
            orth_against(W, Q)
        orth_against(W, Qcur)
    return W


def block_qr_factor(W: np.ndarray, rtol: float = 1e-12):
    """``W ≈ Q B`` truncated; returns ``(Q[:, :r], R[:r, :])`` (``polfed.jl:644-685``)."""
    n, m = W.shape
    Qf, R = np.linalg.qr(W, mode="reduced")
    r = _rank_from_R(R, rtol)
    if r == 0:
        return np.zeros((n, 0)), np.zeros((0, m))
    return np.ascontiguousarray(Qf[:, :r]), np.array(R[:r, :m])


# ----------------------------------------------------------------------------
# reduced matrix, Ritz vectors, Rayleigh quotients        polfed.jl:701-880
# ----------------------------------------------------------------------------
def sym_upper(M: np.ndarray) -> np.ndarray:
    """``Matrix(Symmetric(M))``: mirror the upper triangle."""
    U = np.triu(M)
    return U + np.triu(M, 1).T


def assemble_T(Ablocks: list, Bblocks: list) -> np.ndarray:
    """Dense block-tridiagonal ``T`` (``polfed.jl:701-754``)."""
    nb = len(Ablocks)
    dims = [A.shape[0] for A in Ablocks]
    N = sum(dims)
    T = np.zeros((N, N))
    o = 0
    for j in range(nb):
        r = dims[j]
        T[o:o + r, o:o + r] = Ablocks[j]
        if j < nb - 1:
            B = Bblocks[j]
            T[o + r:o + r + B.shape[0], o:o + r] = B
            T[o:o + r, o + r:o + r + B.shape[0]] = B.T
        o += r
    return sym_upper(T)


def build_X(Qhist: list, U: np.ndarray) -> np.ndarray:
    """``X = Σ_j Q_j U[o_j:o_j+r_j, :]`` (``polfed.jl:767-776``)."""
    n = Qhist[0].shape[0]
    X = np.zeros((n, U.shape[1]))
    o = 0
    for Q in Qhist:
        r = Q.shape[1]
        X += Q @ U[o:o + r, :]
        o += r
    return X


def rayleigh(A, X: np.ndarray):
    """``λ_j = x_jᵀ A x_j``, ``res_j = ‖A x_j − λ_j x_j‖₂`` (``polfed.jl:791-830``)."""
    AX = apply_A_block(A, X)
    lam = np.einsum("ij,ij->j", X, AX)
    R = AX - X * lam[None, :]
    res = np.sqrt(np.einsum("ij,ij->j", R, R))
    return lam, res


def refine_original_rr(A, X: np.ndarray):
    """Original-space Rayleigh–Ritz (``polfed.jl:837-880``)."""
    Q = orth_block_polfed(X.copy())
    k = Q.shape[1]
    n = X.shape[0]
    if k == 0:
        return np.zeros(0), np.zeros((n, 0)), np.zeros(0)
    AQ = apply_A_block(A, Q)
    H = sym_upper(0.5 * (Q.T @ AQ + AQ.T @ Q))
    _, V = np.linalg.eigh(H)
    Y = Q @ V
    lam, res = rayleigh(A, Y)
    p = np.argsort(lam, kind="stable")
    return lam[p], Y[:, p], res[p]


def extract_pairs(Qhist, Ablocks, Bblocks, A, nev: int, emin: float, emax: float,
                  overextract: int = 2, refine: bool = True, extra_keep: int = 8, chunk: int = 64):
    """Extraction of Ritz pairs from the block basis (``polfed.jl:892-984``)."""
    T = assemble_T(Ablocks, Bblocks)
    vals, vecs = np.linalg.eigh(T)
    m = min(len(vals), max(overextract * nev, nev + extra_keep))
    p = np.argsort(-vals, kind="stable")[:m]      # sortperm(vals; rev=true)
    thsel = vals[p]
    Usel = vecs[:, p]
    n = Qhist[0].shape[0]
    lam_keep, res_keep, idx_keep = [], [], []
    for j0 in range(0, m, chunk):
        j1 = min(j0 + chunk, m)
        X = build_X(Qhist, Usel[:, j0:j1])
        lam, res = rayleigh(A, X)
        for j in range(j1 - j0):
            if emin <= lam[j] <= emax:
                lam_keep.append(lam[j])
                res_keep.append(res[j])
                idx_keep.append(j0 + j)
    empty = (np.zeros(0), np.zeros((n, 0)), np.zeros(0), thsel)
    if not idx_keep:
        return empty
    lam_keep = np.array(lam_keep)
    res_keep = np.array(res_keep)
    idx_keep = np.array(idx_keep)
    npre = min(len(res_keep), nev + extra_keep)
    ps0 = np.argsort(res_keep, kind="stable")[:npre]
    lamc, resc, idxc = lam_keep[ps0], res_keep[ps0], idx_keep[ps0]
    Xc = build_X(Qhist, Usel[:, idxc])
    if refine:
        lamc, Xc, resc = refine_original_rr(A, Xc)
        inside2 = np.nonzero((emin <= lamc) & (lamc <= emax))[0]
        if len(inside2) == 0:
            return empty
        lamc, resc, Xc = lamc[inside2], resc[inside2], Xc[:, inside2]
    nkeep = min(len(resc), nev + extra_keep)
    ps = np.argsort(resc, kind="stable")[:nkeep]
    lamf, Xf, rf = lamc[ps], Xc[:, ps], resc[ps]
    pe = np.argsort(lamf, kind="stable")
    return lamf[pe], Xf[:, pe], rf[pe], thsel


def julia_float(v: float) -> str:
    """``v`` as Julia's ``print(::Float64)`` writes it (what ``println`` puts in
    the progress line, ``polfed.jl:999``): the shortest digit string that
    round-trips, ``value = 0.d1…dk × 10^pt``; fixed notation when
    ``-4 < pt <= 6``, otherwise ``d1.d2…dk e<pt-1>`` with no ``+`` and no zero
    padding; a single digit keeps a trailing ``.0`` (``1.0e-5``, ``0.0001``,
    ``123456.0``, ``1.234567e6``)."""
    import decimal
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Inf" if v > 0 else "-Inf"
    if v == 0.0:
        return "-0.0" if math.copysign(1.0, v) < 0 else "0.0"
    # repr gives the shortest round-trip digits (David Gay's algorithm)
    sign, dig, e = decimal.Decimal(repr(v)).as_tuple()
    digits = "".join(map(str, dig)).rstrip("0") or "0"
    e += len(dig) - len(digits)          # value = digits × 10^e
    k = len(digits)
    pt = e + k
    out = "-" if sign else ""
    if -4 < pt <= 6:
        if pt <= 0:
            return out + "0." + "0" * (-pt) + digits
        if pt >= k:
            return out + digits + "0" * (pt - k) + ".0"
        return out + digits[:pt] + "." + digits[pt:]
    return out + digits[0] + "." + (digits[1:] or "0") + "e" + str(pt - 1)


def progress_line(m: int, lam, res) -> str:
    """``m=…  inside=…  r90=…  r95=…  maxres=…`` (``polfed.jl:991-1000``)."""
    if len(res) == 0:
        return f"m={m}  inside=0  maxres=Inf"
    rs = np.sort(res)
    L = len(rs)

    def _ridx(f):  # Julia round(Int, x): round half to even, 1-based clamp
        return min(max(int(np.round(f * L)), 1), L) - 1

    return (f"m={m}  inside={len(lam)}  r90={julia_float(rs[_ridx(0.9)])}"
            f"  r95={julia_float(rs[_ridx(0.95)])}  maxres={julia_float(rs[-1])}")


# ----------------------------------------------------------------------------
# block-Lanczos basis and driver                         polfed.jl:1022-1186
# ----------------------------------------------------------------------------
def build_full_basis(A, Q1: np.ndarray, coeffs, smin: float, smax: float, mmax: int,
                     rtol: float = 1e-12, reorth_passes: int = 1):
    """Block Lanczos on ``p(Â)`` with full re-orthogonalisation (``polfed.jl:1022-1099``).

    Returns ``(Qhist, Ablocks, Bblocks)``.
    """
    Qhist, Ablocks, Bblocks = [], [], []
    n = Q1.shape[0]
    Q = Q1.copy()
    Qprev = np.zeros((n, 0))
    Bprev = np.zeros((0, 0))
    for _ in range(mmax):
        Qhist.append(Q.copy())
        W = apply_poly_block(A, Q, coeffs, smin, smax)
        if Qprev.shape[1] > 0 and Bprev.shape[0] > 0 and Bprev.shape[1] > 0:
            W -= Qprev @ Bprev.T
        Aj = sym_upper(0.5 * (Q.T @ W + W.T @ Q))
        Ablocks.append(Aj)
        W -= Q @ Aj
        for _ in range(reorth_passes):
            reorth_full(W, Qhist, Q, passes=1)
        Qnext, Bj = block_qr_factor(W, rtol)
        if Qnext.shape[1] == 0:
            break
        Bblocks.append(Bj)
        Qprev, Bprev, Q = Q, Bj, Qnext
    return Qhist, Ablocks, Bblocks


def polfed(A, nev: int, degree: int, s: int = 8, smin=None, smax=None, sigma: float = 0.0,
           emin: float = -math.inf, emax: float = math.inf, tol: float = 1e-10,
           overextract: int = 2, refine: bool = True, window_type: str = "step",
           jackson: bool = True, mmax=None, rtol: float = 1e-12, reorth_passes: int = 1,
           extra_keep: int = 8, Q0: Optional[np.ndarray] = None, chunk: int = 64,
           verbose: bool = False, return_info: bool = False):
    """POLFED driver (``polfed.jl:1128-1158``).

    ``Q0`` is the ``n×s`` start block before orthonormalisation (the reference
    draws ``randn(n,s)`` after ``Random.seed!(seed)``, ``polfed.jl:1133-1134``).
    For a sparse ``A`` with missing bounds, Gershgorin bounds are used (the new
    ``SparseMatrixCSC`` method); the matrix-free form requires them
    (``polfed.jl:1130``).
    """
    n = op_size(A)
    if smin is None or smax is None:
        if isinstance(A, MatrixFreeRealSymOp):
            raise ValueError("Matrix-free POLFED requires spectral bounds smin and smax")
        smin, smax = gershgorin_bounds(A)
    coeffs = build_coeffs(degree, smin, smax, sigma, emin, emax, window_type, jackson)
    if mmax is None:
        mmax = math.ceil(2.8 * nev / s)
    if Q0 is None:
        Q0 = np.random.default_rng(1).standard_normal((n, s))
    Q1 = orth_block_polfed(np.array(Q0, dtype=np.float64), rtol)
    Qhist, Ablocks, Bblocks = build_full_basis(A, Q1, coeffs, smin, smax, mmax, rtol, reorth_passes)
    lam, X, res, th = extract_pairs(Qhist, Ablocks, Bblocks, A, nev, emin, emax, overextract,
                                    refine, extra_keep, chunk)
    line = progress_line(len(Qhist), lam, res)
    if verbose:
        print(line)
    info = dict(m=len(Qhist), theta=th, line=line, M=sum(a.shape[0] for a in Ablocks))
    if len(res) > 0 and len(res) >= nev:
        rs = np.sort(res)
        if rs[nev - 1] < tol:
            out = (lam[:nev], X[:, :nev], res[:nev])
            return (*out, info) if return_info else out
    if len(res) > 0:
        p = np.argsort(res, kind="stable")[:min(nev, len(res))]
        out = (lam[p], X[:, p], res[p])
    else:
        out = (np.zeros(0), np.zeros((n, 0)), np.zeros(0))
    return (*out, info) if return_info else out
