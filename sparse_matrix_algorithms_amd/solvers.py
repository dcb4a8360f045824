"""Host-side mirror of the reference's Julia API over the HIP C ABI.

Same function names, argument meaning, defaults and error behaviour as
``polfed.jl`` and ``chebyshev_subspace_iteration.jl``; every numerical step
runs in ``libsmalgo.so`` on the GPU.  Julia's ``ArgumentError`` /
``error(...)`` map to :class:`ValueError`; ``@warn`` maps to
:func:`warnings.warn`.

Operators are SciPy sparse matrices (uploaded to HBM once per call, or once
for all calls through :class:`DeviceMatrix`) or :class:`MatrixFreeRealSymOp`
(host callbacks, FP64).
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp

from . import _lib as L

_DP = C.POINTER(C.c_double)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class SmalgoError(RuntimeError):
    pass


class Context:
    """One HIP stream + rocBLAS handle on one device (``smalgo_ctx``)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.smalgo_ctx_create(device, C.byref(h))
        if rc != L.OK:
            raise SmalgoError(f"smalgo_ctx_create(device={device}) failed with status {rc} "
                              "(no HIP device?)")
        self.h = h
        self.device = device

    def check(self, rc: int):
        if rc == L.OK:
            return
        msg = self.lib.smalgo_last_error(self.h).decode()
        if rc == L.ERR_ARG:
            raise ValueError(msg)
        raise SmalgoError(f"status {rc}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.smalgo_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceMatrix:
    """A symmetric sparse matrix resident in HBM (``smalgo_csr_upload``).

    Accepts any SciPy sparse matrix; CSR row pointers are used directly (for
    the symmetric matrices of this path Julia's CSC arrays are the same
    numbers, so ``SparseMatrixCSC`` fields upload with ``index_base=1``).
    """

    def __init__(self, A, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        A = sp.csr_matrix(A, dtype=np.float64)
        A.sort_indices()
        self.n = A.shape[0]
        if A.shape[1] != self.n:
            raise ValueError("A must be square")
        ptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        idx = np.ascontiguousarray(A.indices, dtype=np.int64)
        val = np.ascontiguousarray(A.data, dtype=np.float64)
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.smalgo_csr_upload(
            self.ctx.h, self.n, ptr.ctypes.data_as(C.POINTER(C.c_int64)),
            idx.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(val), 0, C.byref(h)))
        self.h = h
        self.nnz = int(self.ctx.lib.smalgo_mat_nnz(h))

    @classmethod
    def from_file(cls, path, ctx: Optional[Context] = None) -> "DeviceMatrix":
        """Upload a CSR fixture file (``smalgo_csr_upload_file``; format in
        ``sparse_matrix_algorithms_amd.fixtures``)."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.smalgo_csr_upload_file(self.ctx.h, str(path).encode(),
                                                           C.byref(h)))
        self.h = h
        self.n = int(self.ctx.lib.smalgo_mat_n(h))
        self.nnz = int(self.ctx.lib.smalgo_mat_nnz(h))
        return self

    def is_symmetric(self) -> bool:
        s = C.c_int()
        self.ctx.check(self.ctx.lib.smalgo_mat_is_symmetric(self.h, C.byref(s)))
        return bool(s.value)

    def gershgorin_bounds(self):
        a, b = C.c_double(), C.c_double()
        self.ctx.check(self.ctx.lib.smalgo_gershgorin_bounds(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def free(self):
        if getattr(self, "h", None):
            self.ctx.lib.smalgo_mat_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class MatrixFreeRealSymOp:
    """``MatrixFreeRealSymOp(n, matvec!, matblock!)`` (``polfed.jl:148-165``).

    ``matvec(y, x)`` writes ``A x`` into ``y``; ``matblock(Y, X)`` (optional)
    writes ``A X`` into ``Y`` (``n×b`` arrays).
    """

    n: int
    matvec: Callable
    matblock: Optional[Callable] = None
    _cb: object = field(default=None, repr=False)
    _err: list = field(default_factory=list, repr=False)

    def cfunc(self):
        if self._cb is None:
            def call(user, n, b, X, ldx, Y, ldy):
                try:
                    Xa = np.ctypeslib.as_array(X, shape=(b, ldx)).T[:n]
                    Ya = np.ctypeslib.as_array(Y, shape=(b, ldy)).T[:n]
                    if self.matblock is not None:
                        self.matblock(Ya, Xa)
                    else:
                        for j in range(b):
                            y = np.zeros(n)
                            self.matvec(y, np.ascontiguousarray(Xa[:, j]))
                            Ya[:, j] = y
                    return 0
                except Exception as e:  # surfaced after the C call returns
                    self._err.append(e)
                    return 1
            self._cb = L.MATBLOCK_FN(call)
        return self._cb


def _as_operator(A, ctx):
    if isinstance(A, MatrixFreeRealSymOp):
        return None, A
    if isinstance(A, DeviceMatrix):
        return A, None
    if sp.issparse(A) or isinstance(A, np.ndarray):
        return DeviceMatrix(A, ctx), None
    raise TypeError("A must be a SciPy sparse matrix, DeviceMatrix or MatrixFreeRealSymOp")


# ---------------------------------------------------------------------------
# bounds and coefficients (polfed.jl:254-414)
# ---------------------------------------------------------------------------
def gershgorin_bounds(A, ctx: Optional[Context] = None):
    """``gershgorin_bounds(A)`` (``polfed.jl:254-266``)."""
    dm, mf = _as_operator(A, ctx)
    if mf is not None:
        raise TypeError("gershgorin_bounds has no method for MatrixFreeRealSymOp")
    return dm.gershgorin_bounds()


def jackson_kernel(K: int) -> np.ndarray:
    """``jackson_kernel(K)`` (``polfed.jl:275-284``)."""
    lib = L.load()
    g = np.zeros(max(K, 0) + 1)
    if lib.smalgo_jackson_kernel(K, _dptr(g)) != L.OK:
        raise ValueError("bad K")
    return g


def cheb_delta_coeffs(sigma_hat: float, K: int, jackson: bool = True) -> np.ndarray:
    """``cheb_delta_coeffs(σ̂,K;jackson)`` (``polfed.jl:299-313``)."""
    c = np.zeros(K + 1)
    if L.load().smalgo_cheb_delta_coeffs(sigma_hat, K, int(jackson), _dptr(c)) != L.OK:
        raise ValueError("degree must be nonnegative")
    return c


def cheb_window_coeffs(a: float, b: float, K: int, jackson: bool = True) -> np.ndarray:
    """``cheb_window_coeffs(a,b,K;jackson)`` (``polfed.jl:329-347``)."""
    c = np.zeros(K + 1)
    if L.load().smalgo_cheb_window_coeffs(a, b, K, int(jackson), _dptr(c)) != L.OK:
        raise ValueError("degree must be nonnegative")
    return c


def cheb_smooth_window_coeffs(a: float, b: float, K: int, jackson: bool = True,
                              nsamp: int = 4096, delta: float = 0.02) -> np.ndarray:
    """``cheb_smooth_window_coeffs(a,b,K;jackson,nsamp,δ)`` (``polfed.jl:364-388``)."""
    c = np.zeros(K + 1)
    if L.load().smalgo_cheb_smooth_window_coeffs(a, b, K, int(jackson), nsamp, delta,
                                                 _dptr(c)) != L.OK:
        raise ValueError("degree must be nonnegative")
    return c


def build_coeffs(degree: int, smin: float, smax: float, sigma: float, emin: float, emax: float,
                 window_type: str = "step", jackson: bool = True) -> np.ndarray:
    """``_build_coeffs`` (``polfed.jl:404-414``)."""
    if window_type not in L.WINDOWS:
        raise ValueError(f"Unknown window type: {window_type}")
    c = np.zeros(degree + 1)
    if L.load().smalgo_build_coeffs(degree, smin, smax, sigma, emin, emax, L.WINDOWS[window_type],
                                    int(jackson), _dptr(c)) != L.OK:
        raise ValueError("degree must be nonnegative")
    return c


def apply_poly_block(A, Q: np.ndarray, coeffs: np.ndarray, smin: float, smax: float,
                     precision: str = "fp64", ctx: Optional[Context] = None) -> np.ndarray:
    """``apply_poly_block!(Y,A,Q,coeffs,smin,smax,work)`` (``polfed.jl:464-545``):
    ``Y = p(Â) Q`` on the GPU; returns ``Y`` (FP64, ``n×b``)."""
    ctx = ctx or default_context()
    dm, mf = _as_operator(A, ctx)
    Qf = np.asfortranarray(Q, dtype=np.float64)
    n, b = Qf.shape
    co = np.ascontiguousarray(coeffs, dtype=np.float64)
    Y = np.zeros((n, b), order="F")
    if mf is not None:
        rc = ctx.lib.smalgo_apply_poly_block_mf(ctx.h, n, mf.cfunc(), None, b, _dptr(Qf), n,
                                                _dptr(co), len(co) - 1, smin, smax, _dptr(Y), n)
        if mf._err:
            raise mf._err.pop()
    else:
        rc = ctx.lib.smalgo_apply_poly_block(ctx.h, dm.h, b, _dptr(Qf), n, _dptr(co), len(co) - 1,
                                             smin, smax, L.PRECISIONS[precision], _dptr(Y), n)
    ctx.check(rc)
    return Y


# ---------------------------------------------------------------------------
# results
# ---------------------------------------------------------------------------
_STATS = ("t_total", "t_filter", "t_orth", "t_extract", "t_refine", "filter_steps",
          "filter_launches",
          "block_steps", "basis_dim", "iterations", "spmm_bytes", "gemm_flops",
          "refine_filter_steps", "locked", "smin", "smax", "warn_interval", "tile_vec",
          "tile_lpr", "tile_rt", "tile_lds", "tile_split", "filter_split_steps", "slice_paired",
          "slice_panel_width", "slice_replans", "slice_graphs",
          "slice_filter_bytes", "slice_gemm_flops", "slice_launches", "refine_iterations",
          "t_refine_host")


def _take_result(ctx: Context, h, want_X: bool = True):
    lib = ctx.lib
    try:
        k = int(lib.smalgo_result_count(h))
        n = int(lib.smalgo_result_n(h))
        lam = np.zeros(k)
        res = np.zeros(k)
        X = np.zeros((n, k), order="F") if want_X else None
        ctx.check(lib.smalgo_result_copy(h, _dptr(lam), _dptr(X) if want_X and k else None, n,
                                         _dptr(res)))
        stats = {}
        v = C.c_double()
        for name in _STATS:
            if lib.smalgo_result_stat(h, name.encode(), C.byref(v)) == L.OK:
                stats[name] = v.value
        stats["line"] = lib.smalgo_result_progress_line(h).decode()
        return lam, X, res, stats
    finally:
        lib.smalgo_result_free(h)


# ---------------------------------------------------------------------------
# block kernels of the reference, one call each (polfed.jl:561-880)
# ---------------------------------------------------------------------------
def orth_against(W: np.ndarray, Q: np.ndarray, precision: str = "fp64",
                 ctx: Optional[Context] = None) -> np.ndarray:
    """``orth_against!(W,Q)``: returns ``W − Q(QᵀW)`` (``polfed.jl:605-609``)."""
    Wf = np.array(W, dtype=np.float64, order="F")
    Qf = np.asfortranarray(Q, dtype=np.float64)
    if Wf.ndim != 2 or Qf.ndim != 2 or Qf.shape[0] != Wf.shape[0]:
        raise ValueError(f"Q must have as many rows as W ({Wf.shape[0]})")
    ctx = ctx or default_context()
    n, b = Wf.shape
    ctx.check(ctx.lib.smalgo_orth_against(ctx.h, n, b, _dptr(Wf), n, Qf.shape[1], _dptr(Qf), n,
                                          L.PRECISIONS[precision]))
    return Wf


def block_qr_factor(W: np.ndarray, rtol: float = 1e-12, precision: str = "fp64",
                    ctx: Optional[Context] = None):
    """``block_qr_factor(W; rtol)`` → ``(Q[:, :r], R[:r, :])`` (``polfed.jl:644-685``)."""
    ctx = ctx or default_context()
    Wf = np.asfortranarray(W, dtype=np.float64)
    n, b = Wf.shape
    Q = np.zeros((n, b), order="F")
    R = np.zeros((b, b), order="F")
    r = C.c_int()
    ctx.check(ctx.lib.smalgo_block_qr(ctx.h, n, b, _dptr(Wf), n, rtol, L.PRECISIONS[precision],
                                      _dptr(Q), n, _dptr(R), b, C.byref(r)))
    k = r.value
    return np.ascontiguousarray(Q[:, :k]), np.array(R[:k, :])


def orth_block_polfed(Q: np.ndarray, rtol: float = 1e-12, ctx: Optional[Context] = None):
    """``orth_block_polfed!(Q; rtol)`` (``polfed.jl:561-594``)."""
    return block_qr_factor(Q, rtol, ctx=ctx)[0]


def build_X(V: np.ndarray, U: np.ndarray, ctx: Optional[Context] = None) -> np.ndarray:
    """``build_X!`` with ``V=[Q_1 … Q_m]`` given as one ``n×M`` block (``polfed.jl:767-776``)."""
    ctx = ctx or default_context()
    Vf = np.asfortranarray(V, dtype=np.float64)
    Uf = np.asfortranarray(U, dtype=np.float64)
    n, M = Vf.shape
    k = Uf.shape[1]
    X = np.zeros((n, k), order="F")
    ctx.check(ctx.lib.smalgo_build_X(ctx.h, n, M, k, _dptr(Vf), n, _dptr(Uf), M, _dptr(X), n))
    return X


def _op_args(A, ctx):
    """(matrix handle, callback, n, operator object) — the caller keeps the
    returned operator object alive for the duration of the C call."""
    dm, mf = _as_operator(A, ctx)
    if mf is not None:
        return None, mf.cfunc(), mf.n, mf
    return dm.h, L.MATBLOCK_FN(), dm.n, dm


def rayleigh(A, X: np.ndarray, ctx: Optional[Context] = None):
    """``rayleigh!`` → ``(λ, res)`` (``polfed.jl:791-830``)."""
    ctx = ctx or default_context()
    h, fn, n, mf = _op_args(A, ctx)
    Xf = np.asfortranarray(X, dtype=np.float64)
    k = Xf.shape[1]
    lam, res = np.zeros(k), np.zeros(k)
    rc = ctx.lib.smalgo_rayleigh(ctx.h, h, fn, None, n, k, _dptr(Xf), n, _dptr(lam), _dptr(res))
    if isinstance(mf, MatrixFreeRealSymOp) and mf._err:
        raise mf._err.pop()
    ctx.check(rc)
    return lam, res


def residuals(A, X: np.ndarray, lam: np.ndarray, ctx: Optional[Context] = None) -> np.ndarray:
    """``residuals!`` (``chebyshev_subspace_iteration.jl:311-344``)."""
    ctx = ctx or default_context()
    h, fn, n, mf = _op_args(A, ctx)
    Xf = np.asfortranarray(X, dtype=np.float64)
    k = Xf.shape[1]
    la = np.ascontiguousarray(lam, dtype=np.float64)
    res = np.zeros(k)
    rc = ctx.lib.smalgo_residuals(ctx.h, h, fn, None, n, k, _dptr(Xf), n, _dptr(la), _dptr(res))
    if isinstance(mf, MatrixFreeRealSymOp) and mf._err:
        raise mf._err.pop()
    ctx.check(rc)
    return res


def refine_original_rr(A, X: np.ndarray, ctx: Optional[Context] = None):
    """``refine_original_rr(A,X)`` → ``(λ, Y, res)`` sorted by λ (``polfed.jl:837-880``)."""
    ctx = ctx or default_context()
    h, fn, n, mf = _op_args(A, ctx)
    Xf = np.asfortranarray(X, dtype=np.float64)
    out = C.c_void_p()
    rc = ctx.lib.smalgo_refine_original_rr(ctx.h, h, fn, None, n, Xf.shape[1], _dptr(Xf), n,
                                           C.byref(out))
    if isinstance(mf, MatrixFreeRealSymOp) and mf._err:
        raise mf._err.pop()
    ctx.check(rc)
    lam, Y, res, _ = _take_result(ctx, out)
    return lam, Y, res


# ---------------------------------------------------------------------------
# POLFED (polfed.jl:1128-1186)
# ---------------------------------------------------------------------------
def _start_block(Q0, n: int, m0: int):
    """The caller's ``randn(n, m0)`` replacement (``chebyshev_subspace_iteration.jl:701-702``)
    as an ``n × min(m0, n)`` column-major array; the library reads exactly that many
    values, so a smaller array is an error here, before anything reaches the device."""
    if Q0 is None:
        return None
    Q0 = np.asarray(Q0)
    k = min(m0, n)
    if Q0.ndim != 2 or Q0.shape[0] != n or Q0.shape[1] < k:
        raise ValueError(f"Q0 must be {n} x (at least) {k}, got {Q0.shape}")
    return np.asfortranarray(Q0[:, :k], dtype=np.float64)


def _n_of(A) -> int:
    if isinstance(A, (MatrixFreeRealSymOp, DeviceMatrix)):
        return int(A.n)
    return int(A.shape[0])


def _polfed_params(ctx, s, smin, smax, sigma, emin, emax, tol, overextract, refine, window_type,
                   jackson, mmax, rtol, reorth_passes, extra_keep, seed, chunk, show_progress,
                   precision, refine_guard, refine_degree, refine_maxiter, refine_tol,
                   refine_guard_degree=None):
    if window_type not in L.WINDOWS:
        raise ValueError(f"Unknown window type: {window_type}")
    p = L.PolfedParams()
    ctx.lib.smalgo_polfed_default_params(C.byref(p))
    p.s = s
    if smin is not None and smax is not None:
        p.has_bounds, p.smin, p.smax = 1, float(smin), float(smax)
    p.sigma, p.emin, p.emax, p.tol = sigma, emin, emax, tol
    p.overextract, p.refine = overextract, int(refine)
    p.window_type, p.jackson = L.WINDOWS[window_type], int(jackson)
    p.mmax = mmax if mmax is not None else 0
    p.rtol, p.reorth_passes, p.extra_keep = rtol, reorth_passes, extra_keep
    p.seed, p.chunk, p.show_progress = seed, chunk, int(show_progress)
    p.precision = L.PRECISIONS[precision]
    p.refine_guard = -1 if refine_guard is None else refine_guard
    p.refine_degree = -1 if refine_degree is None else refine_degree
    p.refine_maxiter = refine_maxiter
    p.refine_tol = -1.0 if refine_tol is None else refine_tol
    p.refine_guard_degree = -1 if refine_guard_degree is None else refine_guard_degree
    return p


def polfed(A, nev: int, degree: int, s: int = 8, smin=None, smax=None, sigma: float = 0.0,
           emin: float = -math.inf, emax: float = math.inf, tol: float = 1e-10,
           threaded: bool = True, overextract: int = 2, refine: bool = True,
           window_type: str = "step", jackson: bool = True, mmax: Optional[int] = None,
           rtol: float = 1e-12, reorth_passes: int = 1, extra_keep: int = 8, seed: int = 1,
           show_progress: bool = False, chunk: int = 64, Q1: Optional[np.ndarray] = None,
           precision: str = "fp64", refine_guard: Optional[int] = None,
           refine_degree: Optional[int] = None, refine_maxiter: int = 3,
           refine_tol: Optional[float] = None, refine_guard_degree: Optional[int] = None,
           return_stats: bool = False, ctx: Optional[Context] = None, want_X: bool = True):
    """``polfed(A, nev, degree; kwargs...)`` on the GPU.

    ``A``: SciPy sparse / :class:`DeviceMatrix` (the sparse method; bounds
    default to Gershgorin) or :class:`MatrixFreeRealSymOp` (bounds required,
    ``polfed.jl:1130``).  ``Q1`` is the ``n×s`` start block before
    orthonormalisation (the reference's ``randn(n,s)``); without it the
    library draws one from ``seed``.  ``threaded`` is accepted and unused, as
    in the reference.  Returns ``(λ, X, res)`` (plus a stats dict with
    ``return_stats``) with the reference's return order rule
    (``polfed.jl:1148-1157``).
    """
    n = _n_of(A)
    q = None
    if Q1 is not None:
        q = np.asfortranarray(Q1, dtype=np.float64)
        if q.shape != (n, s):
            raise ValueError(f"Q1 must be {n}x{s}")
    ctx = ctx or default_context()
    dm, mf = _as_operator(A, ctx)
    p = _polfed_params(ctx, s, smin, smax, sigma, emin, emax, tol, overextract, refine,
                       window_type, jackson, mmax, rtol, reorth_passes, extra_keep, seed, chunk,
                       show_progress, precision, refine_guard, refine_degree, refine_maxiter,
                       refine_tol, refine_guard_degree)
    h = C.c_void_p()
    if mf is not None:
        rc = ctx.lib.smalgo_polfed(ctx.h, None, mf.cfunc(), None, n, nev, degree, C.byref(p),
                                   _dptr(q) if q is not None else None, C.byref(h))
        if mf._err:
            raise mf._err.pop()
    else:
        rc = ctx.lib.smalgo_polfed(ctx.h, dm.h, L.MATBLOCK_FN(), None, n, nev, degree, C.byref(p),
                                   _dptr(q) if q is not None else None, C.byref(h))
    ctx.check(rc)
    lam, X, res, stats = _take_result(ctx, h, want_X)
    if show_progress:
        print(stats["line"])
    return (lam, X, res, stats) if return_stats else (lam, X, res)


def polfed_mf(matvec: Callable, n: int, nev: int, degree: int, matblock: Optional[Callable] = None,
              **kwargs):
    """``polfed(matvec!, n, nev, degree; matblock!, kwargs...)`` (``polfed.jl:1183-1186``)."""
    return polfed(MatrixFreeRealSymOp(n, matvec, matblock), nev, degree, **kwargs)


def subspace_rr(A, Y: np.ndarray, return_Q: bool = False, ctx: Optional[Context] = None):
    """The Rayleigh–Ritz tail of the subspace iterations (Chebyshev ``:733-748``,
    FEAST ``feast.jl:540-553``, Zolotarev ``zoloterov.jl:1698-1706``, real
    symmetric case) on the GPU: returns ``(λ, X, res)`` for all columns in
    ascending λ, plus ``Q = thin_orth(Y)`` with ``return_Q``."""
    ctx = ctx or default_context()
    dm, mf = _as_operator(A, ctx)
    Yf = np.asfortranarray(Y, dtype=np.float64)
    n, k = Yf.shape
    Q = np.zeros((n, k), order="F") if return_Q else None
    h = C.c_void_p()
    if mf is not None:
        rc = ctx.lib.smalgo_subspace_rr(ctx.h, None, mf.cfunc(), None, n, k, _dptr(Yf), n,
                                        _dptr(Q) if return_Q else None, n, C.byref(h))
        if mf._err:
            raise mf._err.pop()
    else:
        rc = ctx.lib.smalgo_subspace_rr(ctx.h, dm.h, L.MATBLOCK_FN(), None, n, k, _dptr(Yf), n,
                                        _dptr(Q) if return_Q else None, n, C.byref(h))
    ctx.check(rc)
    lam, X, res, _ = _take_result(ctx, h)
    return (lam, X, res, Q) if return_Q else (lam, X, res)


# ---------------------------------------------------------------------------
# batched spectral slices (no reference counterpart)
# ---------------------------------------------------------------------------
def apply_poly_slices(A, blocks, coeffs, smin: float, smax: float, precision: str = "fp64",
                      panel_width: int = 0, use_graph: bool = True, replays: int = 1,
                      return_stats: bool = False, ctx: Optional[Context] = None):
    """``Y_i = p_i(Â) Q_i`` for several slices in one packed filter.

    ``blocks``: list over replays of lists over slices of ``n×b_i`` arrays
    (or one list over slices for a single replay); ``coeffs[i]``: Chebyshev
    coefficients of slice ``i`` (degree ``len-1``).  Every replay runs the
    same plan (and the same HIP graphs with ``use_graph``).  Returns the same
    nesting of filtered blocks (plus the plan stats with ``return_stats``).
    """
    ctx = ctx or default_context()
    dm, mf = _as_operator(A, ctx)
    if mf is not None:
        raise ValueError("batched slices need a sparse matrix")
    single = isinstance(blocks[0], np.ndarray)
    reps = [blocks] if single else list(blocks)
    ns = len(reps[0])
    widths = [int(b.shape[1]) for b in reps[0]]
    n = dm.n
    wtot = sum(widths)
    Q = np.zeros((n, wtot * len(reps)), order="F")
    for r, rep in enumerate(reps):
        if [int(b.shape[1]) for b in rep] != widths:
            raise ValueError("every replay needs the same slice widths")
        Q[:, r * wtot:(r + 1) * wtot] = np.hstack(rep)
    Y = np.zeros_like(Q)
    cos = [np.ascontiguousarray(c, dtype=np.float64) for c in coeffs]
    if len(cos) != ns:
        raise ValueError("one coefficient vector per slice")
    co = np.concatenate(cos)
    w = (C.c_int * ns)(*widths)
    dg = (C.c_int * ns)(*[len(c) - 1 for c in cos])
    st = np.zeros(6)
    ctx.check(ctx.lib.smalgo_apply_poly_slices(
        ctx.h, dm.h, ns, w, dg, _dptr(co), float(smin), float(smax), L.PRECISIONS[precision],
        int(panel_width), int(use_graph), len(reps), _dptr(Q), n, _dptr(Y), n, _dptr(st)))
    outs = []
    for r in range(len(reps)):
        off = r * wtot
        one = []
        for b in widths:
            one.append(np.array(Y[:, off:off + b]))
            off += b
        outs.append(one)
    res = outs[0] if single else outs
    if return_stats:
        keys = ("panel_width", "panels", "launches", "ms_per_run", "bytes_per_run", "graphs")
        return res, dict(zip(keys, st.tolist()))
    return res


def polfed_slices(A, slices, s: int = 8, smin=None, smax=None, tol: float = 1e-10,
                  overextract: int = 2, refine: bool = True, window_type: str = "step",
                  jackson: bool = True, rtol: float = 1e-12, reorth_passes: int = 1,
                  extra_keep: int = 8, seed: int = 1, precision: str = "fp64",
                  refine_guard: Optional[int] = None, refine_degree: Optional[int] = None,
                  refine_maxiter: int = 3, refine_tol: Optional[float] = None,
                  refine_guard_degree: Optional[int] = None, use_graph: bool = True,
                  show_progress: bool = False, want_X: bool = True,
                  ctx: Optional[Context] = None):
    """POLFED over many intervals of one sparse matrix in one batched run.

    ``slices``: list of dicts with ``emin, emax, nev, degree`` and optional
    ``s`` (default ``s``), ``sigma`` (default the interval midpoint) and
    ``mmax``.  Slice ``i`` is ``polfed(A, nev_i, degree_i; emin_i, emax_i, …)``
    (``polfed.jl:1128-1158``) with start-block seed ``seed + i``; the filters of
    all slices share packed launches (HIP graphs with ``use_graph``).  Returns a
    list of ``(λ, X, res, stats)`` per slice.
    """
    ctx = ctx or default_context()
    dm, mf = _as_operator(A, ctx)
    if mf is not None:
        raise ValueError("batched slices need a sparse matrix")
    p = _polfed_params(ctx, s, smin, smax, 0.0, -math.inf, math.inf, tol, overextract, refine,
                       window_type, jackson, None, rtol, reorth_passes, extra_keep, seed, 64,
                       show_progress, precision, refine_guard, refine_degree, refine_maxiter, refine_tol,
                       refine_guard_degree)
    ns = len(slices)
    arr = (L.SliceParams * ns)()
    for i, d in enumerate(slices):
        a = arr[i]
        a.emin, a.emax = float(d["emin"]), float(d["emax"])
        a.sigma = float(d.get("sigma", 0.5 * (a.emin + a.emax)))
        a.nev, a.s, a.degree = int(d["nev"]), int(d.get("s", s)), int(d["degree"])
        a.mmax = int(d.get("mmax") or 0)
    hs = (C.c_void_p * ns)()
    ctx.check(ctx.lib.smalgo_polfed_slices(ctx.h, dm.h, ns, arr, C.byref(p), int(use_graph), hs))
    out = []
    try:
        for i in range(ns):
            h, hs[i] = hs[i], None
            out.append(_take_result(ctx, C.c_void_p(h), want_X))
    finally:
        for i in range(ns):
            if hs[i]:
                ctx.lib.smalgo_result_free(C.c_void_p(hs[i]))
    return out


# ---------------------------------------------------------------------------
# Chebyshev subspace iteration (chebyshev_subspace_iteration.jl:667-848)
# ---------------------------------------------------------------------------
def _cheb_params(ctx, smin, smax, sigma, m0, degree, maxiter, tol, res_gate, jackson, debug,
                 two_hit, seed, window_type):
    if window_type not in L.WINDOWS:
        raise ValueError("Unknown window type, possible are :step, :delta, and :smooth")
    p = L.ChebParams()
    ctx.lib.smalgo_cheb_default_params(C.byref(p))
    if smin is not None and smax is not None:
        p.has_bounds, p.smin, p.smax = 1, float(smin), float(smax)
    p.sigma, p.m0, p.degree, p.maxiter = sigma, m0, degree, maxiter
    p.tol, p.res_gate, p.jackson, p.debug = tol, res_gate, int(jackson), int(debug)
    p.two_hit, p.seed, p.window_type = int(two_hit), seed, L.WINDOWS[window_type]
    return p


def chebyshev_poly_real_symm_sparse(A, emin: float, emax: float, smin=None, smax=None,
                                    sigma: float = 0.0, m0: int = 200, degree: int = 80,
                                    maxiter: int = 8, tol: float = 1e-10, res_gate: float = 1e-6,
                                    jackson: bool = True, debug: bool = False,
                                    two_hit: bool = True, seed: int = 1,
                                    window_type: str = "step", Q0: Optional[np.ndarray] = None,
                                    return_stats: bool = False, ctx: Optional[Context] = None):
    """Sparse form (``chebyshev_subspace_iteration.jl:667-757``): always runs
    ``maxiter`` iterations; returns ``(λsel, Xsel, ressel)`` sorted by λ."""
    q0 = _start_block(Q0, _n_of(A), m0)
    ctx = ctx or default_context()
    dm, mf = _as_operator(A, ctx)
    if mf is not None:
        raise TypeError("use chebyshev_poly_real_symm_sparse_mf for matrix-free operators")
    if not dm.is_symmetric():
        raise ValueError("Chebyshev filtered solver requires a real symmetric matrix")
    if not emin < emax:
        raise ValueError("need emin<emax")
    if not m0 > 0:
        raise ValueError("m0 must be positive")
    if not degree >= 0:
        raise ValueError("degree must be nonnegative")
    p = _cheb_params(ctx, smin, smax, sigma, m0, degree, maxiter, tol, res_gate, jackson, debug,
                     two_hit, seed, window_type)
    q = None
    if Q0 is not None:
        q = q0
    h = C.c_void_p()
    ctx.check(ctx.lib.smalgo_chebyshev(ctx.h, dm.h, L.MATBLOCK_FN(), None, dm.n, emin, emax,
                                       C.byref(p), _dptr(q) if q is not None else None,
                                       C.byref(h)))
    lam, X, res, stats = _take_result(ctx, h)
    if stats.get("warn_interval"):
        warnings.warn("target interval not contained in supplied spectral bounds; "
                      "filter quality may degrade")
    return (lam, X, res, stats) if return_stats else (lam, X, res)


def chebyshev_poly_real_symm_sparse_mf(matvec: Callable, n: int, emin: float, emax: float,
                                       matblock: Optional[Callable] = None, smin=None, smax=None,
                                       sigma: float = 0.0, m0: int = 200, degree: int = 80,
                                       maxiter: int = 8, tol: float = 1e-10,
                                       res_gate: float = 1e-6, jackson: bool = True,
                                       debug: bool = False, two_hit: bool = True, seed: int = 1,
                                       window_type: str = "step",
                                       Q0: Optional[np.ndarray] = None,
                                       return_stats: bool = False,
                                       ctx: Optional[Context] = None):
    """``matvec!`` form (``chebyshev_subspace_iteration.jl:771-848``): bounds are
    required; ``two_hit`` early exit."""
    q0 = _start_block(Q0, n, m0)
    ctx = ctx or default_context()
    if smin is None or smax is None:
        raise ValueError("matrix-free form requires smin and smax")
    op = MatrixFreeRealSymOp(n, matvec, matblock)
    p = _cheb_params(ctx, smin, smax, sigma, m0, degree, maxiter, tol, res_gate, jackson, debug,
                     two_hit, seed, window_type)
    q = None
    if Q0 is not None:
        q = q0
    h = C.c_void_p()
    rc = ctx.lib.smalgo_chebyshev(ctx.h, None, op.cfunc(), None, n, emin, emax, C.byref(p),
                                  _dptr(q) if q is not None else None, C.byref(h))
    if op._err:
        raise op._err.pop()
    ctx.check(rc)
    lam, X, res, stats = _take_result(ctx, h)
    return (lam, X, res, stats) if return_stats else (lam, X, res)
