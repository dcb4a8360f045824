"""Learned POLFED filter parameters.

A two-layer perceptron (68 → 128 ReLU → 2), trained and evaluated with the
library's HIP kernels (``smalgo_mlp_*``), maps 64 KPM moments of the matrix
plus the target interval (both in the Chebyshev variable), log(nev/s) and the
log stochastic count to the filter
resolution κ (degree K = ⌈κ/h⌉, h = interval half-width in the Chebyshev
variable) and the basis factor f (mmax = ⌈f·nev/s⌉).  The training labels are
the cheapest (κ, f) of a cost-ordered ladder that reproduces the plain
method's converged eigenpairs, logged by ``learned_data.py``.

Fallback: when the prediction would give a Krylov basis smaller than twice the
stochastic eigenvalue count of the interval (the reference extracts 2·nev Ritz
candidates, ``polfed.jl:935``), the plain parameter rule is used instead.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib as L
from .kpm import count as kpm_count
from .kpm import plain_parameters
from .solvers import Context, default_context

_FP = C.POINTER(C.c_float)

NMOM = 64
# safety margins are calibrated per bucket of predicted block steps mmax
MARGIN_BUCKETS = (10.0, 16.0)


def margin_bucket(blocks: float) -> int:
    return int(sum(blocks >= b for b in MARGIN_BUCKETS))
KAPPA_RANGE = (3.0, 10.0)
FACTOR_RANGE = (1.6, 7.0)
WEIGHTS = Path(__file__).resolve().parent / "data" / "mlp_weights.json"
# Learned degree and basis-size factors of the mixed path (basis size rounded
# up), chosen on the 256 training records only: of six (degree, basis) pairs
# from (0.6, 1.08) to (1.0, 1.0), (0.7, 1.08) converges the most training
# instances (254) at the highest eigenpairs/s (data/mixed_scales_train.jsonl).
MIXED_DEGREE_SCALE = 0.7
MIXED_MMAX_SCALE = 1.08


def _fp(a):
    return a.ctypes.data_as(_FP)


class MLP:
    """Handle on a device-resident perceptron (``smalgo_mlp``)."""

    def __init__(self, nin: int = NMOM + 4, nh: int = 128, nout: int = 2, seed: int = 1,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.nin, self.nh, self.nout = nin, nh, nout
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.smalgo_mlp_create(self.ctx.h, nin, nh, nout, seed, C.byref(h)))
        self.h = h
        self.np = int(self.ctx.lib.smalgo_mlp_num_params(h))

    def params(self) -> np.ndarray:
        p = np.zeros(self.np, dtype=np.float32)
        self.ctx.check(self.ctx.lib.smalgo_mlp_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.np
        self.ctx.check(self.ctx.lib.smalgo_mlp_set_params(self.h, _fp(p)))

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.nout), dtype=np.float32)
        self.ctx.check(self.ctx.lib.smalgo_mlp_predict(self.h, _fp(X), X.shape[0], _fp(Y)))
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.np, dtype=np.float32)
        loss = C.c_float()
        self.ctx.check(self.ctx.lib.smalgo_mlp_gradient(self.h, _fp(X), _fp(T), X.shape[0], _fp(g),
                                                        C.byref(loss)))
        return loss.value, g

    def train(self, X, T, epochs: int, lr: float = 1e-3):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(max(epochs, 1), dtype=np.float32)
        loss = C.c_float()
        self.ctx.check(self.ctx.lib.smalgo_mlp_train(self.h, _fp(X), _fp(T), X.shape[0], epochs, lr,
                                                     _fp(hist), C.byref(loss)))
        return hist[:epochs]

    def save(self, path: Path, meta: dict):
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        Path(path).write_text(json.dumps(dict(nin=self.nin, nh=self.nh, nout=self.nout,
                                              params=self.params().tolist(), **meta)))

    @classmethod
    def load(cls, path: Path = WEIGHTS, ctx: Optional[Context] = None):
        d = json.loads(Path(path).read_text())
        m = cls(d["nin"], d["nh"], d["nout"], ctx=ctx)
        m.set_params(np.array(d["params"], dtype=np.float32))
        m.meta = {k: v for k, v in d.items() if k != "params"}
        return m

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.lib.smalgo_mlp_free(self.h)
                self.h = None
        except Exception:
            pass


NIN = NMOM + 4


def features(mu: np.ndarray, emin: float, emax: float, smin: float, smax: float, n: int,
             nev: int, s: int) -> np.ndarray:
    """68 inputs: the Jackson-damped moments μ_0..μ_63 of Â, the interval
    (centre and log half-width in the Chebyshev variable), and two quantities
    derived from those and the request: log(nev/s), the number of block steps
    per unit basis factor, and log of the stochastic eigenvalue count."""
    from .solvers import jackson_kernel
    c0, d = 0.5 * (smin + smax), 0.5 * (smax - smin)
    a, b = (emin - c0) / d, (emax - c0) / d
    g = jackson_kernel(NMOM - 1)
    m = np.asarray(mu[:NMOM], dtype=np.float64) * g
    h = 0.5 * (b - a)
    est = max(kpm_count(mu, n, emin, emax, smin, smax), 1.0)
    return np.concatenate([m, [0.5 * (a + b), math.log(h), math.log(nev / s),
                               math.log(est)]]).astype(np.float32)


def encode_targets(kappa: float, factor: float) -> np.ndarray:
    return np.array([(kappa - 5.0) / 2.0, (factor - 2.4) / 0.4], dtype=np.float32)


def decode_targets(y) -> tuple[float, float]:
    kappa = float(np.clip(5.0 + 2.0 * y[0], *KAPPA_RANGE))
    factor = float(np.clip(2.4 + 0.4 * y[1], *FACTOR_RANGE))
    return kappa, factor


def choose_parameters(model: Optional[MLP], mu: np.ndarray, n: int, nev: int, emin: float,
                      emax: float, smin: float, smax: float, s: int, margin_kappa=None,
                      margin_factor=None, precision: str = "fp64"):
    """(degree, mmax, source) from the network plus its calibrated safety
    margins, or from the plain rule when the predicted basis undercounts the
    stochastic estimate (or no model).  For the mixed path (precision
    "fp32") the learned degree is scaled by MIXED_DEGREE_SCALE and the basis
    size by MIXED_MMAX_SCALE: its FP64 refinement is one more filtered
    subspace step on the candidates, so the FP32 basis converges with a lower
    degree, and the refinement's cost falls with the degree while the basis
    grows only by a block or two (factors chosen on the training records;
    held-out result in data/heldout_rules.jsonl).  The plain-rule answers (no model, or
    fallback) take the plain rule's own mixed-path scale."""
    K0, m0 = plain_parameters(n, nev, emin, emax, smin, smax, s, precision=precision)
    if model is None:
        return K0, m0, "plain"
    meta = getattr(model, "meta", {})
    y = model.predict(features(mu, emin, emax, smin, smax, n, nev, s)[None, :])[0]
    kappa, factor = decode_targets(y)
    margins = meta.get("margins")
    if margins is not None:
        mk, mf = margins[margin_bucket(factor * nev / s)]
    else:
        mk, mf = 0.0, 0.0
    kappa += mk if margin_kappa is None else margin_kappa
    factor += mf if margin_factor is None else margin_factor
    K, mmax = plain_parameters(n, nev, emin, emax, smin, smax, s, kappa, factor)
    est = kpm_count(mu, n, emin, emax, smin, smax)
    if mmax * s < 2.0 * est:
        return K0, m0, "fallback"
    if precision == "fp32":
        K = max(1, math.ceil(MIXED_DEGREE_SCALE * K))
        mmax = max(1, math.ceil(MIXED_MMAX_SCALE * mmax - 1e-9))
    return K, mmax, "learned"


def under_converged(lam, res, emin: float, emax: float, anorm: float,
                    tol_rel: float = 1e-10) -> bool:
    """True when a returned pair inside [emin, emax] has a residual between
    tol_rel·‖A‖ and 1e-6·‖A‖: a real eigenpair the filter did not finish.
    Spurious Ritz values inside an interval sit near 1e-3·‖A‖ and do not
    count."""
    lam = np.asarray(lam)
    res = np.asarray(res)
    inside = (lam >= emin) & (lam <= emax)
    return bool(np.any(inside & (res > tol_rel * anorm) & (res < 1e-6 * anorm)))


def polfed_slices_learned(A, slices, mu: np.ndarray, model: Optional[MLP], smin: float,
                          smax: float, precision: str = "fp32", tol_rel: float = 1e-10,
                          retry: bool = True, **kw):
    """Batched POLFED slices (``solvers.polfed_slices``) with learned filter
    parameters, and the plain rule for every slice whose learned degree fell
    short.

    ``slices``: dicts with ``emin, emax, nev, s`` and optional ``sigma``;
    ``mu``: KPM moments of ``A`` on [smin, smax]; ``kw`` goes to
    ``polfed_slices``.  The learned answer can under-shoot a slice's degree
    where the matrix lies outside the training range; such a slice returns
    real eigenpairs inside its interval whose residuals sit above the
    tolerance (``under_converged``).  With ``retry`` those slices are solved
    again, in one more batch, with the plain rule.  Without a model every
    slice takes the plain rule and nothing is retried.

    Returns ``(outs, params, info)``: one ``(λ, X, res, stats)`` per slice,
    the parameters each answer was computed with (``source``: "learned",
    "fallback", "plain" or "plain retry"), and ``info`` with the retried
    indices (``retried``), their first answers (``first``) and the wall time of
    each batch (``seconds``).
    """
    import time

    from .solvers import polfed_slices
    n = A.shape[0] if hasattr(A, "shape") else A.n
    anorm = max(abs(smin), abs(smax))

    def run(ps):
        t = time.perf_counter()
        o = polfed_slices(A, [{k: v for k, v in p.items() if k != "source"} for p in ps],
                          smin=smin, smax=smax, tol=tol_rel * anorm, precision=precision, **kw)
        return o, time.perf_counter() - t

    params = []
    for d in slices:
        K, mm, src = choose_parameters(model, mu[:NMOM], n, d["nev"], d["emin"], d["emax"],
                                       smin, smax, d["s"], precision=precision)
        params.append(dict(d, degree=K, mmax=mm, source=src))
    outs, dt = run(params)
    info = dict(retried=[], first={}, seconds=[dt])
    if retry and model is not None:
        info["retried"] = [i for i, (d, o) in enumerate(zip(params, outs))
                           if under_converged(o[0], o[2], d["emin"], d["emax"], anorm, tol_rel)]
    if info["retried"]:
        plain = []
        for i in info["retried"]:
            d = params[i]
            K, mm = plain_parameters(n, d["nev"], d["emin"], d["emax"], smin, smax, d["s"],
                                     precision=precision)
            plain.append(dict(d, degree=K, mmax=mm, source="plain retry"))
        outs2, dt2 = run(plain)
        info["seconds"].append(dt2)
        for j, i in enumerate(info["retried"]):
            info["first"][i] = outs[i]
            outs[i], params[i] = outs2[j], plain[j]
    return outs, params, info
