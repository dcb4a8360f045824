"""Training data, training and held-out evaluation of the learned POLFED
parameters on seeded 3D Anderson instances (L ∈ {16..40}, W ∈ [2, 20]).

    python -m sparse_matrix_algorithms_amd.learned_data gen   --first 0 --last 110 --out runs.jsonl
    python -m sparse_matrix_algorithms_amd.learned_data train --data runs.jsonl
    python -m sparse_matrix_algorithms_amd.learned_data eval  --data runs.jsonl

``gen`` logs, per instance, the plain method's run (plain parameter rule,
FP64) — eigenpairs, residuals, block steps, time — and then walks a
cost-ordered ladder of (κ, f) settings, recording the cheapest one that
returns the plain run's eigenvalue set at relative residual ≤ 1e-10.
Instances 0..255 train, 256..319 are held out.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / "data"
LADDER = sorted([(4.0, 1.8), (4.0, 2.2), (5.0, 1.8), (5.0, 2.2), (4.0, 2.8), (6.0, 2.2),
                 (5.0, 2.8), (6.0, 2.8), (8.0, 2.4)], key=lambda t: t[0] * t[1])
# when the plain rule does not converge the window (few block steps), larger
# bases are tried in cost order
RESCUE = sorted([(5.0, 3.4), (8.0, 3.4), (5.0, 4.2), (8.0, 4.2), (5.0, 5.2), (8.0, 5.2),
                 (6.0, 6.4), (8.0, 6.4)], key=lambda t: t[0] * t[1])
PLAIN = (8.0, 2.8)
S = 32
N_TRAIN = 256
N_TOTAL = 320


def instance(i: int, base: int = 2026):
    rng = np.random.default_rng(base * 100003 + i)
    return dict(index=i, L=int(rng.integers(16, 41)), W=float(rng.uniform(2.0, 20.0)),
                seed=int(rng.integers(1, 2 ** 31 - 1)), E0=float(rng.uniform(-1.5, 1.5)),
                target=int(rng.integers(40, 241)))


# mixed-path runs: guard vectors' refinement degree as a fraction of the
# degree (None: the library default, a fiftieth); set by --guard-fraction
GUARD_FRACTION = None


def _solve(sm, dm, nev, K, mmax, emin, emax, smin, smax, anorm, ctx, precision="fp64",
           refine_degree=None):
    t = time.perf_counter()
    gd = None
    if precision == "fp32" and GUARD_FRACTION is not None:
        gd = max(1, int(math.ceil(GUARD_FRACTION * (refine_degree or K))))
    lam, _, res, st = sm.polfed(dm, nev, K, s=S, smin=smin, smax=smax, emin=emin, emax=emax,
                                sigma=0.5 * (emin + emax), mmax=mmax, tol=1e-10 * anorm, ctx=ctx,
                                return_stats=True, want_X=False, precision=precision,
                                refine_degree=refine_degree, refine_guard_degree=gd)
    dt = time.perf_counter() - t
    ok = (lam >= emin) & (lam <= emax) & (res <= 1e-10 * anorm)
    return np.sort(lam[ok]), dt, st


def gen(first: int, last: int, out: Path, indices=None):
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd import kpm
    from sparse_matrix_algorithms_amd.lattices import anderson3d
    ctx = sm.default_context()
    with open(out, "a") as f:
        for i in (indices if indices is not None else range(first, last)):
            inst = instance(i)
            A = anderson3d(inst["L"], inst["W"], inst["seed"])
            n = A.shape[0]
            dm = sm.DeviceMatrix(A, ctx)
            smin, smax = dm.gershgorin_bounds()
            anorm = max(abs(smin), abs(smax))
            mu, _, _ = kpm.kpm_moments(dm, 256, 16, 7, smin, smax, ctx=ctx)
            emin, emax = kpm.interval_for_count(mu, n, inst["target"], smin, smax, inst["E0"])
            nev = math.ceil(inst["target"] / 0.96)
            K0, m0 = kpm.plain_parameters(n, nev, emin, emax, smin, smax, S, *PLAIN)
            lam0, t0, st0 = _solve(sm, dm, nev, K0, m0, emin, emax, smin, smax, anorm, ctx)
            rec = dict(inst, n=n, nev=nev, emin=emin, emax=emax, smin=smin, smax=smax,
                       mu=[float(x) for x in mu[:64]], count_est=kpm.count(mu, n, emin, emax, smin,
                                                                           smax),
                       plain=dict(kappa=PLAIN[0], factor=PLAIN[1], degree=K0, mmax=m0,
                                  seconds=t0, valid=int(len(lam0)),
                                  block_steps=int(st0["block_steps"])),
                       ladder=[])
            best = dict(kappa=PLAIN[0], factor=PLAIN[1], degree=K0, mmax=m0, seconds=t0)
            rec["plain"]["converged"] = bool(len(lam0) >= 0.9 * rec["count_est"])
            if not rec["plain"]["converged"]:
                rec["rescue"] = []
                for kappa, factor in RESCUE:
                    K, mmax = kpm.plain_parameters(n, nev, emin, emax, smin, smax, S, kappa, factor)
                    lam, t, _ = _solve(sm, dm, nev, K, mmax, emin, emax, smin, smax, anorm, ctx)
                    ok = len(lam) >= 0.9 * rec["count_est"]
                    rec["rescue"].append(dict(kappa=kappa, factor=factor, seconds=t,
                                              valid=int(len(lam)), ok=bool(ok)))
                    if ok:
                        best = dict(kappa=kappa, factor=factor, degree=K, mmax=mmax, seconds=t,
                                    valid=int(len(lam)))
                        break
            elif len(lam0) > 0:
                for kappa, factor in LADDER:
                    K, mmax = kpm.plain_parameters(n, nev, emin, emax, smin, smax, S, kappa, factor)
                    lam, t, _ = _solve(sm, dm, nev, K, mmax, emin, emax, smin, smax, anorm, ctx)
                    same = len(lam) == len(lam0) and float(np.max(np.abs(lam - lam0))) <= 1e-10 * anorm
                    rec["ladder"].append(dict(kappa=kappa, factor=factor, seconds=t,
                                              valid=int(len(lam)), same=bool(same)))
                    if same:
                        best = dict(kappa=kappa, factor=factor, degree=K, mmax=mmax, seconds=t)
                        break
            rec["best"] = best
            f.write(json.dumps(rec) + "\n")
            f.flush()
            print(f"[gen] {i} L={inst['L']} W={inst['W']:.2f} n={n} nev={nev} plain {t0:.3f}s "
                  f"valid={len(lam0)} -> best kappa={best['kappa']} f={best['factor']} "
                  f"{best['seconds']:.3f}s ({len(rec['ladder'])} ladder runs)", flush=True)
            dm.free()


def load(path: Path):
    recs = [json.loads(l) for l in open(path) if l.strip()]
    recs.sort(key=lambda r: r["index"])
    return recs


def _xy(recs):
    from sparse_matrix_algorithms_amd.learned import encode_targets, features
    X = np.stack([features(np.array(r["mu"]), r["emin"], r["emax"], r["smin"], r["smax"], r["n"],
                           r["nev"], S) for r in recs])
    T = np.stack([encode_targets(r["best"]["kappa"], r["best"]["factor"]) for r in recs])
    return X, T


def train(path: Path, epochs: int = 3000, lr: float = 3e-3, out: Path = None):
    from sparse_matrix_algorithms_amd.learned import MLP, WEIGHTS
    recs = [r for r in load(path) if r["index"] < N_TRAIN and
            (r["plain"]["converged"] or any(x["ok"] for x in r.get("rescue", [])))]
    X, T = _xy(recs)
    from sparse_matrix_algorithms_amd.learned import decode_targets
    model = MLP(seed=5)
    hist = model.train(X, T, epochs, lr)
    # safety margins: 95th percentile of (label - prediction) on the training set,
    # so the predicted filter under-resolves at most ~5% of the training windows
    from sparse_matrix_algorithms_amd.learned import MARGIN_BUCKETS, margin_bucket
    pred = np.array([decode_targets(y) for y in model.predict(X)])
    lab = np.array([(r["best"]["kappa"], r["best"]["factor"]) for r in recs])
    blocks = np.array([p[1] * r["nev"] / S for p, r in zip(pred, recs)])
    bucket = np.array([margin_bucket(b) for b in blocks])
    margins = []
    for k in range(len(MARGIN_BUCKETS) + 1):
        sel = bucket == k
        if sel.sum() < 5:
            margins.append([0.5, 0.5])
            continue
        margins.append([float(max(0.0, np.quantile(lab[sel, 0] - pred[sel, 0], 0.95))),
                        float(max(0.0, np.quantile(lab[sel, 1] - pred[sel, 1], 0.95)))])
    print(f"[train] {len(recs)} runs, loss {hist[0]:.4f} -> {hist[-1]:.4f}; margins per "
          f"block-step bucket {MARGIN_BUCKETS}: {margins}", flush=True)
    model.save(out or WEIGHTS, dict(train_runs=len(recs), epochs=epochs, lr=lr,
                                    final_loss=float(hist[-1]), margins=margins))
    return model, hist


def evaluate(path: Path, weights: Path = None, margin_kappa=None, margin_factor=None,
             precision: str = "fp64", degree_scale: float = 1.0, refine_scale=None,
             rule: str = "learned", mmax_scale: float = 1.0):
    """Held-out instances (256..319): learned parameters (with the stochastic
    count fallback) versus the plain rule solved on the GPU in FP64.  The
    learned run uses `precision` and the learned degree times `degree_scale`
    (the mixed path's FP64 refinement is one more filtered subspace step, so
    its basis can use a lower degree).  With rule="plain" the candidate run
    takes the plain rule's parameters instead (degree times `degree_scale`).
    `mmax_scale` multiplies the candidate's basis size (rounded up).
    A run succeeds when it returns at least 0.9 of the KPM count of the
    window at relative residual ≤ 1e-10 and, where the plain run succeeded,
    the same eigenvalue set as the plain run."""
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd.learned import MLP, WEIGHTS, choose_parameters
    from sparse_matrix_algorithms_amd.lattices import anderson3d
    ctx = sm.default_context()
    model = MLP.load(weights or WEIGHTS, ctx)
    rows = []
    for r in [r for r in load(path) if r["index"] >= N_TRAIN]:
        A = anderson3d(r["L"], r["W"], r["seed"])
        dm = sm.DeviceMatrix(A, ctx)
        anorm = max(abs(r["smin"]), abs(r["smax"]))
        mu = np.array(r["mu"])
        if rule == "plain":
            K, mmax, src = r["plain"]["degree"], r["plain"]["mmax"], "plain"
        else:
            K, mmax, src = choose_parameters(model, mu, r["n"], r["nev"], r["emin"], r["emax"],
                                             r["smin"], r["smax"], S, margin_kappa,
                                             margin_factor)
        if src != "fallback":  # the scale is applied here, on the FP64 rule's degree
            K = max(1, math.ceil(degree_scale * K))
            mmax = max(1, math.ceil(mmax_scale * mmax - 1e-9))
        kr = None if refine_scale is None else max(1, math.ceil(refine_scale * K))
        lam, t, st = _solve(sm, dm, r["nev"], K, mmax, r["emin"], r["emax"], r["smin"],
                            r["smax"], anorm, ctx, precision, kr)
        K0, m0 = r["plain"]["degree"], r["plain"]["mmax"]
        lam0, t0, st0 = _solve(sm, dm, r["nev"], K0, m0, r["emin"], r["emax"], r["smin"],
                               r["smax"], anorm, ctx)
        need = 0.9 * r["count_est"]
        ok0 = len(lam0) >= need
        # every eigenvalue of a converged plain run must be found (a superset is fine)
        covered = len(lam0) == 0 or (len(lam) > 0 and float(np.max(
            np.abs(lam0[:, None] - lam[None, :]).min(axis=1))) <= 1e-10 * anorm)
        ok = len(lam) >= need and (not ok0 or covered)
        rows.append(dict(index=r["index"], source=src, ok=bool(ok), plain_ok=bool(ok0),
                         t_learned=t, t_plain=t0, degree=K, mmax=mmax, degree_plain=K0,
                         mmax_plain=m0, pairs_learned=int(len(lam)), pairs_plain=int(len(lam0))))
        print(f"[eval] {r['index']} {src} ok={ok} plain_ok={ok0} learned {t:.3f}s "
              f"(K={K}, mmax={mmax}) plain {t0:.3f}s (K={K0}, mmax={m0})", flush=True)
        dm.free()
    both = [x for x in rows if x["plain_ok"]]
    summ = dict(instances=len(rows), rule=rule, precision=precision, degree_scale=degree_scale,
                mmax_scale=mmax_scale,
                refine_scale=refine_scale,
                success_learned=sum(x["ok"] for x in rows),
                success_plain=sum(x["plain_ok"] for x in rows),
                fallback_rate=sum(x["source"] == "fallback" for x in rows) / max(len(rows), 1),
                time_learned=sum(x["t_learned"] for x in rows),
                time_plain=sum(x["t_plain"] for x in rows),
                speedup_where_plain_converges=(sum(x["t_plain"] for x in both) /
                                               max(sum(x["t_learned"] for x in both), 1e-9)),
                eigenpairs_per_s_learned=sum(x["pairs_learned"] for x in rows if x["ok"]) /
                max(sum(x["t_learned"] for x in rows), 1e-9),
                eigenpairs_per_s_plain=sum(x["pairs_plain"] for x in rows if x["plain_ok"]) /
                max(sum(x["t_plain"] for x in rows), 1e-9))
    print(json.dumps(summ), flush=True)
    return summ, rows


# constant (κ, f) rules tried against the learned one (the plain rule is (8, 2.8))
CONST_RULES = [(5.0, 2.8), (6.0, 2.8), (6.0, 3.4), (6.0, 4.2), (8.0, 3.4), (8.0, 4.2)]
# (degree scale, basis scale) candidates of the mixed path's learned parameters
MIXED_SCALES = [(0.6, 1.08), (0.7, 1.0), (0.7, 1.08), (0.8, 1.0), (0.8, 1.08), (1.0, 1.0)]


def _rule_params(rule, model, r):
    """(degree, mmax, precision, source) of one rule on record r: ("plain",),
    ("const", κ, f) and ("learned",) in FP64, or ("mixed", degree scale, basis
    scale) for the learned parameters on the mixed path."""
    from sparse_matrix_algorithms_amd import kpm
    from sparse_matrix_algorithms_amd.learned import choose_parameters
    if rule[0] == "plain":
        return r["plain"]["degree"], r["plain"]["mmax"], "fp64", "plain"
    if rule[0] == "const":
        K, mm = kpm.plain_parameters(r["n"], r["nev"], r["emin"], r["emax"], r["smin"], r["smax"],
                                     S, rule[1], rule[2])
        return K, mm, "fp64", "const"
    K, mm, src = choose_parameters(model, np.array(r["mu"]), r["n"], r["nev"], r["emin"],
                                   r["emax"], r["smin"], r["smax"], S)
    if rule[0] == "learned":
        return K, mm, "fp64", src
    if src != "fallback":
        K = max(1, math.ceil(rule[1] * K))
        mm = max(1, math.ceil(rule[2] * mm - 1e-9))
    return K, mm, "fp32", src


def compare_rules(path: Path, rules, first: int, last: int, out: Path):
    """Solve records [first, last) with every rule and with the plain rule as
    the reference set; one JSON line per record (appended to `out`), with the
    time, the filter steps, the valid pairs and success per rule.  Success is
    the held-out evaluation's: at least 0.9 of the KPM count at relative
    residual <= 1e-10, covering the plain run's set where that run converged."""
    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd.learned import MLP
    from sparse_matrix_algorithms_amd.lattices import anderson3d
    ctx = sm.default_context()
    model = MLP.load(ctx=ctx)
    done = set()
    if out.exists():
        done = {json.loads(l)["index"] for l in open(out) if l.strip()}
    with open(out, "a") as f:
        for r in [r for r in load(path) if first <= r["index"] < last and r["index"] not in done]:
            A = anderson3d(r["L"], r["W"], r["seed"])
            dm = sm.DeviceMatrix(A, ctx)
            anorm = max(abs(r["smin"]), abs(r["smax"]))
            need = 0.9 * r["count_est"]
            row = dict(index=r["index"], nev=r["nev"], count_est=r["count_est"], rules={})
            ref = None
            for rule in [("plain",)] + list(rules):
                K, mm, prec, src = _rule_params(rule, model, r)
                lam, t, st = _solve(sm, dm, r["nev"], K, mm, r["emin"], r["emax"], r["smin"],
                                    r["smax"], anorm, ctx, prec)
                if ref is None:
                    ref = lam if len(lam) >= need else np.zeros(0)
                covered = len(ref) == 0 or (len(lam) > 0 and float(np.max(
                    np.abs(ref[:, None] - lam[None, :]).min(axis=1))) <= 1e-10 * anorm)
                row["rules"]["/".join(map(str, rule))] = dict(
                    ok=bool(len(lam) >= need and covered), seconds=t, degree=K, mmax=mm,
                    source=src, pairs=int(len(lam)),
                    filter_steps=float(st["filter_steps"]),
                    refine_column_steps=float(st.get("refine_filter_steps", 0.0)))
            f.write(json.dumps(row) + "\n")
            f.flush()
            print(f"[rules] {r['index']} " + " ".join(
                f"{k}:{'ok' if v['ok'] else 'FAIL'}:{v['seconds']:.3f}s"
                for k, v in row["rules"].items()), flush=True)
            dm.free()


def summarize_rules(path: Path):
    """Per rule over the records of a compare_rules log: successes, total
    seconds, eigenpairs/s (pairs of successful runs over all seconds) and
    filter steps."""
    rows = [json.loads(l) for l in open(path) if l.strip()]
    out = {}
    for k in rows[0]["rules"]:
        rr = [x["rules"][k] for x in rows]
        t = sum(x["seconds"] for x in rr)
        out[k] = dict(records=len(rr), success=sum(x["ok"] for x in rr), seconds=t,
                      eigenpairs_per_s=sum(x["pairs"] for x in rr if x["ok"]) / max(t, 1e-9),
                      filter_steps=sum(x["filter_steps"] for x in rr))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["gen", "train", "eval", "rules", "summary"])
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--last", type=int, default=N_TOTAL)
    ap.add_argument("--data", type=Path, default=DATA / "anderson_runs.jsonl")
    ap.add_argument("--out", type=Path, default=None)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="eval: precision of the learned-parameter runs")
    ap.add_argument("--degree-scale", type=float, default=1.0,
                    help="eval: factor on the learned degree")
    ap.add_argument("--refine-scale", type=float, default=None,
                    help="eval: refinement degree as a factor of the basis degree")
    ap.add_argument("--mmax-scale", type=float, default=1.0,
                    help="eval: factor on the candidate's basis size")
    ap.add_argument("--rule", choices=["learned", "plain"], default="learned",
                    help="eval: parameters of the candidate runs")
    ap.add_argument("--set", choices=["const", "mixed", "heldout"], default="const",
                    help="rules: the constant (κ, f) rules, the mixed-path scales, or the "
                         "held-out comparison (--const: the chosen constant rules)")
    ap.add_argument("--const", action="append", default=[],
                    help="rules --set heldout: a constant rule 'kappa,factor' (repeatable)")
    ap.add_argument("--guard-fraction", type=float, default=None,
                    help="mixed-path runs: guard vectors' refinement degree as this fraction "
                         "of the degree (default: the library's, a fiftieth)")
    ap.add_argument("--redo-unconverged", type=Path, default=None,
                    help="regenerate the instances of this log whose plain run did not converge")
    a = ap.parse_args()
    global GUARD_FRACTION
    GUARD_FRACTION = a.guard_fraction
    if a.cmd == "gen":
        idx = None
        if a.redo_unconverged:
            idx = [r["index"] for r in load(a.redo_unconverged)
                   if r["plain"]["valid"] < 0.9 * r["count_est"]]
            print(f"[gen] regenerating {len(idx)} instances", flush=True)
        gen(a.first, a.last, a.out or a.data, idx)
    elif a.cmd == "train":
        train(a.data, a.epochs, a.lr, a.out)
    elif a.cmd == "rules":
        if a.set == "const":
            rules = [("const",) + r for r in CONST_RULES]
        elif a.set == "mixed":
            rules = [("mixed",) + r for r in MIXED_SCALES]
        else:
            from sparse_matrix_algorithms_amd.learned import MIXED_DEGREE_SCALE, MIXED_MMAX_SCALE
            rules = [("learned",)] + [("const",) + tuple(float(x) for x in c.split(","))
                                       for c in a.const]
            rules.append(("mixed", MIXED_DEGREE_SCALE, MIXED_MMAX_SCALE))
        compare_rules(a.data, rules, a.first, a.last, a.out)
    elif a.cmd == "summary":
        print(json.dumps(summarize_rules(a.out), indent=1))
    else:
        summ, rows = evaluate(a.data, precision=a.precision, degree_scale=a.degree_scale,
                              refine_scale=a.refine_scale, rule=a.rule, mmax_scale=a.mmax_scale)
        if a.out:
            Path(a.out).write_text(json.dumps(dict(summary=summ, rows=rows), indent=1))


if __name__ == "__main__":
    sys.exit(main())
