"""Filter-kernel profiling driver (run under rocprofv3).

Runs the fused Clenshaw filter on the benchmark workload (3D Anderson,
W=16.5, block width b) for a fixed degree through the C ABI, plus a plain
copy kernel of known size (bytes read = bytes written = ``--copy-mib`` MiB)
so that FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes.

    python -m sparse_matrix_algorithms_amd.profile_filter --L 48 --b 64 --prec fp32 --degree 50
    python -m sparse_matrix_algorithms_amd.profile_filter --lattice xxz --L 20 --seed 20 --b 128
    python -m sparse_matrix_algorithms_amd.profile_filter --copy-sweep
    python -m sparse_matrix_algorithms_amd.profile_filter --trace-span run_kernel_trace.csv

``--trace-span`` reads a rocprofv3 kernel trace (CSV) instead and prints the
time per step of the most frequent filter kernel: the union of its dispatch
intervals divided by the number of steps, where a step of a split tile is two
concurrent dispatches (``--launches-per-step 2``, the default) whose
individual durations overlap.

With ``--slices N`` it runs the packed multi-slice filter instead: N slices of
widths 8, 64, 16, 32, ... (the batched-slices benchmark's widths, cycled),
all of degree ``--degree``, packed and launched as the batched driver does
(HIP graphs), and prints the library's own per-launch time for comparison
with the profiler's kernel average.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time

import numpy as np


# block widths of the batched-slices benchmark (nev 50..400 at L=64)
SLICE_WIDTHS = [8, 64, 16, 32, 8, 32, 16, 64, 8, 32, 16, 32, 16, 64, 32, 32]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=48)
    ap.add_argument("--W", type=float, default=16.5)
    ap.add_argument("--seed", type=int, default=48)
    ap.add_argument("--b", type=int, default=64)
    ap.add_argument("--prec", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--degree", type=int, default=50)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--slices", type=int, default=0)
    ap.add_argument("--lattice", choices=["anderson", "xxz"], default="anderson",
                    help="xxz: random-field XXZ chain, Sz=0 sector, length --L (h=5, seed --seed)")
    ap.add_argument("--copy-sweep", action="store_true",
                    help="only time the copy kernel over buffers of 16 MiB to 4 GiB")
    ap.add_argument("--trace-span", default=None,
                    help="rocprofv3 kernel-trace CSV: per-step time of the filter kernel")
    ap.add_argument("--launches-per-step", type=int, default=2)
    args = ap.parse_args()
    if args.trace_span:
        return trace_span(args.trace_span, args.launches_per_step)

    import sparse_matrix_algorithms_amd as sm
    from sparse_matrix_algorithms_amd.lattices import anderson3d, gaussian_block, xxz_sz0

    ctx = sm.default_context()
    if args.copy_sweep:
        sec = C.c_double()
        for mib in (16, 64, 128, 256, 512, 1024, 4096):
            nbytes = mib << 20
            reps = max(3, min(50, (8 << 30) // nbytes))
            ctx.check(ctx.lib.smalgo_copy_bench(ctx.h, nbytes, reps, C.byref(sec)))
            print(f"copy {mib} MiB (read {mib} MiB + write {mib} MiB): {sec.value * 1e6:.1f} us "
                  f"-> {2 * nbytes / sec.value / 1e9:.0f} GB/s", flush=True)
        return 0
    if args.lattice == "xxz":
        A = xxz_sz0(args.L, 5.0, args.seed)
    else:
        A = anderson3d(args.L, args.W, args.seed)
    n = A.shape[0]
    dm = sm.DeviceMatrix(A, ctx)
    smin, smax = dm.gershgorin_bounds()
    e0 = float(A.diagonal().sum() / n)
    co = sm.build_coeffs(args.degree, smin, smax, e0, e0 - 0.04, e0 + 0.04, "step")
    if args.slices:
        widths = [SLICE_WIDTHS[i % len(SLICE_WIDTHS)] for i in range(args.slices)]
        blocks = [gaussian_block(1 + i, n, w) for i, w in enumerate(widths)]
        _, st = sm.apply_poly_slices(dm, blocks, [co] * len(widths), smin, smax,
                                     precision=args.prec, return_stats=True, ctx=ctx,
                                     replays=args.reps)
        print(f"packed filter: {args.slices} slices, widths {widths}, degree {args.degree}, "
              f"n={n}, {args.prec}: panel width {st['panel_width']:.0f}, "
              f"{st['launches']:.0f} launches per run, "
              f"{st['ms_per_run'] / st['launches'] * 1e3:.2f} us per launch, "
              f"{st['bytes_per_run'] / st['ms_per_run'] / 1e6:.0f} GB/s algorithmic", flush=True)
        return 0
    Q = gaussian_block(1, n, args.b)
    for r in range(args.reps):
        t = time.perf_counter()
        sm.apply_poly_block(dm, Q, co, smin, smax, precision=args.prec, ctx=ctx)
        print(f"filter rep {r}: {time.perf_counter() - t:.4f}s (degree {args.degree}, "
              f"n={n}, b={args.b}, {args.prec})", flush=True)
    w = 4 if args.prec == "fp32" else 8
    per_launch = dm.nnz * (w + 4) + 4 * (n + 1) + 4 * n * args.b * w
    print(f"algorithmic bytes per fused step: {per_launch}", flush=True)
    sec = C.c_double()
    nbytes = args.copy_mib << 20
    ctx.check(ctx.lib.smalgo_copy_bench(ctx.h, nbytes, 20, C.byref(sec)))
    print(f"copy kernel: {nbytes} B read + {nbytes} B written per launch, {sec.value * 1e6:.1f} us "
          f"-> {2 * nbytes / sec.value / 1e9:.1f} GB/s", flush=True)


def trace_span(path, per_step):
    """Per-step time of the most frequent Clenshaw kernel in a kernel trace."""
    import collections
    import csv
    import json
    rows = list(csv.DictReader(open(path)))
    cnt = collections.Counter(r["Kernel_Name"] for r in rows if "clenshaw" in r["Kernel_Name"])
    name = cnt.most_common(1)[0][0]
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Stream_Id"])
                for r in rows if r["Kernel_Name"] == name)
    union, (cs, ce) = 0, iv[0][:2]
    for s0, e0, _ in iv[1:]:
        if s0 > ce:
            union += ce - cs
            cs, ce = s0, e0
        else:
            ce = max(ce, e0)
    union += ce - cs
    steps = len(iv) // per_step
    res = dict(kernel=name.split("(")[0], dispatches=len(iv), launches_per_step=per_step,
               steps=steps, avg_dispatch_us=sum(e - s0 for s0, e, _ in iv) / len(iv) / 1e3,
               union_ms=union / 1e6, avg_step_us=union / max(steps, 1) / 1e3,
               streams=dict(collections.Counter(x[2] for x in iv)))
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
