"""Build ``libsmalgo.so`` (HIP C++ for gfx950) in-tree with hipcc.

Sources: ``csrc/*.hip`` (device kernels) and ``csrc/*.cpp`` (host drivers and
the C ABI), all compiled by hipcc; linked against rocSOLVER/rocBLAS (used
for the small FP64 Cholesky, triangular inverse and symmetric eigensolve).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libsmalgo.so"
BUILD = PKG.parent / "_build"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("SMALGO_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 ops have given silent wrong results on the target boxes
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", str(PKG.parent / "include"), "-Wno-unused-command-line-argument"]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    lang = [] if src.suffix == ".hip" else ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *lang, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src.name}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libsmalgo.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
