"""Build the oracle's C helper (``oracle/clenshaw.c``) with gcc + OpenMP.

Test infrastructure only.  The reference (Julia) has no C/C++ sources, so no
``oracle/_ref`` build exists for this path.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_bin" / "liboracle_clenshaw.so"


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-shared", "-fPIC",
           str(HERE / "clenshaw.c"), "-o", str(tmp)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
