"""Float printing of the progress line (polfed.jl:991-1000 prints Float64 values
with println, i.e. Julia's print(::Float64)).

Pinned three ways: the float tokens of the reference's own published progress
lines (polfed.jl:1288-1407), strings derived by hand from Julia's printing rule
(shortest round-trip digits; fixed notation for 1e-4 <= |x| < 1e6, otherwise
d.ddde<exp> with an unpadded exponent and '.0' after a single digit), and the
C ABI formatter (C++ std::to_chars digits) against the oracle's (Python repr
digits) on many random values.  No GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import polfed_ref as P

# float tokens of the published lines, polfed.jl:1288, 1299, 1310, 1321, 1334,
# 1347, 1369, 1387, 1397, 1407 (the line at 1358 carries hand-typed text and a
# padded "0.01888119337120290"; its value is checked below as a derived case)
PUBLISHED = [
    "m=79  inside=1000  r90=1.263385139348521e-14  r95=1.4465502435677196e-14  maxres=3.1117078311317e-14",
    "m=47  inside=999  r90=2.3431295846617787e-14  r95=2.8476247888023835e-14  maxres=2.175237376209596e-12",
    "m=47  inside=998  r90=3.370908181692764e-14  r95=4.0029332417106185e-14  maxres=7.502190770434304e-14",
    "m=47  inside=998  r90=4.851085163074631e-14  r95=5.668566328697754e-14  maxres=9.447243092189946e-14",
    "m=75  inside=1599  r90=5.406490405565137e-14  r95=6.560545613539887e-14  maxres=2.12138039476538e-11",
    "m=79  inside=998  r90=6.294294088644106e-14  r95=7.424619128109143e-14  maxres=1.1966103393555824e-13",
    "m=79  inside=1000  r90=7.199055102796388e-14  r95=8.595247481961762e-14  maxres=1.2347500904341446e-13",
    "m=32  inside=1000  r90=8.773094039016345e-14  r95=1.0435955357749034e-13  maxres=2.070560874529873e-13",
    "m=47  inside=999  r90=6.2752806221874154e-15  r95=6.709334573082644e-15  maxres=2.576533337148389e-11",
    "m=94  inside=999  r90=5.991394217652792e-15  r95=6.679527965337017e-15  maxres=2.4792930459694443e-11",
]

# (value, Julia's print) derived by hand from the rule above
DERIVED = [
    (1e-5, "1.0e-5"), (1.0927391319480254e-9, "1.0927391319480254e-9"), (1e-10, "1.0e-10"),
    (9.999e-5, "9.999e-5"), (0.0001, "0.0001"), (0.00012345, "0.00012345"),
    (0.01888119337120290, "0.0188811933712029"), (0.1, "0.1"), (1.0, "1.0"), (12.0, "12.0"),
    (123456.7, "123456.7"), (100000.0, "100000.0"), (999999.0, "999999.0"), (1e6, "1.0e6"),
    (1234567.0, "1.234567e6"), (1.5e300, "1.5e300"), (5e-324, "5.0e-324"),
    (2.2250738585072014e-308, "2.2250738585072014e-308"), (-2.5e-10, "-2.5e-10"),
    (-0.5, "-0.5"), (0.0, "0.0"), (float("inf"), "Inf"), (float("nan"), "NaN"),
]


def _abi(v):
    from sparse_matrix_algorithms_amd import _lib
    lib = _lib.load()
    buf = C.create_string_buffer(64)
    assert lib.smalgo_julia_float_string(float(v), buf, 64) == 0
    return buf.value.decode()


def _published_tokens():
    for line in PUBLISHED:
        for tok in re.findall(r"=([0-9][^ ]*e-[0-9]+)", line):
            yield tok


@pytest.mark.parametrize("fmt", ["oracle", "abi"])
def test_published_progress_line_tokens(fmt):
    f = P.julia_float if fmt == "oracle" else _abi
    toks = list(_published_tokens())
    assert len(toks) == 30
    for t in toks:
        assert f(float(t)) == t


@pytest.mark.parametrize("fmt", ["oracle", "abi"])
def test_hand_derived_julia_strings(fmt):
    f = P.julia_float if fmt == "oracle" else _abi
    for v, want in DERIVED:
        assert f(v) == want, (v, f(v), want)


def test_abi_formatter_matches_oracle_on_random_values():
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.standard_normal(400) * 10.0 ** rng.integers(-20, 20, 400),
                           10.0 ** np.arange(-12, 12), 10.0 ** np.arange(-12, 12) * 1.5,
                           rng.integers(1, 10 ** 7, 50).astype(float)])
    for v in vals:
        s = P.julia_float(v)
        assert _abi(v) == s
        assert float(s) == v          # round trip


def test_abi_formatter_buffer_too_small():
    from sparse_matrix_algorithms_amd import _lib
    lib = _lib.load()
    buf = C.create_string_buffer(4)
    assert lib.smalgo_julia_float_string(1.0927391319480254e-9, buf, 4) == _lib.ERR_ARG
    assert lib.smalgo_julia_float_string(1.0, None, 4) == _lib.ERR_ARG


def test_progress_line_uses_julia_floats():
    line = P.progress_line(3, np.arange(20.0), np.array([1e-5] * 19 + [2e-9]))
    assert line == "m=3  inside=20  r90=1.0e-5  r95=1.0e-5  maxres=1.0e-5"
    line = P.progress_line(7, np.arange(10.0), np.array([float(f"{k}e-9") for k in range(1, 10)]
                                                        + [1e-8]))
    # round(Int, 9.0) = 9, round(Int, 9.5) = 10 (ties to even), 1-based
    assert line == "m=7  inside=10  r90=9.0e-9  r95=1.0e-8  maxres=1.0e-8"
