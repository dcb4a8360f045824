"""The C-ABI library loads, exports every symbol include/smalgo.h declares,
and its host-only entry points (coefficients) match the oracle — no GPU
needed for any of this."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import polfed_ref as P

HEADER = Path(__file__).resolve().parent.parent / "include" / "smalgo.h"


def _declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(smalgo_[A-Za-z0-9_]+)\s*\(", txt))
    names.discard("smalgo_matblock_fn")
    return sorted(names)


def test_header_symbols_exported():
    from sparse_matrix_algorithms_amd import _lib
    lib = C.CDLL(str(_lib.LIB_PATH))
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_version_and_null_arguments():
    from sparse_matrix_algorithms_amd import _lib
    lib = _lib.load()
    assert lib.smalgo_version().decode().startswith("smalgo")
    assert lib.smalgo_ctx_create(0, None) == _lib.ERR_ARG
    assert lib.smalgo_jackson_kernel(3, None) == _lib.ERR_ARG
    assert lib.smalgo_result_count(None) == -1
    assert lib.smalgo_polfed(None, None, _lib.MATBLOCK_FN(), None, 10, 1, 1, None, None,
                             None) == _lib.ERR_ARG


@pytest.mark.parametrize("K", [0, 1, 7, 300])
def test_coefficients_through_abi_match_oracle(K):
    import sparse_matrix_algorithms_amd as sm
    assert np.max(np.abs(sm.jackson_kernel(K) - P.jackson_kernel(K))) <= 1e-15
    for a, b in [(-0.3, 0.2), (0.4, -0.9), (-3.0, 3.0)]:
        np.testing.assert_allclose(sm.cheb_window_coeffs(a, b, K), P.cheb_window_coeffs(a, b, K),
                                   rtol=0, atol=1e-14)
    for sh in (-0.5, 0.0, 1.2):
        np.testing.assert_allclose(sm.cheb_delta_coeffs(sh, K), P.cheb_delta_coeffs(sh, K),
                                   rtol=0, atol=1e-14)
    np.testing.assert_allclose(sm.cheb_smooth_window_coeffs(-0.2, 0.5, K),
                               P.cheb_smooth_window_coeffs(-0.2, 0.5, K), rtol=0, atol=1e-13)
    for w in ("step", "delta", "smooth"):
        np.testing.assert_allclose(sm.build_coeffs(K, -14.2, 14.3, 0.1, -0.4, 0.6, w),
                                   P.build_coeffs(K, -14.2, 14.3, 0.1, -0.4, 0.6, w),
                                   rtol=0, atol=1e-13)
    with pytest.raises(ValueError):
        sm.build_coeffs(K, -1.0, 1.0, 0.0, -0.5, 0.5, "box")


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    from sparse_matrix_algorithms_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError, match="no CPU fallback"):
        _lib.load()


def _header_struct_fields(name):
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = re.search(r"typedef struct\s+" + name + r"\s*\{([^}]*)\}\s*" + name + r"\s*;",
                     txt).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, rest = decl.split(None, 1)
        for f in rest.split(","):
            fields.append((ctype, f.strip()))
    return fields


@pytest.mark.parametrize("cname,pyname", [("smalgo_polfed_params", "PolfedParams"),
                                          ("smalgo_cheb_params", "ChebParams"),
                                          ("smalgo_slice", "SliceParams")])
def test_param_structs_mirror_the_header(cname, pyname):
    """The ctypes mirrors list the header's fields in order with the same C types."""
    from sparse_matrix_algorithms_amd import _lib
    ctypes_of = {"int": C.c_int, "double": C.c_double, "uint64_t": C.c_uint64}
    want = [(f, ctypes_of[t]) for t, f in _header_struct_fields(cname)]
    have = [(f, t) for f, t in getattr(_lib, pyname)._fields_]
    assert have == want


def test_default_params_through_abi():
    """smalgo_*_default_params fill the mirrors with the reference's defaults
    (polfed.jl:1128, chebyshev_subspace_iteration.jl:667) and the mixed path's."""
    from sparse_matrix_algorithms_amd import _lib
    lib = _lib.load()
    p = _lib.PolfedParams()
    lib.smalgo_polfed_default_params(C.byref(p))
    assert (p.s, p.has_bounds, p.sigma, p.tol, p.overextract, p.refine) == (8, 0, 0.0, 1e-10, 2, 1)
    assert (p.emin, p.emax) == (-math.inf, math.inf)
    assert (p.jackson, p.mmax, p.rtol, p.reorth_passes, p.extra_keep) == (1, 0, 1e-12, 1, 8)
    assert (p.seed, p.chunk, p.show_progress, p.precision) == (1, 64, 0, 0)
    assert (p.refine_guard, p.refine_degree, p.refine_maxiter) == (-1, -1, 3)
    assert (p.refine_tol, p.refine_guard_degree) == (-1.0, -1)
    q = _lib.ChebParams()
    lib.smalgo_cheb_default_params(C.byref(q))
    assert (q.m0, q.degree, q.maxiter, q.tol, q.res_gate) == (200, 80, 8, 1e-10, 1e-6)
    assert (q.jackson, q.debug, q.two_hit, q.seed, q.window_type) == (1, 0, 1, 1, 0)


def test_wrappers_reject_malformed_start_blocks_before_the_device():
    """Start blocks and projection bases of the wrong shape raise before any
    context or device call (the library would read n·min(m0, n) values)."""
    import scipy.sparse as sps

    import sparse_matrix_algorithms_amd as sm
    A = sps.identity(20, format="csr")
    with pytest.raises(ValueError, match="Q0 must be 20"):
        sm.chebyshev_poly_real_symm_sparse(A, -0.5, 0.5, m0=8, Q0=np.zeros((20, 5)))
    with pytest.raises(ValueError, match="Q0 must be 20"):
        sm.chebyshev_poly_real_symm_sparse(A, -0.5, 0.5, m0=8, Q0=np.zeros((19, 8)))
    with pytest.raises(ValueError, match="Q0 must be 20"):
        sm.chebyshev_poly_real_symm_sparse_mf(lambda y, x: None, 20, -0.5, 0.5, smin=-1, smax=1,
                                              m0=8, Q0=np.zeros(20))
    with pytest.raises(ValueError, match="Q1 must be 20x8"):
        sm.polfed(A, 4, 10, s=8, Q1=np.zeros((20, 4)))
    with pytest.raises(ValueError, match="as many rows"):
        sm.orth_against(np.zeros((20, 3)), np.zeros((19, 2)))

