"""On-disk fixtures (include/smalgo.h, sparse_matrix_algorithms_amd/fixtures.py):
the C ABI and the NumPy module write the same bytes and read each other's
files; corrupted and truncated files are rejected.  Host code only: these
calls need no GPU context."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from sparse_matrix_algorithms_amd import fixtures as F
from sparse_matrix_algorithms_amd.lattices import anderson3d, lap1d


def _lib():
    from sparse_matrix_algorithms_amd import _lib as L
    return L.load()


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64).ctypes.data_as(C.POINTER(C.c_int64))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_checksum_closed_form():
    # one word: (w xor 0) * K2;  two words: + (w1 xor K1) * K2   (mod 2^64)
    k2 = 0xBF58476D1CE4E5B9
    k1 = 0x9E3779B97F4A7C15
    w = np.array([7, 123456789], dtype="<u8").tobytes()
    assert F.checksum(w[:8]) == (7 * k2) % 2**64
    assert F.checksum(w) == (7 * k2 + (123456789 ^ k1) * k2) % 2**64


def test_numpy_round_trip(tmp_path):
    A = anderson3d(5, 16.5, 3)
    F.save_csr(tmp_path / "a.csr", A)
    B, sym = F.load_csr(tmp_path / "a.csr")
    assert sym
    assert (abs(A - B)).max() == 0
    lam = np.linspace(-1, 1, 17)
    res = np.geomspace(1e-15, 1e-10, 17)
    F.save_eig(tmp_path / "e.eig", lam, res)
    l2, r2 = F.load_eig(tmp_path / "e.eig")
    assert np.array_equal(l2, lam) and np.array_equal(r2, res)
    F.save_eig(tmp_path / "e2.eig", lam)
    l3, r3 = F.load_eig(tmp_path / "e2.eig")
    assert np.array_equal(l3, lam) and r3 is None


def test_c_and_numpy_write_identical_bytes(tmp_path):
    lib = _lib()
    A = sp.csr_matrix(lap1d(50))
    A.sort_indices()
    F.save_csr(tmp_path / "np.csr", A)
    rp, ci, v = A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data.astype(np.float64)
    assert lib.smalgo_csr_save(str(tmp_path / "c.csr").encode(), 50, _i64(rp), _i64(ci), _dp(v),
                               0, 1) == 0
    assert (tmp_path / "np.csr").read_bytes() == (tmp_path / "c.csr").read_bytes()
    # 1-based input (Julia CSC of a symmetric matrix) is stored 0-based
    assert lib.smalgo_csr_save(str(tmp_path / "c1.csr").encode(), 50, _i64(rp + 1), _i64(ci + 1),
                               _dp(v), 1, 1) == 0
    assert (tmp_path / "c1.csr").read_bytes() == (tmp_path / "c.csr").read_bytes()
    lam = np.array([0.5, 1.5, 2.25])
    F.save_eig(tmp_path / "np.eig", lam)
    assert lib.smalgo_eig_save(str(tmp_path / "c.eig").encode(), 3, _dp(lam), None) == 0
    assert (tmp_path / "np.eig").read_bytes() == (tmp_path / "c.eig").read_bytes()


def test_c_reads_numpy_files(tmp_path):
    lib = _lib()
    A = anderson3d(4, 10.0, 9)
    F.save_csr(tmp_path / "a.csr", A)
    n, nnz, sym = C.c_int64(), C.c_int64(), C.c_int()
    path = str(tmp_path / "a.csr").encode()
    assert lib.smalgo_csr_info(path, C.byref(n), C.byref(nnz), C.byref(sym)) == 0
    assert (n.value, nnz.value, sym.value) == (64, A.nnz, 1)
    rp = np.zeros(65, np.int64)
    ci = np.zeros(nnz.value, np.int64)
    v = np.zeros(nnz.value)
    assert lib.smalgo_csr_load(path, _i64(rp), ci.ctypes.data_as(C.POINTER(C.c_int64)), _dp(v)) == 0
    B = sp.csr_matrix((v, ci, rp), shape=(64, 64))
    assert (abs(A - B)).max() == 0
    res = np.array([1e-12, 2e-12])
    F.save_eig(tmp_path / "e.eig", [1.0, 2.0], res)
    cnt, hr = C.c_int64(), C.c_int()
    epath = str(tmp_path / "e.eig").encode()
    assert lib.smalgo_eig_load(epath, C.byref(cnt), C.byref(hr), None, None) == 0
    assert (cnt.value, hr.value) == (2, 1)
    lam, r2 = np.zeros(2), np.zeros(2)
    assert lib.smalgo_eig_load(epath, None, None, _dp(lam), _dp(r2)) == 0
    assert np.array_equal(lam, [1.0, 2.0]) and np.array_equal(r2, res)


def test_corruption_and_truncation_rejected(tmp_path):
    lib = _lib()
    A = anderson3d(4, 10.0, 1)
    p = tmp_path / "a.csr"
    F.save_csr(p, A)
    raw = bytearray(p.read_bytes())
    raw[64 + 8 * 70] ^= 0x01  # flip a bit in the payload
    (tmp_path / "bad.csr").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        F.load_csr(tmp_path / "bad.csr")
    rp = np.zeros(65, np.int64)
    ci = np.zeros(A.nnz, np.int64)
    v = np.zeros(A.nnz)
    rc = lib.smalgo_csr_load(str(tmp_path / "bad.csr").encode(), _i64(rp),
                             ci.ctypes.data_as(C.POINTER(C.c_int64)), _dp(v))
    assert rc != 0 and b"checksum" in lib.smalgo_io_last_error()
    (tmp_path / "short.csr").write_bytes(p.read_bytes()[:-16])
    with pytest.raises(ValueError, match="truncated"):
        F.load_csr(tmp_path / "short.csr")
    rc = lib.smalgo_csr_load(str(tmp_path / "short.csr").encode(), _i64(rp),
                             ci.ctypes.data_as(C.POINTER(C.c_int64)), _dp(v))
    assert rc != 0 and b"truncated" in lib.smalgo_io_last_error()
    with pytest.raises(ValueError, match="not an eigenvalue"):
        F.load_eig(p)
