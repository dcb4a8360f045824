"""On-disk fixtures: a CSR matrix and an eigenvalue set (with optional
residuals) in the little-endian binary format that ``include/smalgo.h``
documents (``smalgo_csr_save`` / ``smalgo_eig_save``).  Pure NumPy, so
fixtures can be written and read without the GPU library; the C ABI reads and
writes the same files.  No reference counterpart."""
from __future__ import annotations

from pathlib import Path
from typing import Optional, Tuple

import numpy as np
import scipy.sparse as sp

CSR_MAGIC = b"SMALGO-CSR-v1\0\0\0"
EIG_MAGIC = b"SMALGO-EIG-v1\0\0\0"
_K1 = np.uint64(0x9E3779B97F4A7C15)
_K2 = np.uint64(0xBF58476D1CE4E5B9)


def checksum(payload: bytes) -> int:
    """Σ_i (w_i xor i·K1)·K2 mod 2^64 over the payload's 8-byte words."""
    w = np.frombuffer(payload, dtype="<u8")
    i = np.arange(len(w), dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum((w ^ (i * _K1)) * _K2, dtype=np.uint64))


def save_csr(path, A, symmetric: Optional[bool] = None) -> None:
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("square matrices only")
    if symmetric is None:
        symmetric = (A != A.T).nnz == 0
    payload = (A.indptr.astype("<i8").tobytes() + A.indices.astype("<i8").tobytes()
               + A.data.astype("<f8").tobytes())
    head = (CSR_MAGIC + np.array([n, A.nnz], "<i8").tobytes()
            + np.array([0, 1 if symmetric else 0], "<i4").tobytes()
            + np.array([checksum(payload)], "<u8").tobytes() + bytes(16))
    Path(path).write_bytes(head + payload)


def load_csr(path) -> Tuple[sp.csr_matrix, bool]:
    raw = Path(path).read_bytes()
    if raw[:16] != CSR_MAGIC:
        raise ValueError("not a CSR fixture")
    n, nnz = np.frombuffer(raw, "<i8", 2, 16)
    base, flags = np.frombuffer(raw, "<i4", 2, 32)
    (chk,) = np.frombuffer(raw, "<u8", 1, 40)
    if base != 0 or n < 0 or nnz < 0:
        raise ValueError("bad CSR header")
    payload = raw[64:]
    if len(payload) != 8 * (n + 1 + 2 * nnz):
        raise ValueError("file is truncated")
    if checksum(payload) != int(chk):
        raise ValueError("checksum mismatch")
    rp = np.frombuffer(payload, "<i8", n + 1, 0)
    ci = np.frombuffer(payload, "<i8", nnz, 8 * (n + 1))
    v = np.frombuffer(payload, "<f8", nnz, 8 * (n + 1 + nnz))
    return sp.csr_matrix((v.copy(), ci.copy(), rp.copy()), shape=(n, n)), bool(flags & 1)


def save_eig(path, lam, res=None) -> None:
    lam = np.ascontiguousarray(lam, dtype="<f8")
    payload = lam.tobytes()
    if res is not None:
        res = np.ascontiguousarray(res, dtype="<f8")
        if res.shape != lam.shape:
            raise ValueError("lam and res must have the same length")
        payload += res.tobytes()
    head = (EIG_MAGIC + np.array([len(lam), 0 if res is None else 1], "<i8").tobytes()
            + np.array([checksum(payload)], "<u8").tobytes() + bytes(24))
    Path(path).write_bytes(head + payload)


def load_eig(path):
    """Returns ``(lam, res)``; ``res`` is None when the file has none."""
    raw = Path(path).read_bytes()
    if raw[:16] != EIG_MAGIC:
        raise ValueError("not an eigenvalue fixture")
    count, has_res = np.frombuffer(raw, "<i8", 2, 16)
    (chk,) = np.frombuffer(raw, "<u8", 1, 32)
    payload = raw[64:]
    if has_res not in (0, 1) or len(payload) != 8 * count * (1 + has_res):
        raise ValueError("bad eigenvalue fixture")
    if checksum(payload) != int(chk):
        raise ValueError("checksum mismatch")
    lam = np.frombuffer(payload, "<f8", count, 0).copy()
    res = np.frombuffer(payload, "<f8", count, 8 * count).copy() if has_res else None
    return lam, res
