"""numpy front end of the FP64 host oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by bench.py's cpu_baseline leg.  The shipped library
never imports it.

All arrays follow GPUtils' tensor layout: a (m, n, k) tensor is stored
column-major per matrix with the matrix axis last in memory, which is the
numpy array ``x`` of shape (k, n, m) (C order) -- i.e. x[b].T is matrix b.
The helpers below take and return plain numpy matrices of shape (m, n)
(Fortran order internally) and batches of shape (k, m, n).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> str:
    """Compile oracle.c with gcc (make) into liboracle.so next to this file,
    also when a source is newer than the library."""
    stale = not os.path.exists(_LIB_PATH) or any(
        os.path.getmtime(os.path.join(_HERE, f)) > os.path.getmtime(_LIB_PATH)
        for f in ("oracle.c", "oracle.h", "lapack_loop.c"))
    if force or stale:
        subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_dgemm.argtypes = [_i64, _i64, _i64, ctypes.c_double, _dp, _i64, _dp, _i64,
                                ctypes.c_double, _dp, _i64]
        L.orc_dpotrf_lower.argtypes = [_i64, _dp, _i64]
        L.orc_dpotrf_lower.restype = ctypes.c_int
        L.orc_dpotrs_lower.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64]
        L.orc_dgeqrf.argtypes = [_i64, _i64, _dp, _i64, _dp]
        L.orc_dormqr.argtypes = [ctypes.c_int, _i64, _i64, _i64, _dp, _i64, _dp, _dp, _i64]
        L.orc_dtrsm_lun.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64]
        L.orc_dgels.argtypes = [_i64, _i64, _i64, _dp, _i64, _dp, _i64]
        L.orc_dgesvd.argtypes = [_i64, _i64, _dp, _i64, _dp, _dp, _i64, _dp, _i64]
        L.orc_dgesvd.restype = ctypes.c_int
        L.orc_rank.argtypes = [_i64, _dp, ctypes.c_double]
        L.orc_rank.restype = ctypes.c_int
        L.orc_nullspace.argtypes = [_i64, _i64, _dp, _i64, ctypes.c_double, _dp, _dp]
        L.orc_nullspace.restype = ctypes.c_int
        L.orc_dposv_batched.argtypes = [_i64, _i64, _dp, _dp, _i64, ctypes.c_int]
        L.orc_dposv_batched.restype = ctypes.c_int
        L.orc_dgemm_batched.argtypes = [_i64, _i64, _i64, _dp, _dp, _dp, _i64, ctypes.c_int]
        L.orc_sgemm_batched.argtypes = [_i64, _i64, _i64, _fp, _fp, _fp, _i64, ctypes.c_int]
        L.orc_dgels_ragged.argtypes = [_i64, _i32p, _i32p, _i64p, _i64p, _dp, _dp, ctypes.c_int]
        L.orc_nullspace_batched.argtypes = [_i64, _i64, _dp, ctypes.c_double, _dp, _dp, _i32p, _i64,
                                            ctypes.c_int]
        L.orc_nullspace_batched.restype = ctypes.c_int
        L.orc_drot.argtypes = [_i64, _dp, _i64, _dp, _i64, ctypes.c_double, ctypes.c_double]
        L.orc_dgivens_annihilate.argtypes = [_i64, _i64, _dp, _i64, _i64, _i64, _i64]
        _lib = L
    return _lib


def _f64(a):
    return np.array(a, dtype=np.float64, order="F", copy=True)


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def gemm(A, B, C=None, alpha=1.0, beta=0.0):
    A, B = _f64(A), _f64(B)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), order="F") if C is None else _f64(C)
    lib().orc_dgemm(m, n, k, alpha, _p(A), m, _p(B), k, beta, _p(C), m)
    return C


def potrf(A):
    """Returns (L-in-place matrix, info). Upper triangle left as given."""
    A = _f64(A)
    n = A.shape[0]
    info = lib().orc_dpotrf_lower(n, _p(A), n)
    return A, info


def potrs(L, B):
    L, B = _f64(L), _f64(B)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n, nrhs = B.shape
    lib().orc_dpotrs_lower(n, nrhs, _p(L), L.shape[0], _p(B), n)
    return B


def geqrf(A):
    A = _f64(A)
    m, n = A.shape
    tau = np.zeros(max(min(m, n), 1))
    lib().orc_dgeqrf(m, n, _p(A), m, _p(tau))
    return A, tau[: min(m, n)]


def ormqr(QR, tau, C, trans=True):
    QR, C = _f64(QR), _f64(C)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    if C.ndim == 1:
        C = C.reshape(-1, 1, order="F")
    m, ncols = C.shape
    lib().orc_dormqr(1 if trans else 0, m, ncols, len(tau), _p(QR), QR.shape[0], _p(tau), _p(C), m)
    return C


def trsm_upper(R, B):
    R, B = _f64(R), _f64(B)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    n = R.shape[1]
    lib().orc_dtrsm_lun(n, B.shape[1], _p(R), R.shape[0], _p(B), B.shape[0])
    return B


def gels(A, b):
    """Least-squares solution x (n,) of min ||A x - b|| (m >= n)."""
    A = _f64(A)
    m, n = A.shape
    B = _f64(np.asarray(b, dtype=np.float64).reshape(m, -1))
    lib().orc_dgels(m, n, B.shape[1], _p(A), m, _p(B), m)
    return B[:n].copy()


def gesvd(A, compute_u=True):
    """Returns (S, U or None, VT, sweeps) of a tall m x n matrix."""
    A = _f64(A)
    m, n = A.shape
    S = np.zeros(max(n, 1))
    U = np.zeros((m, m), order="F") if compute_u else None
    VT = np.zeros((n, n), order="F")
    sweeps = lib().orc_dgesvd(m, n, _p(A), m, _p(S),
                              _p(U) if compute_u else None, m, _p(VT), n)
    return S[:n], U, VT, sweeps


def rot_rows(A, i, k, c, s):
    """Rows i and k of A (m x n) rotated as by DTensor::applyLeftGivensRotation:
    row_i <- c row_i + s row_k, row_k <- c row_k - s row_i.  Returns a copy."""
    F = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    m = F.shape[0]
    ptr = F.ctypes.data
    lib().orc_drot(F.shape[1], ctypes.cast(ptr + 8 * i, _dp), m, ctypes.cast(ptr + 8 * k, _dp), m, c, s)
    return F


def rot_cols(A, i, j, c, s):
    """Columns i and j of A rotated as by DTensor::applyRightGivensRotation."""
    F = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    m = F.shape[0]
    ptr = F.ctypes.data
    lib().orc_drot(m, ctypes.cast(ptr + 8 * m * i, _dp), 1, ctypes.cast(ptr + 8 * m * j, _dp), 1, c, s)
    return F


def givens_annihilate(A, i, k, j):
    """GivensAnnihilator::annihilate(i, k, j) on a copy of A (m x n)."""
    F = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    lib().orc_dgivens_annihilate(F.shape[0], F.shape[1], _p(F), F.shape[0], i, k, j)
    return F


def rank(S, eps=1e-6):
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_rank(len(S), _p(S), eps)


def nullspace(A, eps=1e-6):
    """Returns (N, P, nullity) of a fat m x n matrix (m <= n)."""
    A = _f64(A)
    m, n = A.shape
    N = np.zeros((n, n), order="F")
    P = np.zeros((n, n), order="F")
    nullity = lib().orc_nullspace(m, n, _p(A), m, eps, _p(N), _p(P))
    return N, P, nullity


# ---- batched loops (tensor layout: (k, n, m) C-order == column-major per matrix)

def posv_batched(A_t, B_t, threads=None):
    """A_t: (k, n, n) float64, B_t: (k, nrhs, n) float64 (tensor layout).
    Factor + solve in place; returns number of failed factorisations."""
    assert A_t.dtype == np.float64 and A_t.flags.c_contiguous
    assert B_t.dtype == np.float64 and B_t.flags.c_contiguous
    k, n, _ = A_t.shape
    nrhs = B_t.shape[1]
    threads = threads or os.cpu_count() or 1
    return lib().orc_dposv_batched(n, nrhs, _p(A_t), _p(B_t), k, threads)


def gemm_batched(A_t, B_t, threads=None):
    """A_t: (k, kk, m), B_t: (k, n, kk) tensor layout; returns C_t (k, n, m)."""
    k, kk, m = A_t.shape
    n = B_t.shape[1]
    threads = threads or os.cpu_count() or 1
    if A_t.dtype == np.float32:
        C = np.zeros((k, n, m), dtype=np.float32)
        lib().orc_sgemm_batched(m, n, kk, _p(A_t, _fp), _p(B_t, _fp), _p(C, _fp), k, threads)
    else:
        C = np.zeros((k, n, m), dtype=np.float64)
        lib().orc_dgemm_batched(m, n, kk, _p(A_t), _p(B_t), _p(C), k, threads)
    return C


def gels_ragged(ms, ns, offA, offB, A_flat, B_flat, threads=None):
    ms = np.ascontiguousarray(ms, dtype=np.int32)
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    offB = np.ascontiguousarray(offB, dtype=np.int64)
    threads = threads or os.cpu_count() or 1
    lib().orc_dgels_ragged(len(ms), _p(ms, _i32p), _p(ns, _i32p), _p(offA, _i64p),
                           _p(offB, _i64p), _p(A_flat), _p(B_flat), threads)


def nullspace_batched(A_t, eps=1e-6, threads=None):
    """A_t: (k, n, m) float64 tensor layout of k m x n matrices (m <= n).
    Returns (N_t, P_t, nullity) in tensor layout (k, n, n)."""
    A_t = np.ascontiguousarray(A_t, dtype=np.float64)
    k, n, m = A_t.shape
    N = np.zeros((k, n, n))
    P = np.zeros((k, n, n))
    nul = np.zeros(k, dtype=np.int32)
    threads = threads or os.cpu_count() or 1
    lib().orc_nullspace_batched(m, n, _p(A_t), eps, _p(N), _p(P), _p(nul, _i32p), k, threads)
    return N, P, nul
