"""numpy front end of lapack_loop.c: LAPACK batch loops for bench.py's CPU
baseline (OpenBLAS as bundled with scipy, one thread per LAPACK call,
OpenMP over the matrices).  Not the oracle and not the product: bench.py
times these beside the GPU kernels, and tests/test_oracle.py checks them
against the oracle.

Arrays use the GPU side's tensor layout: a batch of k column-major m x n
matrices is the C-ordered numpy array of shape (k, n, m)."""
from __future__ import annotations

import ctypes
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblapackloop.so")
_lib = None
_i64 = ctypes.c_int64


def lapack_path() -> str | None:
    """scipy's bundled OpenBLAS (LAPACK routines with a scipy_ prefix), or None."""
    try:
        import scipy
    except ImportError:
        return None
    hits = sorted(glob.glob(os.path.join(os.path.dirname(scipy.__file__) + ".libs", "libscipy_openblas*.so")))
    return hits[0] if hits else None


def library_name() -> str:
    p = lapack_path()
    return os.path.basename(p) if p else "none"


def lib():
    """The loaded harness, or None when the LAPACK library is not available."""
    global _lib
    if _lib is None:
        path = lapack_path()
        if path is None:
            return None
        import pyoracle
        pyoracle.build()
        L = ctypes.CDLL(_LIB)
        L.ll_open.argtypes = [ctypes.c_char_p]
        if L.ll_open(path.encode()) != 0:
            return None
        vp = ctypes.c_void_p
        L.ll_posv.argtypes = [_i64, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int]
        L.ll_gels_ragged.argtypes = [_i64, vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.ll_sgemm.argtypes = [_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int]
        L.ll_sgemm.restype = None
        L.ll_nullspace.argtypes = [_i64, ctypes.c_int, ctypes.c_int, vp, ctypes.c_double, vp, ctypes.c_int]
        L.ll_project.argtypes = [_i64, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int]
        L.ll_project.restype = None
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def posv(A, B, threads):
    """A (k, n, n) SPD, B (k, nrhs, n): returns (X, failures); inputs unchanged."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    X = np.empty_like(B)
    k, n = A.shape[0], A.shape[1]
    bad = lib().ll_posv(k, n, B.shape[1], _p(A), _p(B), _p(X), threads)
    return X, bad


def gels_ragged(m, n, offA, offB, A_flat, B_flat, threads):
    """Returns (X_flat, failures): the first n[s] entries of each system's
    right-hand side replaced by its least-squares solution."""
    m = np.ascontiguousarray(m, dtype=np.int32)
    n = np.ascontiguousarray(n, dtype=np.int32)
    oa = np.ascontiguousarray(offA, dtype=np.int64)
    ob = np.ascontiguousarray(offB, dtype=np.int64)
    A = np.ascontiguousarray(A_flat, dtype=np.float64)
    X = np.array(B_flat, dtype=np.float64, copy=True)
    bad = lib().ll_gels_ragged(len(m), _p(m), _p(n), _p(oa), _p(ob), _p(A), _p(X), threads)
    return X, bad


def sgemm(A, B, threads):
    """A (k, kk, m), B (k, n, kk) float32 in tensor layout: returns C (k, n, m)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    k, kk, m = A.shape
    n = B.shape[1]
    C = np.empty((k, n, m), dtype=np.float32)
    lib().ll_sgemm(k, m, n, kk, _p(A), _p(B), _p(C), threads)
    return C


def nullspace(A_t, eps, threads):
    """A_t (k, n, m): k fat m x n matrices in tensor layout.  Returns
    (P (k, n, n), failures) with P = N N^T from LAPACK dgesvd of A^T."""
    A = np.ascontiguousarray(A_t, dtype=np.float64)
    k, n, m = A.shape
    P = np.empty((k, n, n))
    bad = lib().ll_nullspace(k, m, n, _p(A), eps, _p(P), threads)
    return P, bad


def project(P, x, calls, threads):
    """calls rounds of x_b <- P_b x_b in FP32 (P (k, n, n), x (k, 1, n)); x is
    updated in place."""
    assert P.dtype == np.float32 and x.dtype == np.float32 and P.flags.c_contiguous and x.flags.c_contiguous
    k, n = P.shape[0], P.shape[1]
    lib().ll_project(k, n, _p(P), _p(x), calls, threads)
    return x
