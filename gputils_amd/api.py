"""Python host API over the C ABI, on torch device memory.

Layout convention: a GPUtils tensor (m, n, k) -- k column-major m x n
matrices -- is the contiguous torch tensor of shape (k, n, m); matrix b is
``t[b].T`` (``t[b]`` is its transpose in row-major order).  Helpers
:func:`from_matrices` / :func:`to_matrices` convert from/to the natural
(k, m, n) stack of matrices.

Every call runs on torch's current CUDA (HIP) stream and raises
:class:`GputilsError` on a non-zero status.  There is no CPU path.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import GputilsError, check, lib

__all__ = [
    "GputilsError", "from_matrices", "to_matrices", "stream_handle",
    "potrf_batched", "potrs_batched", "posv_batched", "posv_mixed_batched",
    "gemm_batched", "dot", "nrm2", "asum", "amax", "amin", "scal", "axpy",
    "transpose_batched",
]


def from_matrices(x: torch.Tensor) -> torch.Tensor:
    """(k, m, n) stack of matrices -> (k, n, m) column-major tensor storage."""
    return x.transpose(-1, -2).contiguous()


def to_matrices(t: torch.Tensor) -> torch.Tensor:
    """(k, n, m) column-major tensor storage -> (k, m, n) stack of matrices."""
    return t.transpose(-1, -2)


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check_dev(t: torch.Tensor, name: str, dtypes=(torch.float32, torch.float64)):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise GputilsError(f"{name} must be a CUDA tensor")
    if t.dtype not in dtypes:
        raise GputilsError(f"{name} has unsupported dtype {t.dtype}")
    if not t.is_contiguous():
        raise GputilsError(f"{name} must be contiguous")


def _pfx(t):
    return "d" if t.dtype == torch.float64 else "s"


def _info(batch, device, info):
    if info is None:
        info = torch.empty(batch, dtype=torch.int32, device=device)
    return info


def potrf_batched(A: torch.Tensor, info: torch.Tensor | None = None) -> torch.Tensor:
    """A: (k, n, n); lower factor in place.  Returns info (k,) int32."""
    _check_dev(A, "A")
    k, n, n2 = A.shape
    assert n == n2
    info = _info(k, A.device, info)
    fn = getattr(lib(), f"gputils_{_pfx(A)}potrf_batched")
    check(fn(n, _ptr(A), n, n * n, _ptr(info), k, stream_handle()), "potrf_batched")
    return info


def potrs_batched(L: torch.Tensor, B: torch.Tensor, info: torch.Tensor | None = None) -> torch.Tensor:
    """L: (k, n, n) lower factors, B: (k, nrhs, n) solved in place."""
    _check_dev(L, "L")
    _check_dev(B, "B")
    k, n, _ = L.shape
    nrhs = B.shape[1]
    info = _info(k, L.device, info)
    fn = getattr(lib(), f"gputils_{_pfx(L)}potrs_batched")
    check(fn(n, nrhs, _ptr(L), n, n * n, _ptr(B), n, n * nrhs, _ptr(info), k, stream_handle()),
          "potrs_batched")
    return info


def posv_batched(A: torch.Tensor, B: torch.Tensor, info: torch.Tensor | None = None) -> torch.Tensor:
    """FP64 fused factor + solve: A <- L, B <- A^{-1} B."""
    _check_dev(A, "A", (torch.float64,))
    _check_dev(B, "B", (torch.float64,))
    k, n, _ = A.shape
    nrhs = B.shape[1]
    info = _info(k, A.device, info)
    check(lib().gputils_dposv_batched(n, nrhs, _ptr(A), n, n * n, _ptr(B), n, n * nrhs, _ptr(info), k,
                                      stream_handle()), "posv_batched")
    return info


def posv_mixed_batched(A: torch.Tensor, B: torch.Tensor, X: torch.Tensor | None = None,
                       tol: float = 1e-12, max_iter: int = 30, iters=None, info=None):
    """Mixed-precision solve; A (k,n,n) and B (k,nrhs,n) FP64 read only.
    Returns (X, iters, info)."""
    _check_dev(A, "A", (torch.float64,))
    _check_dev(B, "B", (torch.float64,))
    k, n, _ = A.shape
    nrhs = B.shape[1]
    if X is None:
        X = torch.empty_like(B)
    iters = _info(k, A.device, iters)
    info = _info(k, A.device, info)
    check(lib().gputils_dposv_mixed_batched(n, nrhs, _ptr(A), n, n * n, _ptr(B), n, n * nrhs, _ptr(X), n,
                                            n * nrhs, tol, max_iter, _ptr(iters), _ptr(info), k,
                                            stream_handle()), "posv_mixed_batched")
    return X, iters, info


def gemm_batched(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor | None = None, alpha=1.0, beta=0.0,
                 mode: str = "fp32") -> torch.Tensor:
    """C_i <- alpha A_i B_i + beta C_i.  A (k, kk, m), B (k, n, kk), C (k, n, m)."""
    _check_dev(A, "A")
    _check_dev(B, "B")
    batch, kk, m = A.shape
    n = B.shape[1]
    if C is None:
        C = torch.zeros((batch, n, m), dtype=A.dtype, device=A.device)
    _check_dev(C, "C")
    st = stream_handle()
    if A.dtype == torch.float64:
        rc = lib().gputils_dgemm_strided_batched(m, n, kk, alpha, _ptr(A), m, m * kk, _ptr(B), kk, kk * n, beta,
                                                 _ptr(C), m, m * n, batch, st)
    else:
        md = {"fp32": 0, "bf16x3": 1}[mode]
        rc = lib().gputils_sgemm_strided_batched(m, n, kk, alpha, _ptr(A), m, m * kk, _ptr(B), kk, kk * n, beta,
                                                 _ptr(C), m, m * n, batch, md, st)
    check(rc, "gemm_batched")
    return C


_WS = {}


def _workspace(device) -> torch.Tensor:
    key = str(device)
    if key not in _WS:
        _WS[key] = torch.empty(lib().gputils_reduce_workspace_bytes(), dtype=torch.uint8, device=device)
    return _WS[key]


def _reduce(name, x, y=None):
    _check_dev(x, "x")
    out = torch.empty(1, dtype=x.dtype, device=x.device)
    ws = _workspace(x.device)
    fn = getattr(lib(), f"gputils_{_pfx(x)}{name}")
    if y is not None:
        _check_dev(y, "y")
        rc = fn(x.numel(), _ptr(x), _ptr(y), _ptr(out), _ptr(ws), stream_handle())
    else:
        rc = fn(x.numel(), _ptr(x), _ptr(out), _ptr(ws), stream_handle())
    check(rc, name)
    return out


def dot(x, y):
    return _reduce("dot", x, y)


def nrm2(x):
    return _reduce("nrm2", x)


def asum(x):
    return _reduce("asum", x)


def amax(x):
    return _reduce("amax", x)


def amin(x):
    return _reduce("amin", x)


def scal(alpha, x):
    _check_dev(x, "x")
    check(getattr(lib(), f"gputils_{_pfx(x)}scal")(x.numel(), alpha, _ptr(x), stream_handle()), "scal")
    return x


def axpy(alpha, x, y):
    _check_dev(x, "x")
    _check_dev(y, "y")
    check(getattr(lib(), f"gputils_{_pfx(x)}axpy")(x.numel(), alpha, _ptr(x), _ptr(y), stream_handle()), "axpy")
    return y


def transpose_batched(A: torch.Tensor) -> torch.Tensor:
    """A: (k, n, m) storage of (m, n, k) -> storage of the (n, m, k) transposes."""
    _check_dev(A, "A")
    k, n, m = A.shape
    B = torch.empty((k, m, n), dtype=A.dtype, device=A.device)
    check(getattr(lib(), f"gputils_{_pfx(A)}transpose_batched")(m, n, _ptr(A), m, m * n, _ptr(B), n, m * n, k,
                                                                 stream_handle()), "transpose_batched")
    return B
