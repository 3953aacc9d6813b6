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
    "transpose_batched", "givens_rotate_cols", "givens_rotate_rows", "givens_annihilate",
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
    wb = lib().gputils_dposv_mixed_workspace_bytes(n, k)
    work = torch.empty(wb, dtype=torch.uint8, device=A.device) if wb > 0 else None
    check(lib().gputils_dposv_mixed_batched(n, nrhs, _ptr(A), n, n * n, _ptr(B), n, n * nrhs, _ptr(X), n,
                                            n * nrhs, tol, max_iter, _ptr(iters), _ptr(info),
                                            _ptr(work) if work is not None else None, k,
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


# ------------------------------------------------------------- Givens --
def _matrix_dims(A: torch.Tensor):
    """(m, n) of one matrix stored as (n, m) or (1, n, m) (column-major)."""
    _check_dev(A, "A")
    if A.dim() == 3:
        if A.shape[0] != 1:
            raise GputilsError("Givens rotations take one matrix (numMats = 1)")
        return A.shape[2], A.shape[1]
    if A.dim() != 2:
        raise GputilsError("A must be (n, m) or (1, n, m)")
    return A.shape[1], A.shape[0]


def givens_rotate_cols(A: torch.Tensor, i: int, j: int, c: float, minus_s: float) -> torch.Tensor:
    """DTensor::applyRightGivensRotation: columns i, j of the m x n matrix
    (storage (n, m)) become c col_i + s col_j and c col_j - s col_i, s = minus_s."""
    m, n = _matrix_dims(A)
    if not (0 <= i < n and 0 <= j < n):
        raise GputilsError("column index out of range")
    es, base = A.element_size(), A.data_ptr()
    fn = getattr(lib(), f"gputils_{_pfx(A)}rot")
    check(fn(m, ctypes.c_void_p(base + es * m * i), 1, ctypes.c_void_p(base + es * m * j), 1, c, minus_s,
             stream_handle()), "rot")
    return A


def givens_rotate_rows(A: torch.Tensor, i: int, j: int, c: float, minus_s: float) -> torch.Tensor:
    """DTensor::applyLeftGivensRotation: rows i, j rotated likewise."""
    m, n = _matrix_dims(A)
    if not (0 <= i < m and 0 <= j < m):
        raise GputilsError("row index out of range")
    es, base = A.element_size(), A.data_ptr()
    fn = getattr(lib(), f"gputils_{_pfx(A)}rot")
    check(fn(n, ctypes.c_void_p(base + es * i), m, ctypes.c_void_p(base + es * j), m, c, minus_s,
             stream_handle()), "rot")
    return A


def givens_annihilate(A: torch.Tensor, i: int, k: int, j: int) -> torch.Tensor:
    """GivensAnnihilator::annihilate(i, k, j): rows i and k rotated so that
    A(k, j) becomes zero (one launch; c and s stay on the device)."""
    m, n = _matrix_dims(A)
    fn = getattr(lib(), f"gputils_{_pfx(A)}givens_annihilate")
    check(fn(m, n, _ptr(A), m, i, k, j, stream_handle()), "givens_annihilate")
    return A


# ------------------------------------------------------------ QR / LS --
def gels_batched(A: torch.Tensor, B: torch.Tensor, info: torch.Tensor | None = None) -> torch.Tensor:
    """Least squares for A (k, n, m) storage of m x n (m >= n) and B (k, nrhs, m):
    A <- QR factors, B[:, :, :n] <- solution."""
    _check_dev(A, "A")
    _check_dev(B, "B")
    k, n, m = A.shape
    nrhs = B.shape[1]
    info = _info(k, A.device, info)
    check(getattr(lib(), f"gputils_{_pfx(A)}gels_batched")(m, n, nrhs, _ptr(A), m, m * n, _ptr(B), m, m * nrhs,
                                                          _ptr(info), k, stream_handle()), "gels_batched")
    return info


def geqrf_batched(A: torch.Tensor) -> torch.Tensor:
    """Householder QR in place (LAPACK storage); returns tau (k, n)."""
    _check_dev(A, "A")
    k, n, m = A.shape
    tau = torch.empty((k, n), dtype=A.dtype, device=A.device)
    check(getattr(lib(), f"gputils_{_pfx(A)}geqrf_batched")(m, n, _ptr(A), m, m * n, _ptr(tau), n, None, k,
                                                           stream_handle()), "geqrf_batched")
    return tau


def ormqr_batched(QR: torch.Tensor, tau: torch.Tensor, C: torch.Tensor, trans: bool = True) -> torch.Tensor:
    """C (k, ncols, m) <- Q^T C (trans) or Q C, Q from geqrf_batched."""
    _check_dev(QR, "QR")
    _check_dev(C, "C")
    k, n, m = QR.shape
    ncols = C.shape[1]
    check(getattr(lib(), f"gputils_{_pfx(QR)}ormqr_batched")(1 if trans else 0, m, ncols, n, _ptr(QR), m, m * n,
                                                            _ptr(tau), n, _ptr(C), m, m * ncols, k,
                                                            stream_handle()), "ormqr_batched")
    return C


def trsm_upper_batched(R: torch.Tensor, B: torch.Tensor, ldr: int | None = None) -> torch.Tensor:
    """B (k, nrhs, ldb>=n) <- R^{-1} B for the leading n x n upper triangle of R
    (storage (k, n, ldr))."""
    _check_dev(R, "R")
    _check_dev(B, "B")
    k, n, ld = R.shape
    nrhs, ldb = B.shape[1], B.shape[2]
    check(getattr(lib(), f"gputils_{_pfx(R)}trsm_upper_batched")(n, nrhs, _ptr(R), ld, ld * n, _ptr(B), ldb,
                                                                ldb * nrhs, k, stream_handle()), "trsm_upper")
    return B


class RaggedLeastSquares:
    """Plan for many least-squares systems of uneven size (FP32 QR + one FP64
    refinement).  System s: m[s] x n[s] column-major at A_flat[offA[s]:],
    right-hand side at B_flat[offB[s]:]; the solution overwrites B[offB[s]:][:n[s]]."""

    def __init__(self, m, n, offA, offB):
        import numpy as np
        self._m = np.ascontiguousarray(m, dtype=np.int32)
        self._n = np.ascontiguousarray(n, dtype=np.int32)
        self._oa = np.ascontiguousarray(offA, dtype=np.int64)
        self._ob = np.ascontiguousarray(offB, dtype=np.int64)
        self._plan = ctypes.c_void_p()
        check(lib().gputils_gels_ragged_plan_create(ctypes.byref(self._plan), len(self._m),
                                                    self._m.ctypes.data_as(ctypes.c_void_p),
                                                    self._n.ctypes.data_as(ctypes.c_void_p),
                                                    self._oa.ctypes.data_as(ctypes.c_void_p),
                                                    self._ob.ctypes.data_as(ctypes.c_void_p)), "plan_create")

    def __call__(self, A_flat: torch.Tensor, B_flat: torch.Tensor, refine_steps: int = 2) -> torch.Tensor:
        _check_dev(A_flat, "A", (torch.float64,))
        _check_dev(B_flat, "B", (torch.float64,))
        check(lib().gputils_dgels_ragged(self._plan, _ptr(A_flat), _ptr(B_flat), refine_steps, stream_handle()),
              "gels_ragged")
        return B_flat

    def describe(self):
        """Launch shape of every bucket: a list of dicts with keys maxe (LDS
        size class, FP32 elements), systems, method, waves, per_cu, grid."""
        nb = ctypes.c_int(0)
        check(lib().gputils_gels_ragged_plan_describe(self._plan, None, 0, ctypes.byref(nb)), "plan_describe")
        desc = (ctypes.c_int32 * (7 * max(nb.value, 1)))()
        check(lib().gputils_gels_ragged_plan_describe(self._plan, desc, nb.value, None), "plan_describe")
        keys = ("maxe", "systems", "method", "waves", "per_cu", "grid", "queue")
        return [dict(zip(keys, desc[7 * i:7 * i + 7])) for i in range(nb.value)]

    def configure(self, bucket: int, method: int, waves: int, per_cu: int, queue: int = 0):
        """Set one bucket's launch shape: method 0 (column steps in LDS,
        waves 2/4/8/16) or 1 (register panels, waves 8/16), workgroups per
        CU, static ranges (queue 0) or a work queue (queue 1)."""
        check(lib().gputils_gels_ragged_plan_configure(self._plan, bucket, method, waves, per_cu, queue),
              "plan_configure")

    @property
    def concurrent(self) -> bool:
        """Buckets run concurrently on the plan's streams (True) or in turn."""
        v = ctypes.c_int(0)
        check(lib().gputils_gels_ragged_plan_get_concurrency(self._plan, ctypes.byref(v)), "plan_concurrency")
        return bool(v.value)

    @concurrent.setter
    def concurrent(self, on: bool):
        check(lib().gputils_gels_ragged_plan_set_concurrency(self._plan, 1 if on else 0), "plan_concurrency")

    def tune(self, A_flat: torch.Tensor, B_flat: torch.Tensor, refine_steps: int = 1):
        """Autotune every bucket's launch shape on these inputs (B_flat is not
        written); returns describe()."""
        _check_dev(A_flat, "A", (torch.float64,))
        _check_dev(B_flat, "B", (torch.float64,))
        check(lib().gputils_gels_ragged_plan_tune(self._plan, _ptr(A_flat), _ptr(B_flat), refine_steps,
                                                  stream_handle()), "plan_tune")
        return self.describe()

    def __del__(self):
        try:
            if self._plan:
                lib().gputils_gels_ragged_plan_destroy(self._plan)
                self._plan = ctypes.c_void_p()
        except Exception:
            pass


# ------------------------------------------------------------------ SVD --
def gesvd_batched(A: torch.Tensor, compute_u: bool = False, sweeps: bool = False, compute_v: bool = True):
    """SVD of every tall m x n matrix of A (tensor layout (k, n, m), m >= n).
    Returns (S (k, 1, n), U (k, m, m) or None, VT (k, n, n) or None, info
    (k,)[, sweeps (k,)]).  A is not modified.  compute_v=False skips the
    right singular vectors (jobvt 'N'): the sweeps then run on R^T, giving
    the same S and column spaces of U (signs may differ).  sweeps holds the
    FP64 sweep count in bits 0..15 and the FP32 one in bits 16..31
    (sweep_counts splits them)."""
    _check_dev(A, "A")
    k, n, m = A.shape
    dev, dt = A.device, A.dtype
    S = torch.empty((k, 1, n), dtype=dt, device=dev)
    VT = torch.empty((k, n, n), dtype=dt, device=dev) if compute_v else None
    U = torch.empty((k, m, m), dtype=dt, device=dev) if compute_u else None
    info = torch.empty(k, dtype=torch.int32, device=dev)
    sw = torch.empty(k, dtype=torch.int32, device=dev)
    L = lib()
    wbytes = L.gputils_gesvd_workspace_bytes(m, n, k)
    if wbytes < 0:
        raise GputilsError(f"unsupported SVD shape m={m}, n={n}")
    work = torch.empty(max(wbytes, 1), dtype=torch.uint8, device=dev) if wbytes > 0 else None
    check(getattr(L, f"gputils_{_pfx(A)}gesvd_batched")(
        m, n, _ptr(A), m, m * n, _ptr(S), n, _ptr(U) if U is not None else None, m, m * m,
        _ptr(VT) if VT is not None else None, n, n * n,
        _ptr(work) if work is not None else None, _ptr(info), _ptr(sw), k, stream_handle()), "gesvd_batched")
    return (S, U, VT, info, sw) if sweeps else (S, U, VT, info)


def sweep_counts(sw: torch.Tensor):
    """(FP64 sweeps, FP32 sweeps) per matrix from gesvd_batched's sweeps."""
    return sw & 0xFFFF, sw >> 16


def set_fp32_sweeps(enable: bool) -> bool:
    """Whether float SVDs start their Jacobi sweeps in FP32 (the default) or
    run them all in FP64; returns the previous setting (process-wide)."""
    return bool(lib().gputils_gesvd_set_fp32_sweeps(1 if enable else 0))


def svd_rank_batched(S: torch.Tensor, eps: float = 1e-6) -> torch.Tensor:
    """rank[i] = #{sigma > eps} for S of shape (k, 1, n); uint32 counts as
    int32 tensor (k,)."""
    _check_dev(S, "S")
    k = S.shape[0]
    n = S.shape[-1] * S.shape[-2]
    rank = torch.empty(k, dtype=torch.int32, device=S.device)
    check(getattr(lib(), f"gputils_{_pfx(S)}svd_rank_batched")(n, _ptr(S), n, eps, _ptr(rank), k,
                                                               stream_handle()), "svd_rank_batched")
    return rank


def nullspace_batched(A: torch.Tensor, eps: float = 1e-6):
    """Nullspace of every m x n (m <= n) matrix of A (tensor layout (k, n, m)),
    as the Nullspace constructor builds it: SVD of A^T with U, rank with
    eps, N = U[:, rank:] padded with zero columns, P = N N^T.  Returns
    (N (k, n, n), P (k, n, n), rank (k,))."""
    _check_dev(A, "A")
    k, n, m = A.shape
    if m > n:
        raise GputilsError("nullspace expects square or fat matrices (m <= n)")
    At = transpose_batched(A)            # (k, m, n): n x m matrices
    S, U, _, _ = gesvd_batched(At, compute_u=True, compute_v=False)
    rank = svd_rank_batched(S, eps)
    N = torch.empty((k, n, n), dtype=A.dtype, device=A.device)
    P = torch.empty((k, n, n), dtype=A.dtype, device=A.device)
    check(getattr(lib(), f"gputils_{_pfx(A)}nullspace_build_batched")(
        n, _ptr(U), n, n * n, _ptr(rank), _ptr(N), n, n * n, _ptr(P), n, n * n, k, stream_handle()),
        "nullspace_build_batched")
    return N, P, rank


def project(P: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """b_i <- P_i b_i in place (Nullspace::project: addAB with B aliasing C)."""
    return gemm_batched(P, b, b)


def nullspace_project(N: torch.Tensor, rank: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """b_i <- N_i (N_i^T b_i) in place from the first n - rank_i columns of
    N_i (as nullspace_batched returns N and rank): the projection P_i b_i
    of Nullspace::project, reading N's nullity columns instead of P.
    N (k, n, n), rank (k,) uint32/int32, b (k, 1, n)."""
    _check_dev(N, "N")
    _check_dev(b, "b")
    k, n, n2 = N.shape
    if n2 != n or b.shape != (k, 1, n) or b.dtype != N.dtype or rank.numel() != k:
        raise GputilsError("nullspace_project: N (k, n, n), rank (k,), b (k, 1, n) of N's dtype")
    if not b.is_contiguous() or not N.is_contiguous():
        raise GputilsError("nullspace_project: contiguous tensors expected")
    check(getattr(lib(), f"gputils_{_pfx(N)}nullspace_project_batched")(
        n, _ptr(N), n, n * n, _ptr(rank), _ptr(b), n, k, stream_handle()), "nullspace_project_batched")
    return b
