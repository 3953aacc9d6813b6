"""gputils_dposv_mixed_batched_noalloc, called through the C ABI, against
gputils_dposv_mixed_batched (its caller-provided workspace) on the same
inputs: identical X, iters and info, bit for bit, on calls in a row whose
second is smaller than the first (the library's scratch then holds the
first call's data), a call that needs no scratch (n <= 128) after them, and
a call beyond the 16 MiB scratch, which is refused without a launch."""
import ctypes

import numpy as np
import pytest
import torch

from gputils_amd import _lib

pytestmark = pytest.mark.gpu


def _spd(k, n, seed):
    r = np.random.default_rng(seed)
    M = r.standard_normal((k, n, n))
    A = M @ np.swapaxes(M, 1, 2) + n * np.eye(n)
    A[k - 1][3, 3] = -1.0  # one non-SPD system: info = 4 on both sides
    return A


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _solve(noalloc, A, B):
    L = _lib.lib()
    k, n, _ = A.shape
    nrhs = B.shape[1]
    X = torch.full_like(B, np.nan)
    iters = torch.full((k,), 77, dtype=torch.int32, device=A.device)
    info = torch.full((k,), 77, dtype=torch.int32, device=A.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = L.gputils_dposv_mixed_batched_noalloc(n, nrhs, _ptr(A), n, n * n, _ptr(B), n, n * nrhs, _ptr(X), n,
                                                   n * nrhs, 1e-12, 30, _ptr(iters), _ptr(info), k, stream)
    else:
        wb = L.gputils_dposv_mixed_workspace_bytes(n, k)
        work = torch.empty(max(wb, 1), dtype=torch.uint8, device=A.device)
        rc = L.gputils_dposv_mixed_batched(n, nrhs, _ptr(A), n, n * n, _ptr(B), n, n * nrhs, _ptr(X), n, n * nrhs,
                                           1e-12, 30, _ptr(iters), _ptr(info), _ptr(work) if wb > 0 else None, k,
                                           stream)
    assert rc == 0, L.gputils_last_error()
    torch.cuda.synchronize()
    return X, iters, info


@pytest.mark.parametrize("k,n,nrhs", [(6, 200, 3), (4, 160, 2), (9, 96, 4)])
def test_noalloc_matches_workspace_entry_point(k, n, nrhs):
    print(f"[no_alloc] k={k} n={n} nrhs={nrhs}", flush=True)
    A = torch.from_numpy(_spd(k, n, seed=n)).cuda()
    B = torch.from_numpy(np.random.default_rng(n + 1).standard_normal((k, nrhs, n))).cuda()
    Xw, iw, fw = _solve(False, A, B)
    Xn, in_, fn = _solve(True, A, B)
    assert torch.equal(iw, in_) and torch.equal(fw, fn)
    assert torch.equal(Xw[: k - 1], Xn[: k - 1])  # the failing system's X is unspecified
    assert fw[k - 1].item() == 4 and int(fw[: k - 1].abs().sum()) == 0
    if n > 128:
        assert iw.tolist() == [-1] * k


def test_noalloc_rejects_work_beyond_its_scratch():
    L = _lib.lib()
    n, k = 1000, 3  # 24 MB of FP64 copies > 16 MiB
    rc = L.gputils_dposv_mixed_batched_noalloc(n, 1, ctypes.c_void_p(16), n, n * n, ctypes.c_void_p(16), n, n,
                                               ctypes.c_void_p(16), n, n, 1e-12, 30, ctypes.c_void_p(16),
                                               ctypes.c_void_p(16), k, None)
    assert rc == -16
    assert "16 MiB" in L.gputils_last_error().decode()
