"""The LAPACK batch loops of bench.py's CPU baseline (oracle/lapack_loop.c
over scipy's bundled OpenBLAS) compute what the GPU kernels compute: each
is checked against the FP64 oracle on small seeded batches.  Skipped where
scipy's OpenBLAS is not installed (bench.py then falls back to the
oracle)."""
import numpy as np
import pytest

import lapack_loop as ll

pytestmark = pytest.mark.skipif(ll.lib() is None, reason="scipy's bundled OpenBLAS is not available")


def test_posv_matches_oracle(oracle):
    r = np.random.default_rng(1)
    k, n = 6, 37
    M = r.standard_normal((k, n, n))
    A = M @ np.swapaxes(M, 1, 2) + n * np.eye(n)
    A[2][4, 4] = -1.0
    B = r.standard_normal((k, 3, n))
    X, bad = ll.posv(A, B, 3)
    assert bad == 1
    A2, B2 = A.copy(), B.copy()
    oracle.posv_batched(A2, B2, 1)
    for b in (0, 1, 3, 4, 5):
        assert np.max(np.abs(X[b] - B2[b])) <= 1e-12 * np.abs(B2[b]).max()


def test_gels_ragged_matches_oracle(oracle):
    r = np.random.default_rng(2)
    ms = np.array([16, 90, 40, 9], dtype=np.int32)
    ns = np.array([8, 30, 40, 9], dtype=np.int32)
    offA = np.concatenate([[0], np.cumsum(ms.astype(np.int64) * ns)[:-1]])
    offB = np.concatenate([[0], np.cumsum(ms.astype(np.int64))[:-1]])
    A = r.standard_normal(int(np.sum(ms.astype(np.int64) * ns)))
    Bf = r.standard_normal(int(np.sum(ms)))
    X, bad = ll.gels_ragged(ms, ns, offA, offB, A, Bf, 2)
    assert bad == 0
    for s in range(len(ms)):
        m, n = int(ms[s]), int(ns[s])
        Am = A[offA[s]:offA[s] + m * n].reshape(n, m).T
        ref = oracle.gels(Am, Bf[offB[s]:offB[s] + m]).ravel()
        assert np.allclose(X[offB[s]:offB[s] + n], ref, rtol=1e-9, atol=1e-10)


def test_sgemm_matches_numpy():
    r = np.random.default_rng(3)
    A = r.uniform(-1, 1, (4, 16, 24)).astype(np.float32)   # (k, kk, m)
    B = r.uniform(-1, 1, (4, 8, 16)).astype(np.float32)    # (k, n, kk)
    C = ll.sgemm(A, B, 2)
    ref = np.matmul(B.astype(np.float64), A.astype(np.float64))   # C^T = B^T A^T per matrix
    assert np.max(np.abs(C - ref)) <= 16 * 2.0 ** -23 * 16


def test_nullspace_and_project_match_oracle(oracle):
    r = np.random.default_rng(4)
    k, m, n, rk = 3, 6, 10, 4
    A = np.einsum("kir,krj->kij", r.standard_normal((k, m, rk)), r.standard_normal((k, rk, n)))
    A_t = np.ascontiguousarray(np.swapaxes(A, 1, 2))
    P, bad = ll.nullspace(A_t, 1e-6, 2)
    assert bad == 0
    for b in range(k):
        _, Po, nul = oracle.nullspace(A[b])
        assert nul == n - rk
        assert np.max(np.abs(P[b] - Po)) < 1e-12
    P32 = np.ascontiguousarray(P.astype(np.float32))
    x = r.standard_normal((k, 1, n)).astype(np.float32)
    y = ll.project(P32, x.copy(), 3, 2)
    for b in range(k):
        assert np.allclose(y[b, 0], P32[b].T.astype(np.float64) @ x[b, 0], atol=1e-5)
