"""Pins the FP64 host oracle (oracle/oracle.c) before it is trusted as the
parity checker: against the reference suite's golden values, against
scipy's LAPACK (OpenBLAS build shipped with scipy, a third-party library)
and against closed forms.  CPU only."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.linalg.lapack as la

import golden_values as G

rng = np.random.default_rng(12345)


def spd(n, shift=None):
    M = rng.standard_normal((n, n))
    return M @ M.T + (n if shift is None else shift) * np.eye(n)


# ------------------------------------------------------------------ GEMM --
def test_gemm_golden(oracle):
    A = G.cm(G.ADDAB_A, 2, 3, 3)
    B = G.cm(G.ADDAB_B, 3, 2, 3)
    C = [oracle.gemm(A[b], B[b]) for b in range(3)]
    got = np.concatenate([c.ravel(order="F") for c in C])
    assert np.array_equal(got, np.array(G.ADDAB_C, dtype=float))


def test_gemm_alpha_beta_against_numpy(oracle):
    A, B, C = rng.standard_normal((7, 5)), rng.standard_normal((5, 3)), rng.standard_normal((7, 3))
    got = oracle.gemm(A, B, C, alpha=-0.5, beta=2.0)
    assert np.allclose(got, -0.5 * A @ B + 2.0 * C, rtol=1e-14, atol=1e-14)


# -------------------------------------------------------------- Cholesky --
def test_potrf_golden(oracle):
    L, info = oracle.potrf(G.cm(G.CHOL_A, 3, 3)[0])
    assert info == 0
    assert abs(L[0, 0] - G.CHOL_L00) < 1e-14
    assert abs(L[2, 1] - G.CHOL_L21) < 1e-14
    assert abs(L[2, 2] - G.CHOL_L22) < 1e-14
    Lg = np.array(G.CHOL_L_ROWMAJOR).reshape(3, 3)
    assert np.allclose(np.tril(L), Lg, atol=1e-14)


def test_potrf_against_lapack_and_upper_untouched(oracle):
    for n in (1, 2, 5, 17, 33, 64):
        A = spd(n)
        A_in = A.copy()
        A_in[np.triu_indices(n, 1)] = 1234.5  # upper triangle must be ignored and preserved
        L, info = oracle.potrf(A_in)
        ref, rinfo = la.dpotrf(A, lower=1, clean=1)
        assert info == 0 and rinfo == 0
        assert np.allclose(np.tril(L), ref, rtol=1e-13, atol=1e-13)
        assert np.all(L[np.triu_indices(n, 1)] == 1234.5)


def test_potrf_not_spd_info(oracle):
    # leading minor of order 2 is indefinite
    A = np.array([[1.0, 2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    assert oracle.potrf(A)[1] == 2 == la.dpotrf(A, lower=1)[1]
    assert oracle.potrf(np.array([[-1.0]]))[1] == 1
    assert oracle.potrf(np.diag([1.0, 1.0, 0.0]))[1] == 3
    assert oracle.potrf(np.array([[np.nan]]))[1] == 1


def test_potrs_golden_and_lapack(oracle):
    L, _ = oracle.potrf(G.cm(G.CHOL_A, 3, 3)[0])
    x = oracle.potrs(L, np.array(G.CHOL_B))
    assert np.allclose(x.ravel(), G.CHOL_X, atol=1e-14)
    for n, nrhs in ((4, 1), (19, 3), (40, 4)):
        A = spd(n)
        B = rng.standard_normal((n, nrhs))
        L, _ = oracle.potrf(A)
        got = oracle.potrs(L, B)
        cref, _ = la.dpotrf(A, lower=1)
        ref, _ = la.dpotrs(cref, B, lower=1)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


# -------------------------------------------------------------------- QR --
def test_geqrf_against_lapack(oracle):
    for m, n in ((4, 3), (20, 3), (7, 7), (50, 12)):
        A = rng.standard_normal((m, n))
        qr, tau = oracle.geqrf(A)
        qr2, tau2, _, info = la.dgeqrf(A)
        assert info == 0
        assert np.allclose(qr, qr2, rtol=1e-12, atol=1e-12)
        assert np.allclose(tau, tau2, rtol=1e-12, atol=1e-12)


def test_ormqr_against_lapack_and_orthogonal(oracle):
    m, n = 9, 4
    A = rng.standard_normal((m, n))
    C = rng.standard_normal((m, 2))
    qr, tau = oracle.geqrf(A)
    got = oracle.ormqr(qr, tau, C, trans=True)
    ref, _, info = la.dormqr("L", "T", qr, tau, C, lwork=64)
    assert info == 0 and np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    Q = oracle.ormqr(qr, tau, np.eye(m), trans=False)
    assert np.allclose(Q.T @ Q, np.eye(m), atol=1e-13)
    R = np.triu(qr[:n])
    assert np.allclose(Q[:, :n] @ R, A, atol=1e-12)


def test_trsm_upper(oracle):
    R = np.triu(rng.standard_normal((6, 6))) + 6 * np.eye(6)
    B = rng.standard_normal((6, 2))
    assert np.allclose(oracle.trsm_upper(R, B), sl.solve_triangular(R, B, lower=False), rtol=1e-13)


def test_gels_golden_residual(oracle):
    A = np.array(G.QRLS_A_ROWMAJOR).reshape(4, 3)
    b = np.array(G.QRLS_B)
    x = oracle.gels(A, b).ravel()
    assert abs(np.linalg.norm(A @ x - b) - G.QRLS_RESIDUAL) < 1e-10


def test_gels_batched_square_golden(oracle):
    A = G.cm(G.LS1_A, 2, 2, 3)
    B = np.array(G.LS1_B).reshape(3, 2)
    for b in range(3):
        x = oracle.gels(A[b], B[b]).ravel()
        assert np.linalg.norm(A[b] @ x - B[b]) < 1e-10
        assert np.allclose(x, np.linalg.solve(A[b], B[b]), rtol=1e-12)


def test_gels_against_lapack(oracle):
    for m, n in ((30, 8), (200, 64), (64, 64), (17, 16)):
        A = rng.standard_normal((m, n))
        b = rng.standard_normal(m)
        x = oracle.gels(A, b).ravel()
        _, ref, info = la.dgels(A, b.reshape(-1, 1))
        assert info == 0
        assert np.allclose(x, ref[:n].ravel(), rtol=1e-9, atol=1e-11)


# ------------------------------------------------------------------- SVD --
def test_svd_golden_sigma(oracle):
    B = G.cm(G.SVD8_B, 8, 3)[0]
    S, U, VT, _ = oracle.gesvd(B)
    assert abs(S[0] - G.SVD8_S01[0]) < 1e-12 and abs(S[1] - G.SVD8_S01[1]) < 1e-12
    assert U.shape == (8, 8)


def test_svd_golden_vectors(oracle):
    A = G.cm(G.SVD3_A, 3, 2, 3)
    Ug = G.cm(G.SVD3_U, 3, 3, 3)
    for b in range(3):
        S, U, VT, _ = oracle.gesvd(A[b])
        assert np.allclose(S, G.SVD3_S[2 * b:2 * b + 2], atol=1e-12)
        if b == 0:
            assert np.allclose(VT.ravel(order="F"), G.SVD3_V4, atol=1e-12)
        r = G.SVD3_RANKS[b]
        # singular vectors of nonzero sigma match element-wise (signs included)
        assert np.allclose(U[:, :r], Ug[b][:, :r], atol=1e-12)
        # the completion spans the same subspace (basis choice is not pinned)
        Pc, Pg = U[:, r:] @ U[:, r:].T, Ug[b][:, r:] @ Ug[b][:, r:].T
        assert np.allclose(Pc, Pg, atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-13)


def test_svd_against_lapack(oracle):
    for m, n in ((10, 4), (96, 48), (30, 30)):
        A = rng.standard_normal((m, n))
        S, U, VT, sweeps = oracle.gesvd(A)
        Uref, Sref, VTref = sl.svd(A, lapack_driver="gesvd")
        assert np.allclose(S, Sref, rtol=1e-12, atol=1e-12)
        assert np.allclose(U[:, :n] * S @ VT, A, atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(m), atol=1e-12)
        assert np.allclose(VT @ VT.T, np.eye(n), atol=1e-12)
        # vectors equal LAPACK's up to sign (distinct singular values)
        sgn = np.sign(np.sum(VT * VTref, axis=1))
        assert np.allclose(VT, sgn[:, None] * VTref, atol=1e-9)
        assert sweeps < 20


def test_rank_golden(oracle):
    A = G.cm(G.RANK_A, 4, 3, 3)
    ranks = [oracle.rank(oracle.gesvd(A[b])[0], G.PRECISION_HIGH) for b in range(3)]
    assert ranks == G.RANK_EXPECTED


# ------------------------------------------------------------- nullspace --
def test_nullspace_golden_invariants(oracle):
    A = G.cm(G.NS_A, 3, 4, 5)
    for b in range(5):
        N, P, nullity = oracle.nullspace(A[b])
        assert nullity == 4 - np.linalg.matrix_rank(A[b])
        assert np.linalg.norm(A[b] @ N) < 1e-12
        NtN = N.T @ N
        assert abs(NtN[0, 0] - 1) < 1e-12
        E = np.zeros((4, 4))
        E[:nullity, :nullity] = np.eye(nullity)
        assert np.allclose(NtN, E, atol=1e-12)
        # closed form: P = I - U_r U_r^T with U_r from LAPACK
        Ur, s, _ = sl.svd(A[b].T)
        r = int(np.sum(s > 1e-6))
        assert np.allclose(P, np.eye(4) - Ur[:, :r] @ Ur[:, :r].T, atol=1e-12)


def test_nullspace_trivial_is_exactly_zero(oracle):
    A = np.array(G.NS_TRIVIAL_ROWMAJOR, dtype=float).reshape(2, 3, 3)
    for b in range(2):
        N, P, nullity = oracle.nullspace(A[b])
        assert nullity == 0 and np.all(N == 0) and np.all(P == 0)


def test_nullspace_projection(oracle):
    A = np.array(G.PROJ_A_ROWMAJOR, dtype=float).reshape(3, 7)
    x = np.array(G.PROJ_X, dtype=float)
    N, P, nullity = oracle.nullspace(A)
    proj = P @ x
    assert np.linalg.norm(A @ proj) < 1e-12
    y = N @ np.array(G.PROJ_OTHER, dtype=float)
    assert (y - proj) @ (proj - x) < 1e-10


# ------------------------------------------------------- batched loops --
def test_batched_loops_match_single(oracle):
    k, n = 6, 12
    A = np.stack([spd(n) for _ in range(k)])
    B = rng.standard_normal((k, 2, n))
    A2, B2 = A.copy(), B.copy()
    assert oracle.posv_batched(A2, B2, threads=2) == 0
    for b in range(k):
        assert np.allclose(A[b] @ B2[b].T, B[b].T, atol=1e-10)
    Af = rng.standard_normal((k, 5, 4)).astype(np.float32)
    Bf = rng.standard_normal((k, 3, 5)).astype(np.float32)
    C = oracle.gemm_batched(Af, Bf, threads=2)
    for b in range(k):
        assert np.allclose(C[b].T, Af[b].T @ Bf[b].T, atol=1e-5)


def _iota(m=10, n=6):
    """The reference suite's test matrix: 1 .. m n, column-major m x n."""
    return np.arange(1, m * n + 1, dtype=np.float64).reshape(n, m).T


def test_rot_against_blas_and_reference_expectations(oracle):
    """orc_drot is BLAS drot: scipy's drot on random vectors, and the
    values the reference suite expects of applyRight/LeftGivensRotation
    (X:580-603, X:615-641)."""
    from scipy.linalg import blas
    rng = np.random.default_rng(11)
    X = rng.standard_normal((40, 2))
    c, s = 0.28, -0.96
    got = oracle.rot_cols(X, 0, 1, c, s)
    ex, ey = blas.drot(X[:, 0].copy(), X[:, 1].copy(), c, s)
    assert np.allclose(got[:, 0], ex, rtol=0, atol=1e-15)
    assert np.allclose(got[:, 1], ey, rtol=0, atol=1e-15)
    A = _iota()
    c, ms = 0.1, np.sqrt(1 - 0.01)
    R = oracle.rot_cols(A, 1, 4, c, ms)
    i = np.arange(10)
    assert np.array_equal(R[:, [0, 2, 3, 5]], A[:, [0, 2, 3, 5]])
    assert np.allclose(R[:, 1], (11 + i) * c - (41 + i) * (-ms), rtol=0, atol=1e-12)
    assert np.allclose(R[:, 4], (11 + i) * (-ms) + (41 + i) * c, rtol=0, atol=1e-12)
    ms = -np.sqrt(1 - 0.01)
    L = oracle.rot_rows(A, 1, 9, c, ms)
    j = np.arange(6)
    assert np.array_equal(L[[0] + list(range(2, 9))], A[[0] + list(range(2, 9))])
    assert np.allclose(L[1], (2 + 10 * j) * c + (10 + 10 * j) * ms, rtol=0, atol=1e-14)
    assert np.allclose(L[9], (2 + 10 * j) * (-ms) + (10 + 10 * j) * c, rtol=0, atol=1e-14)


def test_givens_annihilate_golden_and_invariants(oracle):
    """Golden entries of the reference suite (X:1713-1729), then what the
    mathematics fixes: A(k,j) = 0, A(i,j) = hypot(A(i,j), A(k,j)), the
    other rows untouched and each column's norm over rows i, k kept."""
    A = _iota()
    G = oracle.givens_annihilate(A, 0, 1, 2)
    assert abs(G[1, 2]) <= 1e-14
    assert abs(G[0, 0] - 2.137186834969645) <= 1e-14
    assert abs(G[0, 3] - 44.552125559751822) <= 1e-14
    assert abs(G[1, 3] - (-0.328797974610715)) <= 1e-14
    rng = np.random.default_rng(12)
    B = rng.standard_normal((9, 7))
    for (i, k, j) in [(0, 5, 3), (4, 2, 0), (8, 1, 6)]:
        H = oracle.givens_annihilate(B, i, k, j)
        assert abs(H[k, j]) <= 1e-15 * np.hypot(B[i, j], B[k, j])
        assert np.isclose(H[i, j], np.hypot(B[i, j], B[k, j]), rtol=1e-15, atol=0)
        others = [r for r in range(9) if r not in (i, k)]
        assert np.array_equal(H[others], B[others])
        assert np.allclose(H[i] ** 2 + H[k] ** 2, B[i] ** 2 + B[k] ** 2, rtol=1e-14, atol=0)
    # the reference loop (X:1684-1701): every element below row 0 zeroed in turn
    C = _iota()
    for k in range(1, 10):
        for j in range(6):
            C = oracle.givens_annihilate(C, 0, k, j)
            assert abs(C[k, j]) <= 1e-10
