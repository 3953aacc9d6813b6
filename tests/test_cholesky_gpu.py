"""GPU parity of the batched Cholesky kernels against the FP64 oracle and
the reference suite's golden values (test/testTensor.cu X:1218-1413).

Tolerances: FP64 kernels agree with the oracle to 1e-12 (relative to the
matrix scale); FP32 kernels to 1e-4 (the reference suite's PRECISION_LOW);
the mixed solver must reach the requested backward error 1e-12."""
import numpy as np
import pytest
import torch

import golden_values as G
from gputils_amd import api

pytestmark = pytest.mark.gpu


def spd_batch(k, n, seed, shift=None):
    r = np.random.default_rng(seed)
    M = r.standard_normal((k, n, n))
    return M @ np.swapaxes(M, 1, 2) + (n if shift is None else shift) * np.eye(n)


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def backward_error(A, X, B):
    """max over matrices/columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf).
    A (k,n,n) matrices, X and B in tensor layout (k, nrhs, n)."""
    R = B - np.einsum("kij,krj->kri", A, X)
    an = np.abs(A).sum(axis=2).max(axis=1)[:, None]
    return float(np.max(np.abs(R).max(axis=2) / (an * np.abs(X).max(axis=2) + np.abs(B).max(axis=2))))


# ------------------------------------------------------------- goldens --
@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_batch_factorisation_golden(dtype, eps):
    A = np.stack([G.cm(G.CHOL_A, 3, 3)[0]] * 2)
    t = dev(A, dtype)
    info = api.potrf_batched(t)
    torch.cuda.synchronize()
    L = api.to_matrices(t).cpu().double().numpy()
    for b in range(2):
        assert abs(L[b, 0, 0] - G.CHOL_L00) < eps
        assert abs(L[b, 2, 1] - G.CHOL_L21) < eps
        assert abs(L[b, 2, 2] - G.CHOL_L22) < eps
    assert info.cpu().tolist() == [0, 0]


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_batch_factor_solve_golden(dtype, eps):
    A = np.stack([G.cm(G.CHOL_A, 3, 3)[0]] * 2)
    L = dev(A, dtype)
    api.potrf_batched(L)
    b = dev(np.tile(np.array(G.CHOL_B), (2, 1, 1)), dtype)
    api.potrs_batched(L, b)
    x = b.cpu().double().numpy()
    for k in range(2):
        assert np.allclose(x[k, 0], G.CHOL_X, atol=eps)
        assert np.linalg.norm(A[k] @ x[k, 0] - np.array(G.CHOL_B)) < eps


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_batch_solve_from_given_factor_golden(dtype, eps):
    Lg = np.array(G.CHOL_L_ROWMAJOR).reshape(3, 3)
    L = dev(api.from_matrices(torch.from_numpy(np.stack([Lg, Lg]))).numpy(), dtype)
    b = dev(np.tile(np.array(G.CHOL_B), (2, 1, 1)), dtype)
    api.potrs_batched(L, b)
    x = b.cpu().double().numpy()
    for k in range(2):
        assert np.allclose(x[k, 0], G.CHOL_X, atol=eps)


# ------------------------------------------------------- oracle parity --
@pytest.mark.parametrize("n", [1, 3, 15, 16, 17, 31, 32, 33, 64, 77, 100, 128, 129, 150])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_potrf_matches_oracle(oracle, n, dtype):
    k = 5
    A = spd_batch(k, n, seed=n)
    A_in = A.copy()
    iu = np.triu_indices(n, 1)
    for b in range(k):
        A_in[b][iu] = 777.0  # strictly upper part: must be ignored and left untouched
    t = dev(api.from_matrices(torch.from_numpy(A_in)).numpy(), dtype)
    info = api.potrf_batched(t)
    torch.cuda.synchronize()
    got = api.to_matrices(t).cpu().double().numpy()
    assert info.cpu().tolist() == [0] * k
    tol = 1e-12 if dtype == torch.float64 else 2e-4
    for b in range(k):
        L, oinfo = oracle.potrf(A[b])
        assert oinfo == 0
        scale = np.abs(np.tril(L)).max()
        assert np.max(np.abs(np.tril(got[b]) - np.tril(L))) <= tol * scale, (n, b)
        assert np.all(got[b][iu] == 777.0)


@pytest.mark.parametrize("n", [5, 16, 17, 32])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_small_potrf_reports_failing_minor(oracle, n, dtype):
    """n <= 32 (the register-resident kernel, two or four matrices per
    wavefront): indefinite, zero and NaN pivots give the oracle's info, and
    their neighbours in the same wavefront factorise normally.  The batch
    (7) is not a multiple of the matrices per wavefront."""
    k = 7
    A = spd_batch(k, n, seed=70 + n)
    A[1][n // 2, n // 2] = -3.0        # fails at row n/2 + 1
    A[2][:, :] = 0.0                   # fails at row 1
    A[4][n - 1, n - 1] = -1e6          # fails at the last row
    A[6][n // 3, n // 3] = np.nan      # fails at row n/3 + 1
    t = dev(A, dtype)
    info = api.potrf_batched(t).cpu().tolist()
    expected = [oracle.potrf(A[b])[1] for b in range(k)]
    assert expected == [0, n // 2 + 1, 1, 0, n, 0, n // 3 + 1]
    assert info == expected
    L = api.to_matrices(t).cpu().double().numpy()
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for b in (0, 3, 5):
        Lo, _ = oracle.potrf(A[b])
        assert np.max(np.abs(np.tril(L[b]) - np.tril(Lo))) <= tol * np.abs(Lo).max()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_potrf_reports_failing_minor(oracle, dtype):
    n, k = 40, 6
    A = spd_batch(k, n, seed=7)
    A[1][5, 5] = -3.0            # fails at row 6
    A[3][:, :] = 0.0             # fails at row 1
    A[4][39, 39] = -1e6          # fails at the last row
    A[5][20, 20] = np.nan        # fails at row 21
    t = dev(A, dtype)
    info = api.potrf_batched(t).cpu().tolist()
    expected = [oracle.potrf(A[b])[1] for b in range(k)]
    assert expected == [0, 6, 0, 1, 40, 21]
    assert info == expected


@pytest.mark.parametrize("n,nrhs", [(1, 1), (3, 1), (16, 4), (32, 1), (45, 3), (64, 5), (128, 4), (128, 9), (140, 2)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_potrs_matches_oracle(oracle, n, nrhs, dtype):
    k = 4
    A = spd_batch(k, n, seed=100 + n)
    B = np.random.default_rng(n).standard_normal((k, nrhs, n))
    L = dev(A, dtype)
    api.potrf_batched(L)
    X = dev(B, dtype)
    api.potrs_batched(L, X)
    got = X.cpu().double().numpy()
    tol = 1e-11 if dtype == torch.float64 else 1e-4
    for b in range(k):
        Lo, _ = oracle.potrf(A[b])
        ref = oracle.potrs(Lo, B[b].T)
        assert np.max(np.abs(got[b].T - ref)) <= tol * max(1.0, np.abs(ref).max()), (n, nrhs, b)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (3, 1), (16, 2), (17, 3), (31, 1), (32, 1), (32, 5), (33, 2), (48, 3),
                                    (50, 4), (128, 4), (128, 6), (129, 2), (180, 3)])
def test_fused_posv_matches_oracle(oracle, n, nrhs):
    k = 7   # odd: the n <= 32 kernel pairs matrices per wavefront
    A = spd_batch(k, n, seed=200 + n)
    B = np.random.default_rng(n).standard_normal((k, nrhs, n))
    tA, tB = dev(A), dev(B)
    info = api.posv_batched(tA, tB)
    L = api.to_matrices(tA).cpu().numpy()
    X = tB.cpu().numpy()
    assert info.cpu().tolist() == [0] * k
    for b in range(k):
        Lo, _ = oracle.potrf(A[b])
        assert np.max(np.abs(np.tril(L[b]) - np.tril(Lo))) <= 1e-12 * np.abs(Lo).max()
        ref = oracle.potrs(Lo, B[b].T)
        assert np.max(np.abs(X[b].T - ref)) <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("n", [40, 128, 140])
def test_fp64_posv_tiny_and_huge_pivots(oracle, n):
    """The blocked FP64 kernels form each pivot's square root and its
    reciprocal from the hardware 1/sqrt estimate; pivots below 2^-767 take
    a pre-scaled path.  A scaled by 2^-800 (every pivot below that bound)
    and by 2^+700 gives, up to FP64 rounding, the factor of A scaled by
    2^-400 / 2^+350 and the same solution (exact powers of two, no
    underflow or overflow on the way)."""
    k = 3
    A = spd_batch(k, n, seed=300 + n)
    B = np.random.default_rng(n + 1).standard_normal((k, 2, n))
    Lo = [oracle.potrf(A[b])[0] for b in range(k)]
    Xo = [oracle.potrs(Lo[b], B[b].T) for b in range(k)]
    for e in (-800, 700):
        sc = 2.0 ** e
        tA, tB = dev(A * sc), dev(B * sc)
        info = api.posv_batched(tA, tB)
        assert info.cpu().tolist() == [0] * k
        L = api.to_matrices(tA).cpu().numpy()
        X = tB.cpu().numpy()
        for b in range(k):
            Ls = np.tril(L[b]) * 2.0 ** (-e // 2)
            assert np.max(np.abs(Ls - np.tril(Lo[b]))) <= 1e-13 * np.abs(Lo[b]).max(), (e, b)
            assert np.max(np.abs(X[b].T - Xo[b])) <= 1e-11 * np.abs(Xo[b]).max(), (e, b)


@pytest.mark.parametrize("n", [40, 140])
def test_fused_posv_failure_reports_info(oracle, n):
    """A non-SPD system reports the oracle's info (blocked kernel for
    n <= 128, global-memory path beyond); beyond 128 its B is left as it
    was (LAPACK dposv), and its neighbours are solved."""
    k = 3
    A = spd_batch(k, n, seed=600 + n)
    A[1][9, 9] = -2.0
    B = np.random.default_rng(n).standard_normal((k, 2, n))
    tA, tB = dev(A), dev(B)
    info = api.posv_batched(tA, tB).cpu().tolist()
    assert info == [oracle.potrf(A[b])[1] for b in range(k)] == [0, 10, 0]
    X = tB.cpu().numpy()
    for b in (0, 2):
        Lo, _ = oracle.potrf(A[b])
        ref = oracle.potrs(Lo, B[b].T)
        assert np.max(np.abs(X[b].T - ref)) <= 1e-11 * np.abs(ref).max()
    if n > 128:
        assert np.array_equal(X[1], B[1])


# ---------------------------------------------------- mixed precision --
@pytest.mark.parametrize("n,nrhs", [(1, 1), (8, 1), (16, 2), (17, 4), (33, 4), (48, 5), (65, 1), (80, 4), (100, 3),
                                    (112, 2), (127, 4), (128, 4), (128, 7)])
def test_mixed_reaches_fp64_accuracy(oracle, n, nrhs):
    k = 16
    A = spd_batch(k, n, seed=300 + n, shift=0.5)
    B = np.random.default_rng(n + 1).standard_normal((k, nrhs, n))
    X, iters, info = api.posv_mixed_batched(dev(A), dev(B), tol=1e-12)
    X = X.cpu().numpy()
    assert info.cpu().tolist() == [0] * k
    it = iters.cpu().numpy()
    assert np.all(it >= 0) and np.all(it <= 30)
    assert backward_error(A, X, B) <= 1e-12
    for b in range(k):
        Lo, _ = oracle.potrf(A[b])
        ref = oracle.potrs(Lo, B[b].T)
        cond = np.linalg.cond(A[b])
        assert np.max(np.abs(X[b].T - ref)) <= 1e-12 * cond * np.abs(ref).max()


@pytest.mark.parametrize("n", [24, 48])
def test_mixed_falls_back_to_fp64(oracle, n):
    """cond ~ 1e12 (FP32 factor unusable) and a non-SPD matrix: the FP64
    launch takes over; iters = -1 marks them."""
    k = 4
    r = np.random.default_rng(5)
    Q, _ = np.linalg.qr(r.standard_normal((n, n)))
    A = spd_batch(k, n, seed=9)
    A[1] = (Q * np.logspace(0, -12, n)) @ Q.T
    A[1] = 0.5 * (A[1] + A[1].T)
    A[2] = -A[2]
    B = r.standard_normal((k, 1, n))
    X, iters, info = api.posv_mixed_batched(dev(A), dev(B), tol=1e-12, max_iter=10)
    it, inf = iters.cpu().tolist(), info.cpu().tolist()
    X = X.cpu().numpy()
    assert it[0] >= 0 and it[3] >= 0
    assert it[1] == -1 and it[2] == -1
    assert inf[0] == 0 and inf[2] == 1 and inf[3] == 0
    Lo, oinfo = oracle.potrf(A[1])
    assert inf[1] == oinfo
    if oinfo == 0:
        ref = oracle.potrs(Lo, B[1].T)
        assert np.allclose(X[1].T, ref, rtol=1e-2, atol=1e-2 * np.abs(ref).max())
    for b in (0, 3):
        assert backward_error(A[b:b + 1], X[b:b + 1], B[b:b + 1]) <= 1e-12


@pytest.mark.parametrize("n", [129, 160])
def test_mixed_beyond_the_fused_kernel_solves_in_fp64(oracle, n):
    """n > 128: every system is solved by the FP64 kernels on a copy of A
    (iters = -1); a non-SPD matrix reports the oracle's info; A and B stay
    unchanged."""
    k, nrhs = 3, 2
    A = spd_batch(k, n, seed=400 + n)
    A[2][7, 7] = -5.0
    B = np.random.default_rng(n).standard_normal((k, nrhs, n))
    tA, tB = dev(A), dev(B)
    X, iters, info = api.posv_mixed_batched(tA, tB, tol=1e-12)
    assert iters.cpu().tolist() == [-1] * k
    assert info.cpu().tolist() == [0, 0, oracle.potrf(A[2])[1]] == [0, 0, 8]
    X = X.cpu().numpy()
    for b in range(2):
        Lo, _ = oracle.potrf(A[b])
        ref = oracle.potrs(Lo, B[b].T)
        assert np.max(np.abs(X[b].T - ref)) <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(tA.cpu().numpy(), A) and np.array_equal(tB.cpu().numpy(), B)


def test_mixed_non_finite_iterates_fall_back(oracle):
    """The stopping test never accepts a non-finite iterate: b beyond the
    FP32 range (its FP32 solve overflows) and a NaN right-hand side go to
    the FP64 fallback (iters = -1); the overflow case then matches the
    oracle.  A matrix scaled to 1e-35 (FP32 normal range, tiny) either
    converges or falls back, and its answer matches the oracle."""
    n, k = 48, 4
    A = spd_batch(k, n, seed=33)
    B = np.random.default_rng(4).standard_normal((k, 1, n))
    B[0] *= 1e39                      # > FLT_MAX
    B[1][0, 5] = np.nan
    A[2] *= 1e-35
    X, iters, info = api.posv_mixed_batched(dev(A), dev(B), tol=1e-12, max_iter=10)
    it, inf = iters.cpu().tolist(), info.cpu().tolist()
    X = X.cpu().numpy()
    assert it[0] == -1 and it[1] == -1 and it[3] >= 0
    assert inf == [0, 0, 0, 0]
    for b in (0, 2, 3):
        Lo, _ = oracle.potrf(A[b])
        ref = oracle.potrs(Lo, B[b].T)
        assert np.all(np.isfinite(X[b]))
        assert np.max(np.abs(X[b].T - ref)) <= 1e-10 * np.abs(ref).max(), b
    assert np.isnan(X[1]).any()


def test_mixed_inputs_are_read_only():
    n, k = 64, 3
    A = spd_batch(k, n, seed=11)
    B = np.random.default_rng(2).standard_normal((k, 2, n))
    tA, tB = dev(A), dev(B)
    api.posv_mixed_batched(tA, tB)
    assert np.array_equal(tA.cpu().numpy(), A) and np.array_equal(tB.cpu().numpy(), B)


# --------------------------------------------------- benchmark configs --
def test_config_cholesky_small_fp64(oracle):
    """k=1024 SPD 32x32 (M M^T + 32 I), b 32x1: factor + solve in FP64,
    relative residual <= 1e-13, and agreement with the oracle."""
    from gputils_amd import workloads
    A, b = workloads.cholesky_small()
    tA, tb = dev(A), dev(b)
    info = api.potrf_batched(tA)
    api.potrs_batched(tA, tb)
    x = tb.cpu().numpy()
    assert int(info.abs().sum()) == 0
    res = np.linalg.norm(b - np.einsum("kij,krj->kri", A, x), axis=2).ravel()
    rel = res / np.linalg.norm(b, axis=2).ravel()
    assert rel.max() <= 1e-13
    A2, b2 = A[:64].copy(), b[:64].copy()
    assert oracle.posv_batched(A2, b2) == 0
    assert np.max(np.abs(x[:64] - b2)) <= 1e-13 * np.abs(b2).max()


def test_config_cholesky_mixed_full_size():
    """k=8192 SPD 128x128 (lambda in [1,1e4]), 4 RHS: every system reaches
    backward error <= 1e-12 with the mixed solver; the plain FP64 solve of the
    same systems agrees to cond * 1e-12."""
    from gputils_amd import workloads
    A, B = workloads.cholesky_mixed()
    tA, tB = dev(A), dev(B)
    X, iters, info = api.posv_mixed_batched(tA, tB, tol=1e-12)
    Xm = X.cpu().numpy()
    it = iters.cpu().numpy()
    assert int(info.abs().sum()) == 0
    assert np.all(it >= 0), "FP64 fallback was needed"
    assert backward_error(A, Xm, B) <= 1e-12
    tB2 = tB.clone()
    api.posv_batched(tA, tB2)   # overwrites tA with L
    Xp = tB2.cpu().numpy()
    assert backward_error(A, Xp, B) <= 1e-12
    assert np.max(np.abs(Xm - Xp)) <= 1e4 * 1e-12 * np.abs(Xp).max()


def test_config_cholesky_mixed_sample_against_oracle(oracle):
    """The configuration's own systems (lambda in [1, 1e4], cond 1e4), a
    seeded sample of 96 of the 8192: the mixed solution equals the FP64
    oracle's (dpotrf + dpotrs restated) to cond * 1e-12."""
    from gputils_amd import workloads
    A, B = workloads.cholesky_mixed()
    X, iters, info = api.posv_mixed_batched(dev(A), dev(B), tol=1e-12)
    Xm = X.cpu().numpy()
    pick = np.sort(np.random.default_rng(8192).choice(A.shape[0], 96, replace=False))
    As, Bs = A[pick].copy(), B[pick].copy()
    assert oracle.posv_batched(As, Bs) == 0
    err = np.abs(Xm[pick] - Bs).max(axis=(1, 2)) / np.abs(Bs).max(axis=(1, 2))
    assert err.max() <= 1e4 * 1e-12, err.max()
    assert backward_error(A[pick], Bs, B[pick]) <= 1e-13


@pytest.mark.parametrize("n", [5, 32])
def test_small_fused_posv_failure_and_upper_untouched(oracle, n):
    """n <= 32 fused path: a non-SPD matrix reports the LAPACK info of the
    first non-positive minor, its neighbour in the same wavefront is
    unaffected, and the strictly upper triangle of A is never written."""
    k = 4
    A = spd_batch(k, n, seed=77)
    A[1][n - 1, n - 1] = -1.0          # last leading minor indefinite
    B = np.random.default_rng(3).standard_normal((k, 1, n))
    sentinel = np.triu(np.full((n, n), 7.5), 1)
    M = np.array([np.tril(A[b]) + sentinel for b in range(k)])
    tA = dev(api.from_matrices(torch.from_numpy(M)).numpy())
    tB = dev(B)
    info = api.posv_batched(tA, tB).cpu().tolist()
    _, oinfo = oracle.potrf(A[1])
    assert info[0] == 0 and info[2] == 0 and info[3] == 0 and info[1] == oinfo and oinfo > 0
    L = api.to_matrices(tA).cpu().numpy()
    for b in range(k):
        assert np.array_equal(np.triu(L[b], 1), sentinel)
    X = tB.cpu().numpy()
    for b in (0, 2, 3):
        assert backward_error(A[b:b + 1], X[b:b + 1], B[b:b + 1]) <= 1e-14


@pytest.mark.parametrize("n", [7, 16, 29, 32])
def test_small_fused_posv_ignores_upper_triangle(n):
    """n <= 32 fused path: only the lower triangle is an input.  NaN and inf
    in the strictly upper triangle give the same L and X, bit for bit, as a
    zero upper triangle (potrf's "upper part not referenced")."""
    k = 5
    A = spd_batch(k, n, seed=300 + n)
    B = np.random.default_rng(n).standard_normal((k, 2, n))
    results = []
    upper = np.triu(np.ones((n, n), dtype=bool), 1)
    for fill in (0.0, np.nan, np.inf):
        M = np.array([np.where(upper, fill, np.tril(A[b])) for b in range(k)])
        tA = dev(api.from_matrices(torch.from_numpy(M)).numpy())
        tB = dev(B)
        info = api.posv_batched(tA, tB).cpu().tolist()
        assert info == [0] * k
        L = np.tril(api.to_matrices(tA).cpu().numpy())
        results.append((L, tB.cpu().numpy()))
    for L, X in results[1:]:
        assert np.array_equal(L, results[0][0]) and np.array_equal(X, results[0][1])
    assert backward_error(A, results[0][1], B) <= 1e-14


@pytest.mark.parametrize("n,nrhs", [(9, 2), (32, 3)])
def test_small_fused_posv_padded_leading_dimensions(oracle, n, nrhs):
    """The C entry point with lda = n + 3, ldb = n + 1 and batch strides
    larger than one matrix: same factor and solution as the oracle, and the
    padding rows and the upper triangle are left as they were."""
    k, lda, ldb = 5, n + 3, n + 1
    A = spd_batch(k, n, seed=400 + n)
    B = np.random.default_rng(7).standard_normal((k, nrhs, n))
    Ab = np.full((k, n + 1, lda), 5.25)   # column j of matrix b at Ab[b, j, :]; one spare column per matrix
    for b in range(k):
        Ab[b, :n, :n] = A[b].T
    Bb = np.full((k, nrhs + 1, ldb), -3.5)
    Bb[:, :nrhs, :n] = B
    tA, tB = dev(Ab), dev(Bb)
    info = torch.empty(k, dtype=torch.int32, device="cuda")
    rc = api.lib().gputils_dposv_batched(n, nrhs, api._ptr(tA), lda, (n + 1) * lda, api._ptr(tB), ldb,
                                         (nrhs + 1) * ldb, api._ptr(info), k, api.stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * k
    Ao, Bo = tA.cpu().numpy(), tB.cpu().numpy()
    for b in range(k):
        Lg = Ao[b, :n, :n].T                                  # row i, column j
        Lo, _ = oracle.potrf(A[b])
        assert np.max(np.abs(np.tril(Lg) - np.tril(Lo))) <= 1e-12 * np.abs(Lo).max()
        assert np.array_equal(np.triu(Lg, 1), np.triu(A[b], 1))
        ref = oracle.potrs(Lo, B[b].T)
        assert np.max(np.abs(Bo[b, :nrhs, :n].T - ref)) <= 1e-11 * np.abs(ref).max()
    assert np.all(Ao[:, :n, n:] == 5.25) and np.all(Ao[:, n, :] == 5.25)
    assert np.all(Bo[:, :nrhs, n:] == -3.5) and np.all(Bo[:, nrhs, :] == -3.5)


@pytest.mark.parametrize("stride", [(1 << 28) - 1024, 1 << 28])
def test_small_fused_posv_batch_stride_near_2_gib(oracle, stride):
    """Two 20x20 systems (one wavefront) whose batch stride puts the second
    one about 2 GiB after the first: just below 2^31 bytes the factor is
    stored through the wavefront's range-checked buffer resource (offsets
    up to ~2^31), at 2^31 and beyond through per-lane masked stores.  Both
    give the oracle's L and x and leave the upper triangles and the bytes
    between the matrices as they were."""
    n, nrhs, k = 20, 1, 2
    A = spd_batch(k, n, seed=77)
    B = np.random.default_rng(78).standard_normal((k, nrhs, n))
    big = torch.full((stride + n * n,), 1.5, dtype=torch.float64, device="cuda")
    for b in range(k):
        big[b * stride:b * stride + n * n] = torch.from_numpy(np.ascontiguousarray(A[b].T).ravel()).cuda()
    tB = dev(B)
    info = torch.empty(k, dtype=torch.int32, device="cuda")
    rc = api.lib().gputils_dposv_batched(n, nrhs, api._ptr(big), n, stride, api._ptr(tB), n, n * nrhs,
                                         api._ptr(info), k, api.stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * k
    Bo = tB.cpu().numpy()
    for b in range(k):
        Lg = big[b * stride:b * stride + n * n].cpu().numpy().reshape(n, n).T
        Lo, _ = oracle.potrf(A[b])
        assert np.max(np.abs(np.tril(Lg) - np.tril(Lo))) <= 1e-12 * np.abs(Lo).max()
        assert np.array_equal(np.triu(Lg, 1), np.triu(A[b], 1))
        ref = oracle.potrs(Lo, B[b].T)
        assert np.max(np.abs(Bo[b].T - ref)) <= 1e-11 * np.abs(ref).max()
    gap = big[n * n:stride]
    assert bool((gap == 1.5).all())
    del big


def test_mixed_fallback_many_systems_per_workgroup():
    """The FP64 fallback launch covers the batch with 512 workgroups, so at
    k = 1100 a workgroup owns up to three systems (b, b + 512, b + 1024).
    Systems 5, 517 and 1029 share one workgroup: an indefinite matrix, an
    SPD one and an ill-conditioned one (the FP32 factor is unusable, FP64
    solves it).  Every other system converges in mixed precision."""
    k, n = 1100, 40
    r = np.random.default_rng(11)
    A = spd_batch(k, n, seed=12)
    Q, _ = np.linalg.qr(r.standard_normal((n, n)))
    ill = (Q * np.logspace(0, -12, n)) @ Q.T
    A[1029] = 0.5 * (ill + ill.T)
    A[5] = -A[5]
    A[600] = -A[600]
    B = r.standard_normal((k, 2, n))
    X, iters, info = api.posv_mixed_batched(dev(A), dev(B), tol=1e-12, max_iter=10)
    it, inf = iters.cpu().numpy(), info.cpu().numpy()
    X = X.cpu().numpy()
    assert it[5] == -1 and inf[5] == 1
    assert it[600] == -1 and inf[600] == 1
    assert it[1029] == -1 and inf[1029] == 0
    ok = np.ones(k, dtype=bool)
    ok[[5, 600, 1029]] = False
    assert np.all(it[ok] >= 0) and np.all(inf[ok] == 0)
    assert backward_error(A[ok], X[ok], B[ok]) <= 1e-12
    assert backward_error(A[1029:1030], X[1029:1030], B[1029:1030]) <= 1e-13


@pytest.mark.parametrize("n", [48, 128])
def test_blocked_and_mixed_ignore_upper_triangle(n):
    """n > 32 (LDS-blocked FP64 kernel) and the mixed solver: NaN or inf in
    the strictly upper triangle of A gives the same L, X, iterations and
    info, bit for bit, as a zero upper triangle, and the upper triangle is
    not written."""
    k = 6
    A = spd_batch(k, n, seed=500 + n)
    B = np.random.default_rng(n + 1).standard_normal((k, 3, n))
    upper = np.triu(np.ones((n, n), dtype=bool), 1)
    fused, mixed = [], []
    for fill in (0.0, np.nan, np.inf):
        M = np.array([np.where(upper, fill, np.tril(A[b])) for b in range(k)])
        Mt = api.from_matrices(torch.from_numpy(M)).numpy()
        tA, tB = dev(Mt), dev(B)
        X, it, inf = api.posv_mixed_batched(tA, tB, tol=1e-12, max_iter=10)
        mixed.append((X.cpu().numpy(), it.cpu().numpy(), inf.cpu().numpy()))
        info = api.posv_batched(tA, tB).cpu().tolist()
        assert info == [0] * k
        L = api.to_matrices(tA).cpu().numpy()
        if fill == 0.0:
            assert np.all(L[:, upper] == 0.0)
        elif np.isnan(fill):
            assert np.all(np.isnan(L[:, upper]))
        else:
            assert np.all(L[:, upper] == fill)
        fused.append((np.tril(L), tB.cpu().numpy()))
    for (L, X), (Xm, itm, infm) in zip(fused[1:], mixed[1:]):
        assert np.array_equal(L, fused[0][0]) and np.array_equal(X, fused[0][1])
        assert np.array_equal(Xm, mixed[0][0]) and np.array_equal(itm, mixed[0][1]) and np.array_equal(infm, mixed[0][2])
    assert backward_error(A, fused[0][1], B) <= 1e-14
    assert backward_error(A, mixed[0][0], B) <= 1e-12


@pytest.mark.parametrize("n,nrhs", [(9, 3), (16, 2), (32, 3)])
def test_small_potrs_given_factor_padded(oracle, n, nrhs):
    """n <= 32 solve only (potrs) from the oracle's factor, through the C
    entry point with lda = n + 2, ldb = n + 3: matches the oracle's solve,
    leaves L (both triangles) and the padding untouched."""
    k, lda, ldb = 7, n + 2, n + 3
    A = spd_batch(k, n, seed=600 + n)
    Ls = [oracle.potrf(A[b])[0] for b in range(k)]
    B = np.random.default_rng(9).standard_normal((k, nrhs, n))
    Lb = np.full((k, n, lda), 2.5)
    for b in range(k):
        Lb[b, :, :n] = (np.tril(Ls[b]) + np.triu(np.full((n, n), 4.0), 1)).T   # column j at Lb[b, j, :]
    Bb = np.full((k, nrhs, ldb), -1.5)
    Bb[:, :, :n] = B
    tL, tB = dev(Lb), dev(Bb)
    info = torch.full((k,), 99, dtype=torch.int32, device="cuda")
    rc = api.lib().gputils_dpotrs_batched(n, nrhs, api._ptr(tL), lda, n * lda, api._ptr(tB), ldb, nrhs * ldb,
                                          api._ptr(info), k, api.stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * k
    assert np.array_equal(tL.cpu().numpy(), Lb)
    Xo = tB.cpu().numpy()
    assert np.all(Xo[:, :, n:] == -1.5)
    for b in range(k):
        ref = oracle.potrs(Ls[b], B[b].T)
        assert np.max(np.abs(Xo[b, :, :n].T - ref)) <= 1e-11 * np.abs(ref).max()
