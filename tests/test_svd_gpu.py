"""GPU parity of the batched SVD, rank count, nullspace construction and
projection against the oracle and the reference suite's golden values
(test/testTensor.cu X:1058-1193 for SVD/rank, X:1552-1657 for Nullspace).

Singular values agree with the FP64 oracle to 1e-12 relative to sigma_max
(FP64 data; float data to 2e-6).  Singular vectors of distinct nonzero
singular values agree element-wise, signs included (both sides put the
largest-magnitude entry of v_j negative).  Vectors of zero or repeated
singular values are compared as subspaces (projectors), as the reference
pins them only up to basis."""
import numpy as np
import pytest
import torch

import golden_values as G
from gputils_amd import api

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).cuda()


def tensor_of(mats, dtype):
    """(k, m, n) matrices -> device tensor in (k, n, m) storage."""
    return dev(api.from_matrices(torch.from_numpy(np.ascontiguousarray(mats))).numpy(), dtype)


def svd(mats, dtype, compute_u=True):
    S, U, VT, info, sw = api.gesvd_batched(tensor_of(mats, dtype), compute_u=compute_u, sweeps=True)
    S = S.cpu().double().numpy()[:, 0, :]
    VT = api.to_matrices(VT).cpu().double().numpy()
    U = api.to_matrices(U).cpu().double().numpy() if U is not None else None
    return S, U, VT, info.cpu().numpy(), sw.cpu().numpy()


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_singular_values_golden(dtype, eps):
    """X:1058-1078: sigma_0, sigma_1 of an 8 x 3 matrix (MATLAB values)."""
    S, U, VT, info, _ = svd(G.cm(G.SVD8_B, 8, 3), dtype)
    assert info.tolist() == [0]
    assert abs(S[0, 0] - G.SVD8_S01[0]) < eps and abs(S[0, 1] - G.SVD8_S01[1]) < eps
    assert U is not None and U.shape == (1, 8, 8)


@pytest.mark.parametrize("dtype,eps", [(torch.float32, 10 * G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_singular_vectors_multiple_matrices_golden(dtype, eps):
    """X:1117-1166: three 3 x 2 matrices (ranks 2, 1, 1): all six sigma, the
    first four entries of V' and all 27 entries of U, element-wise."""
    A = G.cm(G.SVD3_A, 3, 2, 3)
    S, U, VT, info, _ = svd(A, dtype)
    assert np.allclose(S.ravel(), G.SVD3_S, atol=eps)
    vt_flat = np.concatenate([VT[b].ravel(order="F") for b in range(3)])
    assert np.allclose(vt_flat[:4], G.SVD3_V4, atol=eps)
    u_flat = np.concatenate([U[b].ravel(order="F") for b in range(3)])
    assert np.allclose(u_flat, G.SVD3_U, atol=eps), np.round(u_flat - np.array(G.SVD3_U), 6)


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_rank_golden(dtype, eps):
    """X:1175-1193: ranks 3, 2, 1 of three 4 x 3 matrices."""
    A = tensor_of(G.cm(G.RANK_A, 4, 3, 3), dtype)
    S, _, _, _ = api.gesvd_batched(A)
    r = api.svd_rank_batched(S, eps)
    assert r.cpu().tolist() == G.RANK_EXPECTED
    # the count is written, not accumulated: a second call gives the same
    assert api.svd_rank_batched(S, eps).cpu().tolist() == G.RANK_EXPECTED


SHAPES = [(1, 1), (5, 1), (3, 2), (10, 4), (17, 9), (30, 30), (64, 64), (96, 48), (128, 64), (129, 40),
          (200, 50), (300, 64), (1000, 20),  # U in registers up to m = 128, in column chunks beyond
          (65, 65), (100, 96), (128, 128), (200, 130)]  # n > 64: G, W, V in the workspace


@pytest.mark.parametrize("m,n", SHAPES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_svd_matches_oracle(oracle, m, n, dtype):
    k = 3
    r = np.random.default_rng(1000 * m + n)
    A = r.standard_normal((k, m, n))
    if dtype == torch.float32:
        A = A.astype(np.float32).astype(np.float64)
    S, U, VT, info, sw = svd(A, dtype)
    assert info.tolist() == [0] * k
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    for b in range(k):
        So, Uo, VTo, _ = oracle.gesvd(A[b].copy())
        smax = So[0]
        assert np.max(np.abs(S[b] - So)) <= tol * smax * max(1, n / 8)
        # distinct, well separated sigma: vectors agree element-wise
        gap = np.min(np.abs(np.diff(So))) / smax if n > 1 else 1.0
        vtol = tol * 100 / max(gap, 1e-3)
        assert np.max(np.abs(VT[b] - VTo)) <= vtol, (b, np.max(np.abs(VT[b] - VTo)))
        assert np.max(np.abs(U[b][:, :n] - Uo[:, :n])) <= vtol
        # U orthogonal, completion spans the complement, reconstruction
        assert np.allclose(U[b].T @ U[b], np.eye(m), atol=tol * 100)
        Pc, Po = U[b][:, n:] @ U[b][:, n:].T, Uo[:, n:] @ Uo[:, n:].T
        assert np.allclose(Pc, Po, atol=tol * 100)
        assert np.linalg.norm(U[b][:, :n] * S[b] @ VT[b] - A[b]) <= tol * 10 * smax * np.sqrt(m * n)
    s64, s32 = sw & 0xFFFF, sw >> 16
    assert s64.max() <= 20
    assert s32.max() == 0  # with VT the sweeps are all FP64


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_svd_without_u_and_read_only_input(dtype):
    r = np.random.default_rng(5)
    A = r.standard_normal((4, 40, 12))
    tA = tensor_of(A, dtype)
    before = tA.clone()
    S1, U1, VT1, _ = api.gesvd_batched(tA, compute_u=False)
    S2, U2, VT2, _ = api.gesvd_batched(tA, compute_u=True)
    assert U1 is None
    assert torch.equal(tA, before)
    assert torch.equal(S1, S2) and torch.equal(VT1, VT2)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_svd_rank_deficient_and_zero(oracle, dtype):
    """Exact low-rank constructions, a zero matrix and zero columns: the
    trailing sigma vanish, rank is exact, U and V stay orthogonal and the
    null directions agree with the oracle as subspaces."""
    r = np.random.default_rng(11)
    mats = []
    for (m, n, rk) in [(20, 8, 3)] * 3:
        mats.append(r.standard_normal((m, rk)) @ r.standard_normal((rk, n)))
    mats = np.stack(mats)
    mats[1] = 0.0
    mats[2][:, [1, 4]] = 0.0
    S, U, VT, info, _ = svd(mats, dtype)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    ranks = api.svd_rank_batched(dev(S[:, None, :], dtype), 1e-6 if dtype == torch.float64 else 1e-3)
    for b in range(3):
        So, Uo, VTo, _ = oracle.gesvd(mats[b].copy())
        smax = max(So[0], 1.0)
        assert np.max(np.abs(S[b] - So)) <= tol * smax * 10
        assert np.allclose(U[b].T @ U[b], np.eye(20), atol=tol * 10)
        assert np.allclose(VT[b] @ VT[b].T, np.eye(8), atol=tol * 10)
        rk = int(np.sum(So > 1e-6 * smax))
        assert ranks[b].item() == rk
        assert np.allclose(U[b][:, :rk] @ U[b][:, :rk].T, Uo[:, :rk] @ Uo[:, :rk].T, atol=tol * 100)
        assert np.allclose(VT[b][:rk].T @ VT[b][:rk], VTo[:rk].T @ VTo[:rk], atol=tol * 100)
    # the zero matrix: U = I and V' = I, as LAPACK leaves them
    assert np.array_equal(S[1], np.zeros(8))
    assert np.allclose(U[1], np.eye(20)) and np.allclose(VT[1], np.eye(8))


@pytest.mark.parametrize("m,n", [(96, 48), (20, 8), (64, 64), (5, 1), (30, 17), (120, 80), (96, 96)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_svd_without_vt_same_values_and_subspaces(m, n, dtype):
    """VT = NULL (the nullspace builds' mode) runs the sweeps on R^T: the
    singular values equal the full SVD's, U is orthogonal, ||A^T u_j|| =
    sigma_j for the nonzero sigma, and the projectors onto the leading and
    the trailing columns of U agree with the full SVD's (the column signs
    may differ).  The second matrix is rank deficient."""
    r = np.random.default_rng(100 * m + n)
    rk = max(1, n // 2)
    A = np.stack([r.standard_normal((m, n)), r.standard_normal((m, rk)) @ r.standard_normal((rk, n))])
    if dtype == torch.float32:
        A = A.astype(np.float32).astype(np.float64)
    tA = tensor_of(A, dtype)
    S1, U1, VT1, info1 = api.gesvd_batched(tA, compute_u=True)
    S2, U2, VT2, info2 = api.gesvd_batched(tA, compute_u=True, compute_v=False)
    assert VT2 is None
    assert info1.tolist() == [0, 0] and info2.tolist() == [0, 0]
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    S1, S2 = S1.cpu().double().numpy()[:, 0, :], S2.cpu().double().numpy()[:, 0, :]
    U1 = np.swapaxes(U1.cpu().double().numpy(), 1, 2)
    U2 = np.swapaxes(U2.cpu().double().numpy(), 1, 2)
    for b in range(2):
        smax = max(S1[b, 0], 1.0)
        assert np.max(np.abs(S1[b] - S2[b])) <= 10 * tol * smax
        assert np.allclose(U2[b].T @ U2[b], np.eye(m), atol=100 * tol)
        rb = int(np.sum(S1[b] > 1e-6 * smax))
        for j in range(rb):
            assert abs(np.linalg.norm(A[b].T @ U2[b][:, j]) - S2[b, j]) <= 100 * tol * smax
        for lo, hi in ((0, rb), (rb, m)):
            P1 = U1[b][:, lo:hi] @ U1[b][:, lo:hi].T
            P2 = U2[b][:, lo:hi] @ U2[b][:, lo:hi].T
            assert np.allclose(P1, P2, atol=1e3 * tol), (b, lo, hi, np.abs(P1 - P2).max())


@pytest.mark.parametrize("m,n,rk", [(96, 48, 48), (96, 48, 40), (64, 64, 64), (200, 50, 50), (40, 12, 5),
                                    (10, 4, 4)])
def test_fp32_first_sweeps_match_fp64_sweeps(oracle, m, n, rk):
    """Float data without VT (the nullspace builds' SVD): FP32 sweeps, then
    FP64 sweeps from their result, against all-FP64 sweeps
    (gputils_gesvd_set_fp32_sweeps(0)) and the FP64 oracle: the same
    singular values to FP32 rounding (2e-6 sigma_max), U orthogonal to FP32
    rounding, the same subspaces for the nonzero and the zero singular
    values, and at most four FP64 sweeps (fewer than all-FP64 takes).  With
    VT the sweeps stay FP64: bit-identical to the switch turned off."""
    k = 4
    r = np.random.default_rng(7 * m + n + rk)
    A = np.einsum("kir,krj->kij", r.standard_normal((k, m, rk)), r.standard_normal((k, rk, n)))
    A = A.astype(np.float32).astype(np.float64)
    tA = tensor_of(A, torch.float32)
    runs = {}
    for mixed in (False, True):
        prev = api.set_fp32_sweeps(mixed)
        try:
            runs[mixed] = [api.gesvd_batched(tA, compute_u=True, sweeps=True, compute_v=cv) for cv in (False, True)]
        finally:
            api.set_fp32_sweeps(prev)
    assert api.set_fp32_sweeps(True) is True  # the default was restored
    (S0, U0, _, info0, sw0), (Sv0, Uv0, VT0, _, _) = runs[False]
    (S1, U1, _, info1, sw1), (Sv1, Uv1, VT1, _, swv1) = runs[True]
    assert torch.equal(Sv0, Sv1) and torch.equal(Uv0, Uv1) and torch.equal(VT0, VT1)
    assert int(api.sweep_counts(swv1)[1].max()) == 0
    assert info0.tolist() == [0] * k and info1.tolist() == [0] * k
    (a64, a32), (b64, b32) = api.sweep_counts(sw0), api.sweep_counts(sw1)
    assert int(a32.max()) == 0 and int(b32.min()) >= 1
    assert int(b64.max()) <= 4 and int(b64.max()) < int(a64.min()), (a64, b64, b32)
    S0, S1 = S0.cpu().double().numpy()[:, 0, :], S1.cpu().double().numpy()[:, 0, :]
    U0, U1 = api.to_matrices(U0).cpu().double().numpy(), api.to_matrices(U1).cpu().double().numpy()
    for b in range(k):
        So = oracle.gesvd(A[b].copy())[0]
        smax = So[0]
        assert np.max(np.abs(S1[b] - So)) <= 2e-6 * smax
        assert np.max(np.abs(S1[b] - S0[b])) <= 2e-6 * smax
        assert np.allclose(U1[b].T @ U1[b], np.eye(m), atol=2e-5)
        for lo, hi in ((0, rk), (rk, m)):
            P0 = U0[b][:, lo:hi] @ U0[b][:, lo:hi].T
            P1 = U1[b][:, lo:hi] @ U1[b][:, lo:hi].T
            assert np.abs(P1 - P0).max() <= 1e-4, (b, lo, hi)


def test_svd_argument_checks():
    A = torch.zeros((2, 3, 5), dtype=torch.float64, device="cuda")   # 5 x 3: fine
    api.gesvd_batched(A)
    with pytest.raises(api.GputilsError):
        api.gesvd_batched(torch.zeros((2, 5, 3), dtype=torch.float64, device="cuda"))  # fat
    S, _, _, info = api.gesvd_batched(torch.zeros((1, 65, 80), dtype=torch.float64, device="cuda"))  # n > 64
    assert info.tolist() == [0] and torch.equal(S, torch.zeros_like(S))


# ------------------------------------------------------------ nullspace --
@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_nullspace_golden(oracle, dtype, eps):
    """X:1556-1585: A N = 0 and N^T N = diag(I, 0) for five 3 x 4 matrices,
    and P equal to the oracle's (basis independent)."""
    A = G.cm(G.NS_A, 3, 4, 5)
    N, P, rank = api.nullspace_batched(tensor_of(A, dtype))
    N = api.to_matrices(N).cpu().double().numpy()
    P = api.to_matrices(P).cpu().double().numpy()
    for b in range(5):
        assert np.linalg.norm(A[b] @ N[b]) < eps
        NtN = N[b].T @ N[b]
        nul = 4 - np.linalg.matrix_rank(A[b])
        assert rank[b].item() == 4 - nul if nul < 4 else rank[b].item() == 0
        E = np.zeros((4, 4))
        E[:nul, :nul] = np.eye(nul)
        assert np.allclose(NtN, E, atol=eps)
        No, Po, nul_o = oracle.nullspace(A[b])
        assert nul_o == nul
        assert np.allclose(P[b], Po, atol=eps)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_nullspace_trivial_is_exactly_zero(dtype):
    """X:1597-1613: full-rank square matrices give N = 0 exactly."""
    A = np.array(G.NS_TRIVIAL_ROWMAJOR, dtype=float).reshape(2, 3, 3)
    N, P, rank = api.nullspace_batched(tensor_of(A, dtype))
    assert rank.cpu().tolist() == [3, 3]
    assert torch.count_nonzero(N).item() == 0 and torch.count_nonzero(P).item() == 0


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_project_on_nullspace_golden(dtype, eps):
    """X:1624-1657: A (P x) = 0 and (N o - P x) . (P x - x) = 0."""
    A = np.array(G.PROJ_A_ROWMAJOR, dtype=float).reshape(1, 3, 7)
    N, P, _ = api.nullspace_batched(tensor_of(A, dtype))
    x = dev(np.array(G.PROJ_X, dtype=float).reshape(1, 1, 7), dtype)
    proj = api.project(P, x.clone())
    pn = proj.cpu().double().numpy()[0, 0]
    assert np.linalg.norm(A[0] @ pn) < eps
    Nm = api.to_matrices(N).cpu().double().numpy()[0]
    y = Nm @ np.array(G.PROJ_OTHER, dtype=float)
    assert (y - pn) @ (pn - np.array(G.PROJ_X, dtype=float)) < eps


def test_nullspace_config(oracle):
    """2048 rank-40 48 x 96 FP32 matrices (sigma in [0.1, 10]): rank 40
    everywhere, P equal to the oracle's on a sample, A P x ~ 0 for all."""
    from gputils_amd import workloads
    A_t, x = workloads.svd_nullspace()
    tA = dev(A_t)
    N, P, rank = api.nullspace_batched(tA)
    assert rank.cpu().tolist() == [40] * 2048
    Pm = api.to_matrices(P).cpu().double().numpy()
    Am = api.to_matrices(torch.from_numpy(A_t)).double().numpy()
    for b in np.random.default_rng(0).choice(2048, 16, replace=False):
        _, Po, nul = oracle.nullspace(Am[b])
        assert nul == 56
        assert np.max(np.abs(Pm[b] - Po)) < 2e-6
    tb = dev(x)
    api.project(P, tb)
    z = tb.cpu().double().numpy()[:, 0, :]
    res = np.einsum("kij,kj->ki", Am, z)
    assert np.max(np.linalg.norm(res, axis=1) / np.linalg.norm(x[:, 0, :], axis=1)) < 1e-5


@pytest.mark.parametrize("m,n,rk", [(80, 120, 70), (96, 130, 96), (65, 65, 30)])
def test_nullspace_of_wide_matrices(oracle, m, n, rk):
    """More than 64 rows (the SVD of A^T then has more than 64 columns):
    rank and P agree with the oracle; A N = 0 and N^T N = diag(I, 0)."""
    k = 2
    r = np.random.default_rng(m + n)
    A = np.einsum("kir,krj->kij", r.standard_normal((k, m, rk)), r.standard_normal((k, rk, n)))
    tA = dev(api.from_matrices(torch.from_numpy(A)).numpy())
    N, P, rank = api.nullspace_batched(tA)
    assert rank.cpu().tolist() == [rk] * k
    Nm = api.to_matrices(N).cpu().numpy()
    Pm = api.to_matrices(P).cpu().numpy()
    nul = n - rk
    for b in range(k):
        _, Po, onul = oracle.nullspace(A[b])
        assert onul == nul
        assert np.max(np.abs(Pm[b] - Po)) < 1e-11
        assert np.linalg.norm(A[b] @ Nm[b]) < 1e-10 * np.linalg.norm(A[b])
        G_ = Nm[b].T @ Nm[b]
        want = np.zeros((n, n))
        want[:nul, :nul] = np.eye(nul)
        assert np.max(np.abs(G_ - want)) < 1e-12


def test_config_project_1000_calls_in_one_graph():
    """The configuration's projection at full size: P of the 2048 rank-40
    48 x 96 FP32 matrices; 1000 Nullspace::project calls captured in one HIP
    graph and replayed twice, with eager work between the replays.  Each
    replay equals 1000 eager calls bit for bit, and ||A P x|| <= 1e-5 ||x||
    for every matrix (FP64 on the host)."""
    from gputils_amd import workloads
    A_t, x = workloads.svd_nullspace()
    _, P, rank = api.nullspace_batched(dev(A_t))
    assert rank.cpu().tolist() == [40] * 2048
    x0 = dev(x)
    eager = x0.clone()
    for _ in range(1000):
        api.project(P, eager)
    b = x0.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        api.project(P, b)            # warm-up outside capture
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            for _ in range(1000):
                api.project(P, b)
    torch.cuda.synchronize()
    Am = api.to_matrices(torch.from_numpy(A_t)).double().numpy()
    for rep in range(2):
        b.copy_(x0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b, eager), rep
        z = b.cpu().double().numpy()[:, 0, :]
        res = np.einsum("kij,kj->ki", Am, z)
        assert np.max(np.linalg.norm(res, axis=1) / np.linalg.norm(x[:, 0, :], axis=1)) <= 1e-5
        _ = (x0 * 2).sum().item()   # eager work between replays


@pytest.mark.parametrize("m,n,rk", [(48, 96, 40), (3, 4, 2), (5, 5, 5), (20, 33, 7), (64, 130, 64), (10, 12, 0),
                                    (16, 32, 16), (32, 32, 32), (30, 64, 20), (40, 48, 30), (120, 128, 100),
                                    (96, 96, 90), (4, 96, 1)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_nullspace_project_from_n_equals_p(m, n, rk, dtype):
    """Nullspace::project as N (N^T b) from N's nullity columns against
    P b: the same projection to the type's rounding; A y = 0 for the
    result; full rank (nullity 0) gives y = 0 exactly."""
    k = 5
    r = np.random.default_rng(10 * m + n + rk)
    A = (np.einsum("kir,krj->kij", r.standard_normal((k, m, rk)), r.standard_normal((k, rk, n)))
         if rk > 0 else np.zeros((k, m, n)))
    tA = dev(api.from_matrices(torch.from_numpy(A)).numpy(), dtype)
    N, P, rank = api.nullspace_batched(tA)
    x = dev(r.standard_normal((k, 1, n)), dtype)
    yp = api.project(P, x.clone())
    yn = api.nullspace_project(N, rank, x.clone())
    tol = 1e-5 if dtype == torch.float32 else 1e-12
    scale = float(x.abs().max())
    assert float((yp - yn).abs().max()) <= tol * scale
    Z = yn.cpu().double().numpy()[:, 0, :]
    for b in range(k):
        assert np.linalg.norm(A[b] @ Z[b]) <= 10 * tol * max(np.linalg.norm(A[b]), 1.0) * scale * n
        if int(rank[b]) >= n:
            assert np.all(Z[b] == 0.0)


def test_config_nullspace_project_1000_calls_in_one_graph():
    """The configuration's projection through Nullspace::project's N form
    at full size: 1000 calls captured in one HIP graph, replayed twice with
    eager work between; each replay equals 1000 eager calls bit for bit and
    ||A y|| <= 1e-5 ||x|| for every matrix."""
    from gputils_amd import workloads
    A_t, x = workloads.svd_nullspace()
    N, P, rank = api.nullspace_batched(dev(A_t))
    x0 = dev(x)
    eager = x0.clone()
    for _ in range(1000):
        api.nullspace_project(N, rank, eager)
    b = x0.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        api.nullspace_project(N, rank, b)
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            for _ in range(1000):
                api.nullspace_project(N, rank, b)
    torch.cuda.synchronize()
    Am = api.to_matrices(torch.from_numpy(A_t)).double().numpy()
    for rep in range(2):
        b.copy_(x0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b, eager), rep
        z = b.cpu().double().numpy()[:, 0, :]
        res = np.einsum("kij,kj->ki", Am, z)
        assert np.max(np.linalg.norm(res, axis=1) / np.linalg.norm(x[:, 0, :], axis=1)) <= 1e-5
        _ = (x0 * 2).sum().item()


def _project_reference(N, rank, b):
    """y = N_q (N_q^T b) in float64 on the device, N_q the first n - rank
    columns of each N (column j of matrix i is N[i, j, :])."""
    k, n, _ = N.shape
    q = (n - rank.to(torch.int64).clamp(max=n)).to(N.device)
    keep = (torch.arange(n, device=N.device)[None, :] < q[:, None]).to(torch.float64)
    Nd = N.to(torch.float64) * keep[:, :, None]
    z = torch.bmm(Nd, b.to(torch.float64).transpose(1, 2).contiguous())       # (k, n, 1)
    return torch.bmm(Nd.transpose(1, 2).contiguous(), z).transpose(1, 2)     # (k, 1, n)


@pytest.mark.parametrize("n,dtype,k", [(32, torch.float32, 65540),   # lane groups, two grid rows
                                       (64, torch.float64, 300),     # lane groups, FP64
                                       (12, torch.float32, 300),     # vector path
                                       (96, torch.float64, 300),     # vector path, FP64
                                       (33, torch.float32, 300)])    # generic path
def test_nullspace_project_paths_match_float64(n, dtype, k):
    """Every path of the projection kernel (8-lane column groups, vector
    slices, scalar) on random N and ranks from 0 to beyond n, against
    float64 on the device: |y - y64| <= 4 u n max|N|^2 max|b| (u the type's
    unit roundoff), and y = 0 exactly where rank >= n."""
    g = torch.Generator(device="cuda").manual_seed(n * 7 + k)
    N = torch.randn((k, n, n), generator=g, device="cuda", dtype=dtype)
    rank = torch.randint(0, n + 3, (k,), generator=g, device="cuda", dtype=torch.int32)
    b = torch.randn((k, 1, n), generator=g, device="cuda", dtype=dtype)
    ref = _project_reference(N, rank, b)
    y = api.nullspace_project(N, rank, b.clone())
    u = torch.finfo(dtype).eps / 2
    bound = 4 * u * n * float(N.abs().max()) ** 2 * float(b.abs().max()) * n
    assert float((y.to(torch.float64) - ref).abs().max()) <= bound
    full = rank.to(torch.int64) >= n
    assert bool((y[full] == 0).all())


def test_nullspace_project_unaligned_b_takes_the_scalar_path():
    """b 4 bytes past a 16-byte boundary: the library falls back to the
    scalar path and gives the same projection (checked against float64)."""
    from gputils_amd import _lib
    import ctypes
    k, n = 257, 32
    g = torch.Generator(device="cuda").manual_seed(5)
    N = torch.randn((k, n, n), generator=g, device="cuda")
    rank = torch.randint(0, n, (k,), generator=g, device="cuda", dtype=torch.int32)
    buf = torch.zeros(k * n + 1, device="cuda")
    b = buf[1:].view(k, 1, n)
    b.copy_(torch.randn((k, 1, n), generator=g, device="cuda"))
    ref = _project_reference(N, rank, b)
    bmax = float(b.abs().max())
    rc = _lib.lib().gputils_snullspace_project_batched(n, ctypes.c_void_p(N.data_ptr()), n, n * n,
                                                       ctypes.c_void_p(rank.data_ptr()),
                                                       ctypes.c_void_p(b.data_ptr()), n, k,
                                                       ctypes.c_void_p(api.stream_handle()))
    assert rc == 0
    torch.cuda.synchronize()
    bound = 4 * 2.0 ** -24 * n * float(N.abs().max()) ** 2 * bmax * n
    assert float((b.to(torch.float64) - ref).abs().max()) <= bound
    assert float(buf[0]) == 0.0   # nothing written before b


def test_nullspace_project_argument_checks():
    N = torch.zeros((2, 4, 4), device="cuda")
    rank = torch.zeros(2, dtype=torch.int32, device="cuda")
    with pytest.raises(api.GputilsError):
        api.nullspace_project(N, rank, torch.zeros((2, 1, 5), device="cuda"))
    with pytest.raises(api.GputilsError):
        api.nullspace_project(N, rank, torch.zeros((2, 1, 4), dtype=torch.float64, device="cuda"))


def test_project_graph_replays_match_eager():
    """Nullspace::project captured in a HIP graph and replayed repeatedly,
    with eager work in between: every replay equals the eager result."""
    r = np.random.default_rng(2)
    k, n = 64, 96
    Q = np.linalg.qr(r.standard_normal((k, n, n)))[0][:, :, :30]
    P = dev(api.from_matrices(torch.from_numpy(Q @ np.swapaxes(Q, 1, 2))).numpy(), torch.float32)
    b0 = dev(r.standard_normal((k, 1, n)), torch.float32)
    eager = api.project(P, b0.clone())
    b = b0.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        api.project(P, b)            # warm-up outside capture
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            api.project(P, b)
    torch.cuda.synchronize()
    for rep in range(4):
        b.copy_(b0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b, eager), rep
        _ = (b0 * 2).sum().item()   # eager work between replays
