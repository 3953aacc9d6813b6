"""GPU parity of batched GEMM, the BLAS-1 reductions, scal/axpy and the
batched transpose against the FP64 oracle and the reference suite's golden
values (test/testTensor.cu X:488-573, 828-853, 934-950).

GEMM tolerances are the forward error bounds for the arithmetic used:
FP32 |dC| <= k * 2^-24 |A||B| (times a safety factor); BF16X3 drops the
lo*lo term and rounds the split, bounded by k * 2^-15 |A||B|; FP64 k * 2^-53."""
import numpy as np
import pytest
import torch

import golden_values as G
from gputils_amd import api

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).cuda()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_addab_golden(dtype):
    A = dev(np.array(G.ADDAB_A, dtype=np.float64).reshape(3, 3, 2), dtype)   # (k, n=3, m=2)
    B = dev(np.array(G.ADDAB_B, dtype=np.float64).reshape(3, 2, 3), dtype)   # (k, n=2, kk=3)
    C = api.gemm_batched(A, B)
    assert C.cpu().double().numpy().ravel().tolist() == G.ADDAB_C


def _gemm_case(oracle, m, n, kk, batch, dtype, mode, alpha=1.0, beta=0.0, seed=0):
    r = np.random.default_rng(seed)
    A = r.uniform(-1, 1, (batch, kk, m))
    B = r.uniform(-1, 1, (batch, n, kk))
    C0 = r.uniform(-1, 1, (batch, n, m))
    if dtype == torch.float32:
        A, B, C0 = (x.astype(np.float32).astype(np.float64) for x in (A, B, C0))
    tC = dev(C0, dtype)
    api.gemm_batched(dev(A, dtype), dev(B, dtype), tC, alpha=alpha, beta=beta, mode=mode)
    got = tC.cpu().double().numpy()
    for b in range(batch):
        ref = oracle.gemm(A[b].T, B[b].T, C0[b].T, alpha=alpha, beta=beta)
        bound = np.abs(A[b].T) @ np.abs(B[b].T) * abs(alpha) + abs(beta) * np.abs(C0[b].T)
        unit = {"fp32": 2.0 ** -24, "bf16x3": 2.0 ** -15}.get(mode, 2.0 ** -53) if dtype == torch.float32 \
            else 2.0 ** -53
        err = np.abs(got[b].T - ref)
        assert np.all(err <= 4 * (kk + 2) * unit * bound + 1e-30), (m, n, kk, b, err.max())


@pytest.mark.parametrize("m,n,kk", [(1, 1, 1), (2, 2, 3), (5, 7, 3), (64, 64, 64), (65, 63, 33),
                                    (96, 1, 96), (128, 200, 70), (16, 3, 100),
                                    # register-direct path: m, n <= 64, k a multiple of 16 up to 64
                                    (50, 33, 48), (64, 64, 16), (17, 64, 32), (64, 5, 64), (3, 64, 64),
                                    # matrix-vector products (row-quad GEMV path when m % 4 == 0)
                                    (100, 1, 7), (1024, 1, 40), (98, 1, 5)])
@pytest.mark.parametrize("dtype,mode", [(torch.float32, "fp32"), (torch.float32, "bf16x3"),
                                        (torch.float64, "fp64")])
def test_gemm_matches_oracle(oracle, m, n, kk, dtype, mode):
    _gemm_case(oracle, m, n, kk, 3, dtype, mode, seed=m * 7 + n)


@pytest.mark.parametrize("m,n,kk", [(40, 24, 17), (64, 40, 32), (30, 64, 64), (64, 1, 33)])
@pytest.mark.parametrize("dtype,mode", [(torch.float32, "fp32"), (torch.float32, "bf16x3"),
                                        (torch.float64, "fp64")])
def test_gemm_alpha_beta(oracle, m, n, kk, dtype, mode):
    _gemm_case(oracle, m, n, kk, 4, dtype, mode, alpha=-1.5, beta=0.5, seed=3)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gemm_in_place_b_equals_c(oracle, dtype):
    """C aliases B as in Nullspace::project (tensor.cuh:2084)."""
    r = np.random.default_rng(1)
    batch, n = 5, 96
    P = r.standard_normal((batch, n, n))
    x = r.standard_normal((batch, 1, n))
    tP, tx = dev(P, dtype), dev(x, dtype)
    api.gemm_batched(tP, tx, tx)
    got = tx.cpu().double().numpy()
    for b in range(batch):
        ref = P[b].T @ x[b, 0]
        assert np.allclose(got[b, 0], ref, rtol=1e-4 if dtype == torch.float32 else 1e-12,
                           atol=1e-3 if dtype == torch.float32 else 1e-11)


def test_gemm_many_batches_grid_split(oracle):
    """More than 65535 matrices exercises the 2-D batch grid."""
    batch = 70000
    r = np.random.default_rng(2)
    A = r.uniform(-1, 1, (batch, 2, 3)).astype(np.float32)
    B = r.uniform(-1, 1, (batch, 2, 2)).astype(np.float32)
    C = api.gemm_batched(dev(A), dev(B)).cpu().numpy()
    for b in (0, 65534, 65535, 65536, batch - 1):
        assert np.allclose(C[b].T, A[b].T.astype(np.float64) @ B[b].T, atol=1e-5)
    A2 = r.uniform(-1, 1, (batch, 64, 64)).astype(np.float32)
    C2 = api.gemm_batched(dev(A2), dev(A2)).cpu().numpy()
    for b in (0, 65535, batch - 1):
        assert np.allclose(C2[b].T, A2[b].T.astype(np.float64) @ A2[b].T, atol=1e-4)


@pytest.mark.parametrize("mode,bound", [("fp32", 64 * 2.0 ** -24), ("bf16x3", 2.0 ** -14)])
def test_config_gemm_full_size(mode, bound):
    """The configuration itself: 16384 products of 64x64 by 64x64 FP32
    matrices (U[-1,1], seeded), each mode checked on every entry against
    the FP64 host product: max |dC| / (|A||B|) within the mode's bound
    (FP32: k u32; split-BF16: 2^-14, twice the 2^-15 split error)."""
    from gputils_amd import workloads
    A, B = workloads.gemm()
    C = api.gemm_batched(dev(A), dev(B), mode=mode).cpu().numpy().astype(np.float64)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    # tensor layout (k, n, m): C^T = B^T A^T per matrix
    ref = np.matmul(B64, A64)
    mag = np.matmul(np.abs(B64), np.abs(A64))
    err = np.max(np.abs(C - ref) / mag)
    assert err <= bound, err
    if mode == "bf16x3":
        assert err > 64 * 2.0 ** -24 / 16  # the split error is visible: this is not the FP32 path


@pytest.mark.parametrize("dtype,eps", [(torch.float32, G.PRECISION_LOW), (torch.float64, G.PRECISION_HIGH)])
def test_reductions_golden(dtype, eps):
    a = dev(np.array(G.T234A, dtype=np.float64), dtype)
    b = dev(np.array(G.T234B, dtype=np.float64), dtype)
    amb = dev(np.array(G.T234AMB, dtype=np.float64), dtype)
    assert float(api.dot(a, b)) == G.DOT_AB
    assert abs(float(api.nrm2(a)) - G.NORMF_A) < eps
    assert abs(float(api.asum(a)) - G.SUMABS_A) < 1e-10
    assert float(api.amax(amb)) == G.MAXABS_AMB
    assert float(api.amin(amb)) == G.MINABS_AMB


@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 100003, 3_000_001])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_reductions_match_numpy(n, dtype):
    r = np.random.default_rng(n)
    x = r.standard_normal(n)
    y = r.standard_normal(n)
    if dtype == torch.float32:
        x, y = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
    tx, ty = dev(x, dtype), dev(y, dtype)
    rt = 1e-6 if dtype == torch.float32 else 1e-12
    assert np.isclose(float(api.dot(tx, ty)), x @ y, rtol=rt, atol=rt * np.sqrt(max(n, 1)))
    assert np.isclose(float(api.nrm2(tx)), np.linalg.norm(x), rtol=rt)
    assert np.isclose(float(api.asum(tx)), np.abs(x).sum(), rtol=rt)
    if n:
        assert float(api.amax(tx)) == np.abs(x).max()
        assert float(api.amin(tx)) == np.abs(x).min()


def test_nrm2_extreme_range():
    for scale in (1e-200, 1e200):
        x = np.full(1000, 3.0 * scale)
        assert np.isclose(float(api.nrm2(dev(x))), np.sqrt(1000) * 3.0 * scale, rtol=1e-13)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_scal_axpy_golden(dtype):
    a = dev(np.array(G.T234A, dtype=np.float64), dtype)
    b = dev(np.array(G.T234B, dtype=np.float64), dtype)
    api.axpy(1.0, b, a)
    assert a.cpu().double().numpy().tolist() == G.T234APB
    a = dev(np.array(G.T234A, dtype=np.float64), dtype)
    api.axpy(-1.0, b, a)
    assert a.cpu().double().numpy().tolist() == G.T234AMB
    a = dev(np.array(G.T234A, dtype=np.float64), dtype)
    api.scal(3.0, a)
    assert a.cpu().double().numpy().tolist() == [3 * v for v in G.T234A]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_transpose_golden_and_random(dtype):
    A = dev(np.array(G.TR_A, dtype=np.float64).reshape(2, 2, 3), dtype)  # (k=2, n=2, m=3)
    assert api.transpose_batched(A).cpu().double().numpy().ravel().tolist() == G.TR_AT
    r = np.random.default_rng(0)
    X = r.standard_normal((7, 96, 48))
    T = api.transpose_batched(dev(X, dtype)).cpu().double().numpy()
    ref = np.swapaxes(X.astype(np.float32) if dtype == torch.float32 else X, 1, 2)
    assert np.array_equal(T, ref)
    # larger than the whole-matrix-in-LDS form (tiled kernel), and odd sizes
    for shape in ((3, 300, 200), (5, 7, 513), (4, 33, 17), (2, 1, 1), (3, 1, 40)):
        X = r.standard_normal(shape)
        T = api.transpose_batched(dev(X, dtype)).cpu().double().numpy()
        ref = np.swapaxes(X.astype(np.float32) if dtype == torch.float32 else X, 1, 2)
        assert np.array_equal(T, ref), shape
