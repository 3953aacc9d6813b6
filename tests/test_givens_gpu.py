"""Givens rotations and GivensAnnihilator on the GPU (include/tensor.cuh:
1074-1104, 2204-2312) against the oracle and the reference suite's
expectations (X:580-641, X:1684-1729)."""
import numpy as np
import pytest
import torch

from gputils_amd import api

pytestmark = pytest.mark.gpu


def _dev(M, dtype):
    """m x n numpy matrix -> (n, m) column-major device storage."""
    return torch.from_numpy(np.ascontiguousarray(M.T)).to(dtype).cuda()


def _host(t):
    return t.double().cpu().numpy().T


def _iota(m=10, n=6):
    return np.arange(1, m * n + 1, dtype=np.float64).reshape(n, m).T


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-4), (torch.float64, 1e-10)])
def test_right_givens_reference_case(dtype, tol):
    A = _iota()
    c = 0.1
    ms = float(np.sqrt(1 - c * c))
    t = _dev(A, dtype)
    api.givens_rotate_cols(t, 1, 4, c, ms)
    R = _host(t)
    i = np.arange(10)
    for col in (0, 2, 3, 5):
        assert np.allclose(R[:, col], A[:, col], rtol=0, atol=tol)
    assert np.allclose(R[:, 1], (11 + i) * c - (41 + i) * (-ms), rtol=0, atol=tol)
    assert np.allclose(R[:, 4], (11 + i) * (-ms) + (41 + i) * c, rtol=0, atol=tol)


def test_left_givens_reference_case():
    A = _iota()
    c = 0.1
    ms = -float(np.sqrt(1 - c * c))
    t = _dev(A, torch.float64)
    api.givens_rotate_rows(t, 1, 9, c, ms)
    L = _host(t)
    j = np.arange(6)
    assert np.array_equal(L[[0] + list(range(2, 9))], A[[0] + list(range(2, 9))])
    assert np.allclose(L[1], (2 + 10 * j) * c + (10 + 10 * j) * ms, rtol=0, atol=1e-14)
    assert np.allclose(L[9], (2 + 10 * j) * (-ms) + (10 + 10 * j) * c, rtol=0, atol=1e-14)


@pytest.mark.parametrize("m,n", [(7, 5), (300, 4), (5, 1000)])
def test_rotations_match_oracle(oracle, m, n):
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    c, s = 0.6, -0.8
    t = _dev(A, torch.float64)
    api.givens_rotate_cols(t, n - 1, 0, c, s)
    assert np.allclose(_host(t), oracle.rot_cols(A, n - 1, 0, c, s), rtol=0, atol=4e-16 * np.abs(A).max() * 4)
    t = _dev(A, torch.float64)
    api.givens_rotate_rows(t, 0, m - 1, c, s)
    assert np.allclose(_host(t), oracle.rot_rows(A, 0, m - 1, c, s), rtol=0, atol=4e-16 * np.abs(A).max() * 4)


def test_annihilate_golden():
    t = _dev(_iota(), torch.float64)
    api.givens_annihilate(t, 0, 1, 2)
    G = _host(t)
    assert abs(G[1, 2]) <= 1e-14
    assert abs(G[0, 0] - 2.137186834969645) <= 1e-14
    assert abs(G[0, 3] - 44.552125559751822) <= 1e-14
    assert abs(G[1, 3] - (-0.328797974610715)) <= 1e-14


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-4), (torch.float64, 1e-10)])
def test_annihilate_every_element(dtype, tol):
    t = _dev(_iota(), dtype)
    for k in range(1, 10):
        for j in range(6):
            api.givens_annihilate(t, 0, k, j)
            assert abs(float(t[j, k])) <= tol


@pytest.mark.parametrize("m,n", [(9, 7), (64, 300), (513, 3)])
def test_annihilate_matches_oracle(oracle, m, n):
    rng = np.random.default_rng(m + 7 * n)
    A = rng.standard_normal((m, n))
    t = _dev(A, torch.float64)
    want = A
    for (i, k, j) in [(0, m - 1, n - 1), (m // 2, 1, 0), (m - 1, m // 3, n // 2)]:
        api.givens_annihilate(t, i, k, j)
        want = oracle.givens_annihilate(want, i, k, j)
        got = _host(t)
        assert abs(got[k, j]) <= 1e-15 * np.abs(want).max() * 4
        assert np.allclose(got, want, rtol=0, atol=1e-14 * max(1.0, np.abs(want).max()))


def test_annihilate_zero_pivots_follow_the_reference_formula(oracle):
    """A(i,j) = A(k,j) = 0: 1 / hypot is +inf and c = s = 0 * inf = NaN, as
    in the reference's rhypot formula; rows i and k become NaN there and in
    the oracle alike, the other rows stay."""
    A = np.ones((4, 3))
    A[1, 2] = A[3, 2] = 0.0
    t = _dev(A, torch.float64)
    api.givens_annihilate(t, 1, 3, 2)
    got, want = _host(t), oracle.givens_annihilate(A, 1, 3, 2)
    assert np.isnan(got[[1, 3]]).all() and np.isnan(want[[1, 3]]).all()
    assert np.array_equal(got[[0, 2]], A[[0, 2]])


def test_givens_argument_checks():
    t = _dev(np.ones((4, 3)), torch.float64)
    with pytest.raises(api.GputilsError):
        api.givens_annihilate(t, 0, 4, 0)   # row k out of range
    with pytest.raises(api.GputilsError):
        api.givens_annihilate(t, 0, 1, 3)   # column j out of range
    with pytest.raises(api.GputilsError):
        api.givens_rotate_cols(t, 0, 3, 1.0, 0.0)
    with pytest.raises(api.GputilsError):
        api.givens_rotate_rows(t, 4, 0, 1.0, 0.0)
