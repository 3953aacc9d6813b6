"""Seeded synthetic inputs of the five benchmark workloads.

Random numbers come from a counter-based SplitMix64 stream (Steele, Lea &
Flood 2014) with explicit Box-Muller normals, so every platform produces the
same bits for the same seed (numpy's and C++'s distribution objects do not
promise that).  All batches are returned in GPUtils tensor layout: a tensor
(m, n, k) is the C-contiguous numpy array of shape (k, n, m), i.e. each
matrix column-major.

Workloads
  cholesky_small   k=1024 SPD 32x32, A = M M^T + 32 I, b (32x1)        [FP64]
  gemm             k=16384, 64x64 times 64x64, entries U[-1,1]         [FP32]
  cholesky_mixed   k=8192 SPD 128x128, A = Q diag(lambda) Q^T,
                   lambda_j = 10^(4j/127), 4 right-hand sides            [FP64]
  ls_ragged        4096 tall systems, n ~ U{8..64}, m ~ U{max(16,n)..512} [FP64]
  svd_nullspace    k=2048 48x96 of exact rank 40, x (96x1)              [FP32]
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """count outputs of SplitMix64 started at state seed*2^20 + stream."""
    with np.errstate(over="ignore"):
        base = np.uint64((int(seed) * (1 << 20) + int(stream)) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, count + 1, dtype=np.uint64) * _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, stream: int = 0, low=0.0, high=1.0) -> np.ndarray:
    """U[low, high) doubles from the top 53 bits."""
    u = (splitmix64(seed, count, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return low + (high - low) * u


def normal(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """N(0,1) doubles by Box-Muller on (0,1] uniforms."""
    half = (count + 1) // 2
    z = splitmix64(seed, 2 * half, stream)
    u = ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * half)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:count]


def randint(seed: int, count: int, low: int, high: int, stream: int = 0) -> np.ndarray:
    """Integers in [low, high] (inclusive) by multiply-shift on 32 high bits."""
    z = splitmix64(seed, count, stream) >> np.uint64(32)
    span = np.uint64(high - low + 1)
    return (low + ((z * span) >> np.uint64(32)).astype(np.int64)).astype(np.int64)


def _orthonormal(seed: int, k: int, m: int, n: int, stream: int) -> np.ndarray:
    """k matrices m x n with orthonormal columns: Householder QR of N(0,1),
    signs fixed so that diag(R) > 0 (the Q is then unique)."""
    G = normal(seed, k * m * n, stream).reshape(k, m, n)
    Q, R = np.linalg.qr(G)
    s = np.sign(np.diagonal(R, axis1=1, axis2=2))
    s[s == 0] = 1.0
    return Q * s[:, None, :]


def _to_tensor(mats: np.ndarray, dtype) -> np.ndarray:
    """(k, m, n) matrices -> (k, n, m) tensor layout (column-major per matrix)."""
    return np.ascontiguousarray(np.swapaxes(mats, 1, 2), dtype=dtype)


def cholesky_small(seed: int = 1, k: int = 1024, n: int = 32):
    M = normal(seed, k * n * n, 0).reshape(k, n, n)
    A = M @ np.swapaxes(M, 1, 2) + n * np.eye(n)
    b = normal(seed, k * n, 1).reshape(k, 1, n)
    return _to_tensor(A, np.float64), np.ascontiguousarray(b)


def gemm(seed: int = 2, k: int = 16384, n: int = 64):
    A = uniform(seed, k * n * n, 0, -1.0, 1.0).astype(np.float32).reshape(k, n, n)
    B = uniform(seed, k * n * n, 1, -1.0, 1.0).astype(np.float32).reshape(k, n, n)
    return A, B  # already in tensor layout (i.i.d. entries)


def spd_spectrum(n: int = 128) -> np.ndarray:
    return 10.0 ** (4.0 * np.arange(n) / (n - 1))


def cholesky_mixed(seed: int = 3, k: int = 8192, n: int = 128, nrhs: int = 4, chunk: int = 1024):
    lam = spd_spectrum(n)
    A = np.empty((k, n, n))
    for s in range(0, k, chunk):
        c = min(chunk, k - s)
        Q = _orthonormal(seed + 7919 * (s // chunk + 1), c, n, n, 0)
        Ac = (Q * lam[None, None, :]) @ np.swapaxes(Q, 1, 2)
        A[s:s + c] = 0.5 * (Ac + np.swapaxes(Ac, 1, 2))
    B = normal(seed, k * nrhs * n, 1).reshape(k, nrhs, n)
    return A, B  # symmetric: tensor layout equals matrix layout


def ls_ragged(seed: int = 4, count: int = 4096, nmin: int = 8, nmax: int = 64, mmax: int = 512):
    """Returns (ms, ns, offA, offB, A_flat, B_flat): system s is the m x n
    column-major matrix at A_flat[offA[s]:] and the m-vector at B_flat[offB[s]:]."""
    ns = randint(seed, count, nmin, nmax, 0)
    lo = np.maximum(16, ns)
    u = uniform(seed, count, 1)
    ms = (lo + np.floor(u * (mmax - lo + 1))).astype(np.int64)
    ms = np.minimum(ms, mmax)
    sizes = ms * ns
    offA = np.zeros(count, dtype=np.int64)
    offA[1:] = np.cumsum(sizes)[:-1]
    offB = np.zeros(count, dtype=np.int64)
    offB[1:] = np.cumsum(ms)[:-1]
    A = normal(seed, int(sizes.sum()), 2)
    B = normal(seed, int(ms.sum()), 3)
    return ms.astype(np.int32), ns.astype(np.int32), offA, offB, A, B


def svd_nullspace(seed: int = 5, k: int = 2048, m: int = 48, n: int = 96, r: int = 40):
    U = _orthonormal(seed, k, m, r, 0)
    V = _orthonormal(seed, k, n, r, 1)
    sig = uniform(seed, k * r, 2, 0.1, 10.0).reshape(k, r)
    A = (U * sig[:, None, :]) @ np.swapaxes(V, 1, 2)
    x = normal(seed, k * n, 3).reshape(k, 1, n)
    return _to_tensor(A, np.float32), x.astype(np.float32)
