// chol.hip -- batched Cholesky factorisation / solve for gfx950.
//
// Replaces cusolverDn?potrfBatched / ?potrsBatched of GPUtils'
// CholeskyBatchFactoriser (include/tensor.cuh:2098-2197) and adds a fused
// factor+solve and a mixed-precision (FP32 factor, FP64 refinement) solve.
//
// Design (one 256-thread workgroup = 4 wavefronts per matrix):
//  * The lower block triangle of the matrix lives in LDS as 16x16 blocks
//    (block (I,J), I >= J, at index I(I+1)/2 + J, column-major inside the
//    block).  n is padded to a multiple of 16 with an identity tail, so the
//    padded factor is diag(L, I) and no kernel path needs edge masks.
//  * Right-looking blocked algorithm, per block column J:
//      1. wave 0 factors the 16x16 diagonal block in registers (one row per
//         lane, broadcasts through v_readlane) and forms its triangular
//         inverse D_J; D_J replaces L_JJ in LDS (L_JJ goes straight to HBM
//         when the factor is an output);
//      2. panel   L_IJ = A_IJ D_J^T      (16x16x4 MFMA, 4 per block);
//      3. update  A_IK -= L_IJ L_KJ^T    (16x16x4 MFMA, 4 per tile).
//    Both MFMA operands of a 16x16x4 step read LDS at index lane + 64*kk.
//  * Triangular solves use the same block inverses: per block row a 16x16
//    matrix-vector product replaces the serial substitution.
#include "common.hpp"

namespace gputils {
namespace chol {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kRhsChunk = 4;  // right-hand sides per solve sweep (lane = row + 16*rhs)

__host__ __device__ constexpr int tri(int I) { return (I * (I + 1)) >> 1; }

// Load the lower triangle of the n x n matrix A (column-major, lda) into
// LDS block storage with identity padding up to 16*nb.  Columns are read
// whole (contiguous rows) by one wavefront each.
template <typename Tin, typename T>
__device__ void load_lower(const Tin *__restrict__ A, int64_t lda, int n, int nb, T *L) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int npad = nb * 16;
    for (int gk = wave; gk < npad; gk += kWaves) {
        const int J = gk >> 4, k = gk & 15;
        for (int gi = (gk & ~15) + lane; gi < npad; gi += kWave) {
            T v;
            if (gi < n && gk < n)
                v = (gi >= gk) ? (T)A[gi + (int64_t)gk * lda] : T(0);
            else
                v = (gi == gk) ? T(1) : T(0);
            const int I = gi >> 4;
            L[(tri(I) + J) * 256 + (gi & 15) + 16 * k] = v;
        }
    }
}

// Store the strictly-below-diagonal blocks (I > J) back to A.
template <typename T>
__device__ void store_offdiag(T *__restrict__ A, int64_t lda, int n, int nb, const T *L) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int gk = wave; gk < n; gk += kWaves) {
        const int J = gk >> 4, k = gk & 15;
        for (int gi = 16 * (J + 1) + lane; gi < n; gi += kWave) {
            const int I = gi >> 4;
            A[gi + (int64_t)gk * lda] = L[(tri(I) + J) * 256 + (gi & 15) + 16 * k];
        }
    }
}

// Factor the 16x16 diagonal block `blk` (lower part valid) with wave 0,
// write D = L^{-1} over it.  If Lout != nullptr the factor L itself is
// written there (lower part, rows < nvalid, column-major with ldo).
// Returns the index (0..15) of the first non-positive pivot or -1.
template <typename T>
__device__ int factor_diag_block(T *blk, T *Lout, int64_t ldo, int nvalid) {
    const int lane = threadIdx.x & 63;
    T row[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) row[c] = (lane < 16) ? blk[lane + 16 * c] : T(0);
    int fail = -1;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const T d = readlane(row[k], k);
        if (!(d > T(0))) {  // uniform: also catches NaN
            fail = k;
            break;
        }
        const T s = dev_sqrt(d);
        const T inv = T(1) / s;
        const T lk = (lane > k) ? row[k] * inv : ((lane == k) ? s : T(0));
        row[k] = lk;
#pragma unroll
        for (int j = k + 1; j < 16; ++j) {
            const T ljk = readlane(lk, j);
            if (lane >= j) row[j] -= lk * ljk;
        }
    }
    if (fail >= 0) return fail;
    if (Lout != nullptr && lane < nvalid) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c <= lane && c < nvalid) Lout[lane + (int64_t)c * ldo] = row[c];
    }
    // D = L^{-1}: lane c builds column c by forward substitution; the
    // L entries are uniform (v_readlane of row r).
    T x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        T acc = (lane == r) ? T(1) : T(0);
#pragma unroll
        for (int p = 0; p < r; ++p) acc -= readlane(row[p], r) * x[p];
        x[r] = acc / readlane(row[r], r);
    }
    if (lane < 16) {
#pragma unroll
        for (int r = 0; r < 16; ++r) blk[r + 16 * lane] = x[r];
    }
    return -1;
}

// Blocked right-looking factorisation of the LDS block triangle.
// On return the diagonal slots hold D_J = L_JJ^{-1}; off-diagonal slots
// hold L_IJ.  When Lout != nullptr each L_JJ is written to Lout as it is
// formed.  Returns 0 or the 1-based failing row (LAPACK info).
template <typename T>
__device__ int factor_blocked(T *L, int nb, int n, T *Lout, int64_t ldo, int *s_flag) {
    using M = Mfma16<T>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int J = 0; J < nb; ++J) {
        T *dJ = L + (tri(J) + J) * 256;
        if (wave == 0) {
            T *lo = (Lout != nullptr) ? Lout + 16 * J + (int64_t)16 * J * ldo : nullptr;
            const int f = factor_diag_block(dJ, lo, ldo, n - 16 * J);
            if (lane == 0) *s_flag = (f >= 0) ? 16 * J + f + 1 : 0;
        }
        __syncthreads();
        const int fail = *s_flag;
        if (fail) return fail;
        // panel: L_IJ = A_IJ * D_J^T
        for (int I = J + 1 + wave; I < nb; I += kWaves) {
            T *t = L + (tri(I) + J) * 256;
            typename M::acc_t acc = M::zero();
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) acc = M::fma(t[lane + 64 * kk], dJ[lane + 64 * kk], acc);
            const int base = 4 * (lane >> 4) + 16 * (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) t[base + r] = acc[r];
        }
        __syncthreads();
        // trailing update of tiles (I,K), J < K <= I
        int t = 0;
        for (int K = J + 1; K < nb; ++K) {
            const T *lk = L + (tri(K) + J) * 256;
            for (int I = K; I < nb; ++I, ++t) {
                if ((t & (kWaves - 1)) != wave) continue;
                const T *li = L + (tri(I) + J) * 256;
                T *c = L + (tri(I) + K) * 256;
                const int base = 4 * (lane >> 4) + 16 * (lane & 15);
                typename M::acc_t acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] = c[base + r];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) acc = M::fma(-li[lane + 64 * kk], lk[lane + 64 * kk], acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) c[base + r] = acc[r];
            }
        }
        __syncthreads();
    }
    return 0;
}

// In LDS, replace each diagonal block L_JJ by its inverse (wave J % 4).
template <typename T>
__device__ void invert_diag_blocks(T *L, int nb) {
    const int wave = threadIdx.x >> 6;
    for (int J = wave; J < nb; J += kWaves) {
        T *dJ = L + (tri(J) + J) * 256;
        const int lane = threadIdx.x & 63;
        T row[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) row[c] = (lane < 16) ? dJ[lane + 16 * c] : T(0);
        T x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            T acc = (lane == r) ? T(1) : T(0);
#pragma unroll
            for (int p = 0; p < r; ++p) acc -= readlane(row[p], r) * x[p];
            x[r] = acc / readlane(row[r], r);
        }
        if (lane < 16) {
#pragma unroll
            for (int r = 0; r < 16; ++r) dJ[r + 16 * lane] = x[r];
        }
    }
}

// Solve L L^T Y = Y in place for the npad x 4 block Y (LDS, column-major,
// ld = npad) using the block triangle (diagonal slots = inverses).
// Lane = row + 16 * rhs inside a 16-row block.
template <typename T>
__device__ void solve_blocked(const T *L, int nb, T *Y) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, r = lane >> 4;
    const int ld = 16 * nb;
    // forward: L y = b
    for (int I = 0; I < nb; ++I) {
        if (wave == (I & (kWaves - 1))) {
            const T *d = L + (tri(I) + I) * 256;
            T *y = Y + 16 * I + ld * r;
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < 16; ++k) acc += d[i + 16 * k] * y[k];
            y[i] = acc;
        }
        __syncthreads();
        for (int K = I + 1 + wave; K < nb; K += kWaves) {
            const T *l = L + (tri(K) + I) * 256;
            const T *y = Y + 16 * I + ld * r;
            T *x = Y + 16 * K + ld * r;
            T acc = x[i];
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= l[i + 16 * k] * y[k];
            x[i] = acc;
        }
        __syncthreads();
    }
    // backward: L^T x = y
    for (int I = nb - 1; I >= 0; --I) {
        if (wave == (I & (kWaves - 1))) {
            const T *d = L + (tri(I) + I) * 256;
            T *y = Y + 16 * I + ld * r;
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < 16; ++k) acc += d[k + 16 * i] * y[k];
            y[i] = acc;
        }
        __syncthreads();
        for (int K = wave; K < I; K += kWaves) {
            const T *l = L + (tri(I) + K) * 256;
            const T *z = Y + 16 * I + ld * r;
            T *x = Y + 16 * K + ld * r;
            T acc = x[i];
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= l[k + 16 * i] * z[k];
            x[i] = acc;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ kernels --
// Factor only (API: CholeskyBatchFactoriser::factorise).
template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void potrf_kernel(int n, T *A, int64_t lda, int64_t strideA,
                                                         int *info, int64_t batch) {
    __shared__ T L[tri(NB) * 256];
    __shared__ int s_flag;
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int nb = (n + 15) >> 4;
    T *a = A + b * strideA;
    load_lower(a, lda, n, nb, L);
    __syncthreads();
    const int fail = factor_blocked(L, nb, n, a, lda, &s_flag);
    if (fail == 0) store_offdiag(a, lda, n, nb, L);
    if (threadIdx.x == 0 && info != nullptr) info[b] = fail;
}

// Solve only (API: CholeskyBatchFactoriser::solve), any nrhs.
template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void potrs_kernel(int n, int nrhs, const T *Lg, int64_t lda,
                                                         int64_t strideL, T *B, int64_t ldb,
                                                         int64_t strideB, int *info, int64_t batch) {
    __shared__ T L[tri(NB) * 256];
    __shared__ T Y[16 * NB * kRhsChunk];
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int nb = (n + 15) >> 4;
    const int npad = 16 * nb;
    load_lower(Lg + b * strideL, lda, n, nb, L);
    __syncthreads();
    invert_diag_blocks(L, nb);
    T *bb = B + b * strideB;
    for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
        __syncthreads();
        for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) {
            const int row = e % npad, c = e / npad;
            Y[e] = (row < n && r0 + c < nrhs) ? bb[row + (int64_t)(r0 + c) * ldb] : T(0);
        }
        __syncthreads();
        solve_blocked(L, nb, Y);
        for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) {
            const int row = e % npad, c = e / npad;
            if (row < n && r0 + c < nrhs) bb[row + (int64_t)(r0 + c) * ldb] = Y[e];
        }
    }
    if (threadIdx.x == 0 && info != nullptr) info[b] = 0;
}

// Fused factor + solve.  WRITE_L: A <- L (API shape).  X may alias B.
// mask != nullptr: only matrices with mask[b] < 0 are processed (the FP64
// fallback of the mixed solver); then A is read only.
template <typename T, int NB, bool WRITE_L>
__global__ __launch_bounds__(kThreads) void posv_kernel(int n, int nrhs, T *A, int64_t lda,
                                                        int64_t strideA, const T *B, int64_t ldb,
                                                        int64_t strideB, T *X, int64_t ldx,
                                                        int64_t strideX, int *info, int64_t batch,
                                                        const int *mask) {
    __shared__ T L[tri(NB) * 256];
    __shared__ T Y[16 * NB * kRhsChunk];
    __shared__ int s_flag;
    for (int64_t b = blockIdx.x; b < batch; b += gridDim.x) {
        if (mask != nullptr && mask[b] >= 0) continue;
        const int nb = (n + 15) >> 4;
        const int npad = 16 * nb;
        T *a = A + b * strideA;
        __syncthreads();
        load_lower(a, lda, n, nb, L);
        __syncthreads();
        const int fail = factor_blocked(L, nb, n, WRITE_L ? a : (T *)nullptr, lda, &s_flag);
        if (threadIdx.x == 0 && info != nullptr) info[b] = fail;
        if (fail) continue;  // uniform
        if (WRITE_L) store_offdiag(a, lda, n, nb, L);
        const T *bb = B + b * strideB;
        T *xx = X + b * strideX;
        for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
            __syncthreads();
            for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) {
                const int row = e % npad, c = e / npad;
                Y[e] = (row < n && r0 + c < nrhs) ? bb[row + (int64_t)(r0 + c) * ldb] : T(0);
            }
            __syncthreads();
            solve_blocked(L, nb, Y);
            for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) {
                const int row = e % npad, c = e / npad;
                if (row < n && r0 + c < nrhs) xx[row + (int64_t)(r0 + c) * ldx] = Y[e];
            }
        }
    }
}

// ----------------------------------------------- large-n global kernels --
// Correctness path for n > 128: unblocked right-looking algorithm on the
// matrix in global memory, one workgroup per matrix.
template <typename T>
__global__ __launch_bounds__(kThreads) void potrf_global_kernel(int n, T *A, int64_t lda,
                                                                int64_t strideA, int *info,
                                                                int64_t batch) {
    __shared__ int s_fail;
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    T *a = A + b * strideA;
    if (threadIdx.x == 0) s_fail = 0;
    for (int k = 0; k < n; ++k) {
        __syncthreads();
        const T d = a[k + (int64_t)k * lda];
        if (!(d > T(0))) {
            if (threadIdx.x == 0) s_fail = k + 1;
            break;
        }
        const T s = dev_sqrt(d);
        __syncthreads();
        for (int i = k + 1 + threadIdx.x; i < n; i += kThreads) a[i + (int64_t)k * lda] /= s;
        if (threadIdx.x == 0) a[k + (int64_t)k * lda] = s;
        __syncthreads();
        const int m = n - k - 1;
        for (int64_t e = threadIdx.x; e < (int64_t)m * m; e += kThreads) {
            const int i = k + 1 + (int)(e % m), j = k + 1 + (int)(e / m);
            if (i >= j) a[i + (int64_t)j * lda] -= a[i + (int64_t)k * lda] * a[j + (int64_t)k * lda];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && info != nullptr) info[b] = s_fail;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void potrs_global_kernel(int n, int nrhs, const T *L,
                                                                int64_t lda, int64_t strideL, T *B,
                                                                int64_t ldb, int64_t strideB,
                                                                int *info, int64_t batch) {
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const T *l = L + b * strideL;
    for (int r = 0; r < nrhs; ++r) {
        T *x = B + b * strideB + (int64_t)r * ldb;
        for (int k = 0; k < n; ++k) {
            __syncthreads();
            const T xk = x[k] / l[k + (int64_t)k * lda];
            __syncthreads();
            if (threadIdx.x == 0) x[k] = xk;
            for (int i = k + 1 + threadIdx.x; i < n; i += kThreads) x[i] -= l[i + (int64_t)k * lda] * xk;
        }
        for (int k = n - 1; k >= 0; --k) {
            __syncthreads();
            // x_k = (x_k - sum_{i>k} l_ik x_i) / l_kk : block reduction
            __shared__ T red[kThreads];
            T part = T(0);
            for (int i = k + 1 + threadIdx.x; i < n; i += kThreads) part += l[i + (int64_t)k * lda] * x[i];
            red[threadIdx.x] = part;
            __syncthreads();
            for (int s = kThreads / 2; s > 0; s >>= 1) {
                if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) x[k] = (x[k] - red[0]) / l[k + (int64_t)k * lda];
        }
    }
    if (threadIdx.x == 0 && info != nullptr) info[b] = 0;
}

// --------------------------------------------------- mixed precision --
// FP32 factor in LDS; FP64 iterate x (LDS); FP64 residual straight from
// the FP64 matrix in HBM/L2 (symmetric: element (p,q) = A[max + min*lda]).
template <int NB>
__global__ __launch_bounds__(kThreads) void posv_mixed_kernel(
    int n, int nrhs, const double *__restrict__ A, int64_t lda, int64_t strideA,
    const double *__restrict__ B, int64_t ldb, int64_t strideB, double *__restrict__ X,
    int64_t ldx, int64_t strideX, double tol, int max_iter, int *iters, int64_t batch) {
    __shared__ float L[tri(NB) * 256];
    __shared__ float R[16 * NB * kRhsChunk];   // FP32 right-hand side / correction
    __shared__ double XD[16 * NB * kRhsChunk];  // FP64 iterate
    __shared__ double s_red[kWaves][3][kRhsChunk];
    __shared__ double s_amax[kWaves];
    __shared__ int s_flag;
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb = (n + 15) >> 4;
    const int npad = 16 * nb;
    const double *a = A + b * strideA;
    const double *bb = B + b * strideB;
    double *xx = X + b * strideX;

    load_lower(a, lda, n, nb, L);
    __syncthreads();
    const int fail = factor_blocked(L, nb, n, (float *)nullptr, 0, &s_flag);
    if (fail) {
        if (threadIdx.x == 0) iters[b] = -1;
        return;
    }
    int used = 0;
    for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
        const int nr = (nrhs - r0 < kRhsChunk) ? (nrhs - r0) : kRhsChunk;
        __syncthreads();
        for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) {
            const int row = e % npad, c = e / npad;
            R[e] = (row < n && c < nr) ? (float)bb[row + (int64_t)(r0 + c) * ldb] : 0.f;
            XD[e] = 0.0;
        }
        int it = 0;
        bool converged = false;
        for (;;) {
            // correction: solve L L^T d = R (FP32), x += d (FP64)
            __syncthreads();
            solve_blocked(L, nb, R);
            for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) XD[e] += (double)R[e];
            __syncthreads();
            // FP64 residual r = b - A x, norms per right-hand side
            const int i = lane & 15, rr = lane >> 4;
            double rmax = 0.0, xmax = 0.0, bmax = 0.0, amax = 0.0;
            for (int I = wave; I < nb; I += kWaves) {
                const int p = 16 * I + i;
                double acc = 0.0, arow = 0.0;
                if (p < n) {
                    acc = (rr < nr) ? bb[p + (int64_t)(r0 + rr) * ldb] : 0.0;
                    bmax = fmax(bmax, fabs(acc));
                    for (int q = 0; q < n; ++q) {
                        const int hi = p > q ? p : q, lo = p > q ? q : p;
                        const double aq = a[hi + (int64_t)lo * lda];
                        acc -= aq * XD[q + npad * rr];
                        arow += fabs(aq);
                    }
                    xmax = fmax(xmax, fabs(XD[p + npad * rr]));
                }
                R[p + npad * rr] = (float)acc;
                rmax = fmax(rmax, fabs(acc));
                amax = fmax(amax, arow);
            }
            // reduce over the 16 rows of a lane group, then over waves
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) {
                rmax = fmax(rmax, __shfl_xor(rmax, off, kWave));
                xmax = fmax(xmax, __shfl_xor(xmax, off, kWave));
                bmax = fmax(bmax, __shfl_xor(bmax, off, kWave));
                amax = fmax(amax, __shfl_xor(amax, off, kWave));
            }
            amax = fmax(amax, __shfl_xor(amax, 16, kWave));
            amax = fmax(amax, __shfl_xor(amax, 32, kWave));
            if (i == 0) {
                s_red[wave][0][rr] = rmax;
                s_red[wave][1][rr] = xmax;
                s_red[wave][2][rr] = bmax;
            }
            if (lane == 0) s_amax[wave] = amax;
            __syncthreads();
            // every thread evaluates the same test (uniform branch)
            double anorm = 0.0;
            for (int w = 0; w < kWaves; ++w) anorm = fmax(anorm, s_amax[w]);
            bool ok = true;
            for (int c = 0; c < nr; ++c) {
                double rn = 0.0, xn = 0.0, bn = 0.0;
                for (int w = 0; w < kWaves; ++w) {
                    rn = fmax(rn, s_red[w][0][c]);
                    xn = fmax(xn, s_red[w][1][c]);
                    bn = fmax(bn, s_red[w][2][c]);
                }
                ok = ok && (rn <= tol * (anorm * xn + bn));
            }
            if (ok) {
                converged = true;
                break;
            }
            if (it >= max_iter) break;
            ++it;
        }
        used = (it > used) ? it : used;
        if (!converged) {
            used = -1;
            break;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < npad * kRhsChunk; e += kThreads) {
            const int row = e % npad, c = e / npad;
            if (row < n && c < nr) xx[row + (int64_t)(r0 + c) * ldx] = XD[e];
        }
    }
    if (threadIdx.x == 0) iters[b] = used;
}

// ----------------------------------------------------------- dispatch --
template <typename T>
int potrf_dispatch(int64_t n, T *A, int64_t lda, int64_t strideA, int *info, int64_t batch,
                   hipStream_t st) {
    const dim3 grid((unsigned)batch), block(kThreads);
    if (n <= 16) hipLaunchKernelGGL((potrf_kernel<T, 1>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else if (n <= 32) hipLaunchKernelGGL((potrf_kernel<T, 2>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else if (n <= 64) hipLaunchKernelGGL((potrf_kernel<T, 4>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else if (n <= 128) hipLaunchKernelGGL((potrf_kernel<T, 8>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else hipLaunchKernelGGL((potrf_global_kernel<T>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    return launch_check("potrf_kernel");
}

template <typename T>
int potrs_dispatch(int64_t n, int64_t nrhs, const T *L, int64_t lda, int64_t strideL, T *B,
                   int64_t ldb, int64_t strideB, int *info, int64_t batch, hipStream_t st) {
    const dim3 grid((unsigned)batch), block(kThreads);
#define GPUTILS_POTRS(NBV) \
    hipLaunchKernelGGL((potrs_kernel<T, NBV>), grid, block, 0, st, (int)n, (int)nrhs, L, lda, strideL, B, ldb, strideB, info, batch)
    if (n <= 16) GPUTILS_POTRS(1);
    else if (n <= 32) GPUTILS_POTRS(2);
    else if (n <= 64) GPUTILS_POTRS(4);
    else if (n <= 128) GPUTILS_POTRS(8);
    else hipLaunchKernelGGL((potrs_global_kernel<T>), grid, block, 0, st, (int)n, (int)nrhs, L, lda, strideL, B, ldb, strideB, info, batch);
#undef GPUTILS_POTRS
    return launch_check("potrs_kernel");
}

template <typename T, bool WRITE_L>
int posv_dispatch(int64_t n, int64_t nrhs, T *A, int64_t lda, int64_t strideA, const T *B,
                  int64_t ldb, int64_t strideB, T *X, int64_t ldx, int64_t strideX, int *info,
                  int64_t batch, const int *mask, hipStream_t st) {
    const int64_t g = (mask != nullptr) ? (batch < 2048 ? batch : 2048) : batch;
    const dim3 grid((unsigned)g), block(kThreads);
#define GPUTILS_POSV(NBV)                                                                     \
    hipLaunchKernelGGL((posv_kernel<T, NBV, WRITE_L>), grid, block, 0, st, (int)n, (int)nrhs, A, \
                       lda, strideA, B, ldb, strideB, X, ldx, strideX, info, batch, mask)
    if (n <= 16) GPUTILS_POSV(1);
    else if (n <= 32) GPUTILS_POSV(2);
    else if (n <= 64) GPUTILS_POSV(4);
    else GPUTILS_POSV(8);
#undef GPUTILS_POSV
    return launch_check("posv_kernel");
}

}  // namespace chol
}  // namespace gputils

using namespace gputils;

extern "C" {

int gputils_spotrf_batched(int64_t n, float *A, int64_t lda, int64_t strideA, int *info,
                           int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 3, "lda < max(1,n)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 6, "batch out of range");
    if (n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 2, "A is NULL");
    return chol::potrf_dispatch<float>(n, A, lda, strideA, info, batch, as_stream(stream));
}

int gputils_dpotrf_batched(int64_t n, double *A, int64_t lda, int64_t strideA, int *info,
                           int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 3, "lda < max(1,n)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 6, "batch out of range");
    if (n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 2, "A is NULL");
    return chol::potrf_dispatch<double>(n, A, lda, strideA, info, batch, as_stream(stream));
}

int gputils_spotrs_batched(int64_t n, int64_t nrhs, const float *L, int64_t lda, int64_t strideL,
                           float *B, int64_t ldb, int64_t strideB, int *info, int64_t batch,
                           void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 4, "lda < max(1,n)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 10, "batch out of range");
    if (n == 0 || nrhs == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(L != nullptr, 3, "L is NULL");
    GPUTILS_ARG(B != nullptr, 6, "B is NULL");
    return chol::potrs_dispatch<float>(n, nrhs, L, lda, strideL, B, ldb, strideB, info, batch,
                                       as_stream(stream));
}

int gputils_dpotrs_batched(int64_t n, int64_t nrhs, const double *L, int64_t lda, int64_t strideL,
                           double *B, int64_t ldb, int64_t strideB, int *info, int64_t batch,
                           void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 4, "lda < max(1,n)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 10, "batch out of range");
    if (n == 0 || nrhs == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(L != nullptr, 3, "L is NULL");
    GPUTILS_ARG(B != nullptr, 6, "B is NULL");
    return chol::potrs_dispatch<double>(n, nrhs, L, lda, strideL, B, ldb, strideB, info, batch,
                                        as_stream(stream));
}

int gputils_dposv_batched(int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t strideA,
                          double *B, int64_t ldb, int64_t strideB, int *info, int64_t batch,
                          void *stream) {
    GPUTILS_ARG(n >= 0 && n <= 128, 1, "n must be in [0, 128]");
    GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 4, "lda < max(1,n)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 10, "batch out of range");
    if (n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(nrhs == 0 || B != nullptr, 6, "B is NULL");
    return chol::posv_dispatch<double, true>(n, nrhs, A, lda, strideA, B, ldb, strideB, B, ldb,
                                             strideB, info, batch, nullptr, as_stream(stream));
}

int gputils_dposv_mixed_batched(int64_t n, int64_t nrhs, const double *A, int64_t lda,
                                int64_t strideA, const double *B, int64_t ldb, int64_t strideB,
                                double *X, int64_t ldx, int64_t strideX, double tol, int max_iter,
                                int *iters, int *info, int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 0 && n <= 128, 1, "n must be in [0, 128]");
    GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 4, "lda < max(1,n)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(ldx >= (n > 1 ? n : 1), 10, "ldx < max(1,n)");
    GPUTILS_ARG(tol > 0.0, 12, "tol <= 0");
    GPUTILS_ARG(max_iter >= 0, 13, "max_iter < 0");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 16, "batch out of range");
    if (n == 0 || nrhs == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(B != nullptr, 6, "B is NULL");
    GPUTILS_ARG(X != nullptr, 9, "X is NULL");
    GPUTILS_ARG(iters != nullptr, 14, "iters is NULL");
    GPUTILS_ARG(info != nullptr, 15, "info is NULL");
    hipStream_t st = as_stream(stream);
    // info defaults to 0 for matrices that the mixed kernel solves
    {
        int rc = gputils_fill_bytes_async(info, 0, sizeof(int) * (size_t)batch, stream);
        if (rc != GPUTILS_OK) return rc;
    }
    const dim3 grid((unsigned)batch), block(chol::kThreads);
#define GPUTILS_MIXED(NBV)                                                                      \
    hipLaunchKernelGGL((chol::posv_mixed_kernel<NBV>), grid, block, 0, st, (int)n, (int)nrhs, A, \
                       lda, strideA, B, ldb, strideB, X, ldx, strideX, tol, max_iter, iters, batch)
    if (n <= 16) GPUTILS_MIXED(1);
    else if (n <= 32) GPUTILS_MIXED(2);
    else if (n <= 64) GPUTILS_MIXED(4);
    else GPUTILS_MIXED(8);
#undef GPUTILS_MIXED
    int rc = launch_check("posv_mixed_kernel");
    if (rc != GPUTILS_OK) return rc;
    // FP64 fallback for matrices flagged with iters < 0 (A, B read only)
    return chol::posv_dispatch<double, false>(n, nrhs, const_cast<double *>(A), lda, strideA, B, ldb,
                                              strideB, X, ldx, strideX, info, batch, iters, st);
}

}  // extern "C"
