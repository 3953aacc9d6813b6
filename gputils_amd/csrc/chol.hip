// chol.hip -- batched Cholesky factorisation / solve for gfx950.
//
// Replaces cusolverDn?potrfBatched / ?potrsBatched of GPUtils'
// CholeskyBatchFactoriser (include/tensor.cuh:2098-2197) and adds a fused
// factor+solve.  The mixed-precision solve (FP32 factor, FP64 refinement)
// is in chol_mixed.hip and n <= 32 in chol_small.hip; its C entry point is
// here, with this file's FP64 kernel as the fallback for systems that do
// not converge.
//
// Design (one 256-thread workgroup = 4 wavefronts per matrix):
//  * The lower block triangle of the matrix lives in LDS as 16x16 blocks
//    (block (I,J), I >= J, at index I(I+1)/2 + J, column-major inside the
//    block with the row index XOR-ed by the column, see sw()).  n is padded to a multiple of 16 with an identity tail, so the
//    padded factor is diag(L, I) and no kernel path needs edge masks.
//  * Right-looking blocked algorithm, per block column J:
//      1. wave 0 factors the 16x16 diagonal block in registers (one row per
//         lane, broadcasts through v_readlane) and forms its triangular
//         inverse D_J; D_J replaces L_JJ in LDS (L_JJ goes straight to HBM
//         when the factor is an output);
//      2. panel   L_IJ = A_IJ D_J^T      (16x16x4 MFMA, 4 per block);
//      3. update  A_IK -= L_IJ L_KJ^T    (16x16x4 MFMA, 4 per tile).
//    Both MFMA operands of a 16x16x4 step read element
//    (lane % 16, lane / 16 + 4 kk) of their block.
//  * Triangular solves use the same block inverses: per block row a 16x16
//    matrix-vector product replaces the serial substitution.
#include <mutex>
#include <type_traits>

#include "chol_mixed.hpp"
#include "chol_small.hpp"
#include "common.hpp"

namespace gputils {
namespace chol {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kRhsChunk = 4;  // right-hand sides per solve sweep (lane = row + 16*rhs)

__host__ __device__ constexpr int tri(int I) { return (I * (I + 1)) >> 1; }

// Element (r, c) of a 16x16 block in LDS sits at sw(r, c): column-major with
// the row index XOR-ed by the column.  A column (fixed c) still fills 16
// consecutive slots, and a row (fixed r, c = 0..15), which the MFMA
// accumulator layout and the transposed solves read and write, falls in
// 16 different banks instead of one (a plain column-major 16x16 block of
// doubles puts a row on a 128-byte stride).
__device__ __forceinline__ int sw(int r, int c) { return (r ^ c) + 16 * c; }

// Load the lower triangle of the n x n matrix A (column-major, lda) into
// LDS block storage with identity padding up to 16*nb.  Thread t owns
// element (t % 16, t / 16) of every block; blocks are fetched eight at a
// time so that each thread keeps eight independent loads in flight.
template <typename Tin, typename T>
__device__ __forceinline__ void load_lower(const Tin *__restrict__ A, int64_t lda, int n, int nb, T *L) {
    const int i = threadIdx.x & 15, k = threadIdx.x >> 4;
    const int nblk = tri(nb);
    int I = 0, J = 0;  // coordinates of block b0 (uniform)
    for (int b0 = 0; b0 < nblk; b0 += 8) {
        T v[8];
        int Iu = I, Ju = J;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int gi = 16 * Iu + i, gk = 16 * Ju + k;
            T x = (gi == gk) ? T(1) : T(0);
            if (b0 + u < nblk && gi < n && gk < n && gi >= gk) x = (T)A[gi + (int64_t)gk * lda];
            v[u] = x;
            if (++Ju > Iu) { ++Iu; Ju = 0; }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (b0 + u < nblk) L[(b0 + u) * 256 + sw(i, k)] = v[u];
        I = Iu;
        J = Ju;
    }
}

// Store the strictly-below-diagonal blocks (I > J) back to A.
template <typename T>
__device__ __forceinline__ void store_offdiag(T *__restrict__ A, int64_t lda, int n, int nb, const T *L) {
    const int i = threadIdx.x & 15, k = threadIdx.x >> 4;
    for (int I = 1; I < nb; ++I) {
        const int gi = 16 * I + i;
        for (int J = 0; J < I; ++J) {
            const int gk = 16 * J + k;
            if (gi < n && gk < n) A[gi + (int64_t)gk * lda] = L[(tri(I) + J) * 256 + sw(i, k)];
        }
    }
}

// Pivot of column k: s = sqrt(d) and inv = 1/s, for d > 0.  FP32 uses the
// hardware reciprocal square root (v_rsq_f32, 1 ulp).  FP64 refines the
// hardware estimate v_rsq_f64 instead of running the IEEE square root and
// divide sequences (about 30 dependent operations on the critical path of
// every column): one coupled Goldschmidt step for sqrt(d) and 1/(2 sqrt(d)),
// then one Newton correction of each.  s comes out correctly rounded and
// inv within 1.5 ulp of 1/sqrt(d) (an estimate error up to 2^-22 checked
// against exact arithmetic).  Pivots below 2^-767 are scaled by 2^256
// first, as the library square root does, to keep the estimate accurate.
__device__ __forceinline__ void pivot(float d, float &s, float &inv) {
    inv = __builtin_amdgcn_rsqf(d);
    s = d * inv;
}
__device__ __forceinline__ void pivot(double d, double &s, double &inv) {
    const bool tiny = d < 0x1p-767;
    const double dd = tiny ? d * 0x1p+256 : d;
    const double r = __builtin_amdgcn_rsq(dd);
    double g = dd * r, h = 0.5 * r;
    double e = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    e = __builtin_fma(-g, g, dd);
    g = __builtin_fma(e, h, g);
    double q = h + h;
    e = __builtin_fma(-g, q, 1.0);
    q = __builtin_fma(q, e, q);
    s = tiny ? g * 0x1p-128 : g;
    inv = tiny ? q * 0x1p+128 : q;
}

// In registers of one wavefront (lane r < 16 holds row r of the block),
// invert the lower triangular L (row[]) given the reciprocals of its
// diagonal (uniform): lane c returns column c of L^{-1} in x[].
template <typename T>
__device__ __forceinline__ void tri_inverse_cols(const T (&row)[16], const T (&invd)[16], T (&x)[16]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        T acc = (lane == r) ? T(1) : T(0);
#pragma unroll
        for (int p = 0; p < r; ++p) acc -= readlane(row[p], r) * x[p];
        x[r] = acc * invd[r];
        // x[r] is computed here: left to the compiler, every row's readlane
        // results were hoisted into SGPRs and spilled (about 450 spills)
        asm volatile("" : "+v"(x[r]));
    }
}

// Factor the 16x16 diagonal block `blk` (lower part valid) with one
// wavefront and write D = L^{-1} over it (column-major).  If Lout != nullptr
// the factor L itself is written there (rows/cols < nvalid).
// Returns the index (0..15) of the first non-positive pivot or -1.
//
// Lane r < 16 holds row r.  Step k scales column k, applies its rank-1
// update to the columns right of it and forms row k of L^{-1} (lane c
// holding column c), whose inputs -- row k of L -- are final at that
// point.  The 16 steps have no branch between them (a failed pivot is only
// recorded; the NaNs it spreads are never stored), so the scheduler can
// start the pivot of column k + 1 as soon as its own update is done and
// overlap it with the rest of step k.  Entries above the diagonal are
// updated with garbage and never read: only lanes r > j read row[j].
template <typename T>
__device__ __forceinline__ int factor_diag_block(T *blk, T *Lout, int64_t ldo, int nvalid) {
    // a lane index the compiler cannot relate to the one of the other
    // phases, so the per-step lane masks below are formed where they are
    // used instead of being kept live in SGPRs across the kernel
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    T row[16], x[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) row[c] = blk[sw(lane & 15, c)];
    int fail = -1;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const T d = readlane(row[k], k);
        if (fail < 0 && !(d > T(0))) fail = k;  // uniform; also catches NaN
        T sq, inv;
        pivot(d, sq, inv);
        const T lk = (lane == k) ? sq : row[k] * inv;
        row[k] = lk;
#pragma unroll
        for (int j = k + 1; j < 16; ++j) row[j] -= lk * readlane(lk, j);
        // row k of L^{-1}: x[k] = (e_k - sum_{p<k} L(k, p) x[p]) / L(k, k)
        T acc = (lane == k) ? T(1) : T(0);
#pragma unroll
        for (int p = 0; p < k; ++p) acc -= readlane(row[p], k) * x[p];
        x[k] = acc * inv;
        // left to the compiler, every row's readlane results were hoisted
        // into SGPRs and spilled
        asm volatile("" : "+v"(x[k]));
    }
    if (fail >= 0) return fail;
    if (Lout != nullptr && lane < nvalid) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c <= lane && c < nvalid) Lout[lane + (int64_t)c * ldo] = row[c];
    }
    if (lane < 16) {
#pragma unroll
        for (int r = 0; r < 16; ++r) blk[sw(r, lane)] = x[r];
    }
    return -1;
}

// C -= A_op * B_op^T for one 16x16 tile (all operands 16x16 LDS blocks,
// column-major); two tiles are interleaved for MFMA latency hiding.
template <typename T>
__device__ __forceinline__ void tile_update2(T *c0, const T *a0, const T *b0, T *c1, const T *a1,
                                    const T *b1, bool two) {
    using M = Mfma16<T>;
    const int lane = threadIdx.x & 63;
    const int col = lane & 15;
    typename M::acc_t x0, x1;
#pragma unroll
    for (int r = 0; r < 4; ++r) x0[r] = c0[sw(M::drow(lane, r), col)];
    if (two) {
#pragma unroll
        for (int r = 0; r < 4; ++r) x1[r] = c1[sw(M::drow(lane, r), col)];
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const int o = sw(col, (lane >> 4) + 4 * kk);  // operand element (lane % 16, lane / 16 + 4 kk)
        x0 = M::fma(-a0[o], b0[o], x0);
        if (two) x1 = M::fma(-a1[o], b1[o], x1);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) c0[sw(M::drow(lane, r), col)] = x0[r];
    if (two) {
#pragma unroll
        for (int r = 0; r < 4; ++r) c1[sw(M::drow(lane, r), col)] = x1[r];
    }
}

// Blocked right-looking factorisation of the LDS block triangle with
// lookahead: while waves 1..3 apply the trailing update of block column
// J, wave 0 first updates the next diagonal block and factors it, so the
// serial 16x16 factorisations overlap the MFMA work.  Two barriers per
// block column.  On return the diagonal slots hold D_J = L_JJ^{-1} and the
// off-diagonal slots L_IJ; each L_JJ goes to Lout when Lout != nullptr.
// Returns 0 or the 1-based failing row (LAPACK info).
template <typename T>
__device__ __forceinline__ int factor_blocked(T *L, int nb, int n, T *Lout, int64_t ldo, int *s_flag) {
    using M = Mfma16<T>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto blk = [&](int I, int J) { return L + (tri(I) + J) * 256; };
    auto lout = [&](int J) { return (Lout != nullptr) ? Lout + 16 * J + (int64_t)16 * J * ldo : (T *)nullptr; };
    if (wave == 0) {
        const int f = factor_diag_block(blk(0, 0), lout(0), ldo, n);
        if (lane == 0) *s_flag = (f >= 0) ? f + 1 : 0;
    }
    __syncthreads();
    if (*s_flag) return *s_flag;
    for (int J = 0; J < nb; ++J) {
        const T *dJ = blk(J, J);
        // panel: L_IJ = A_IJ * D_J^T
        for (int I = J + 1 + wave; I < nb; I += kWaves) {
            T *t = blk(I, J);
            typename M::acc_t acc = M::zero();
            const int col = lane & 15;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int o = sw(col, (lane >> 4) + 4 * kk);
                acc = M::fma(t[o], dJ[o], acc);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) t[sw(M::drow(lane, r), col)] = acc[r];
        }
        __syncthreads();
        if (J + 1 == nb) break;
        if (wave == 0) {
            // lookahead: next diagonal tile first, then its factorisation
            tile_update2(blk(J + 1, J + 1), blk(J + 1, J), blk(J + 1, J), (T *)nullptr, (const T *)nullptr,
                         (const T *)nullptr, false);
            const int f = factor_diag_block(blk(J + 1, J + 1), lout(J + 1), ldo, n - 16 * (J + 1));
            if (lane == 0) *s_flag = (f >= 0) ? 16 * (J + 1) + f + 1 : 0;
        } else {
            // remaining tiles (I,K), J < K <= I, except (J+1,J+1), over waves 1..3
            int t = 0, pend = -1;
            T *pc = nullptr;
            const T *pa = nullptr, *pb = nullptr;
            for (int K = J + 1; K < nb; ++K) {
                for (int I = K; I < nb; ++I) {
                    if (I == J + 1 && K == J + 1) continue;
                    if ((t++ % (kWaves - 1)) != wave - 1) continue;
                    if (pend < 0) {
                        pc = blk(I, K);
                        pa = blk(I, J);
                        pb = blk(K, J);
                        pend = 1;
                    } else {
                        tile_update2(pc, pa, pb, blk(I, K), blk(I, J), blk(K, J), true);
                        pend = -1;
                    }
                }
            }
            if (pend > 0) tile_update2(pc, pa, pb, (T *)nullptr, (const T *)nullptr, (const T *)nullptr, false);
        }
        __syncthreads();
        if (*s_flag) return *s_flag;
    }
    return 0;
}

// In LDS, replace each diagonal block L_JJ by its inverse (wave J % 4).
template <typename T>
__device__ __forceinline__ void invert_diag_blocks(T *L, int nb) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int J = wave; J < nb; J += kWaves) {
        T *dJ = L + (tri(J) + J) * 256;
        T row[16], invd[16], x[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) row[c] = dJ[sw(lane & 15, c)];
#pragma unroll
        for (int r = 0; r < 16; ++r) invd[r] = T(1) / readlane(row[r], r);
        tri_inverse_cols(row, invd, x);
        if (lane < 16) {
#pragma unroll
            for (int r = 0; r < 16; ++r) dJ[sw(r, lane)] = x[r];
        }
    }
}

// y_i = sum_k M(i, k) v[k], or sum_k M(k, i) v[k] when TRANS (M = a D or L
// block), two partial sums.
template <bool TRANS, typename T>
__device__ __forceinline__ T dot16(const T *m, int i, const T *v) {
    T a0 = T(0), a1 = T(0);
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
        a0 += m[TRANS ? sw(k, i) : sw(i, k)] * v[k];
        a1 += m[TRANS ? sw(k + 1, i) : sw(i, k + 1)] * v[k + 1];
    }
    return a0 + a1;
}

// Solve L L^T Y = Y in place for the npad x 4 block Y (LDS, column-major,
// ld = npad) with the block triangle (diagonal slots = inverses).  Lane =
// row + 16 * rhs inside a 16-row block.  Lookahead: the wave that owns the
// next block row applies the current update to it first and then its
// diagonal product, so each block step costs one barrier.
template <typename T>
__device__ __forceinline__ void solve_blocked(const T *L, int nb, T *Y) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, r = lane >> 4;
    const int ld = 16 * nb;
    auto blk = [&](int I, int J) { return L + (tri(I) + J) * 256; };
    // forward: L y = b
    if (wave == 0) {
        T *y = Y + ld * r;
        const T v = dot16<false>(blk(0, 0), i, y);
        y[i] = v;
    }
    __syncthreads();
    for (int I = 0; I + 1 < nb; ++I) {
        const T *yI = Y + 16 * I + ld * r;
        const int owner = (I + 1) & (kWaves - 1);
        if (wave == owner) {
            T *x = Y + 16 * (I + 1) + ld * r;
            x[i] -= dot16<false>(blk(I + 1, I), i, yI);
            const T v = dot16<false>(blk(I + 1, I + 1), i, x);
            x[i] = v;
        } else {
            int t = 0;
            for (int K = I + 2; K < nb; ++K, ++t) {
                if ((t % (kWaves - 1)) != ((wave - owner - 1) & (kWaves - 1))) continue;
                T *x = Y + 16 * K + ld * r;
                x[i] -= dot16<false>(blk(K, I), i, yI);
            }
        }
        __syncthreads();
    }
    // backward: L^T x = y
    const int last = nb - 1;
    if (wave == (last & (kWaves - 1))) {
        T *y = Y + 16 * last + ld * r;
        const T v = dot16<true>(blk(last, last), i, y);
        y[i] = v;
    }
    __syncthreads();
    for (int I = last; I > 0; --I) {
        const T *zI = Y + 16 * I + ld * r;
        const int owner = (I - 1) & (kWaves - 1);
        if (wave == owner) {
            T *x = Y + 16 * (I - 1) + ld * r;
            x[i] -= dot16<true>(blk(I, I - 1), i, zI);
            const T v = dot16<true>(blk(I - 1, I - 1), i, x);
            x[i] = v;
        } else {
            int t = 0;
            for (int K = I - 2; K >= 0; --K, ++t) {
                if ((t % (kWaves - 1)) != ((wave - owner - 1) & (kWaves - 1))) continue;
                T *x = Y + 16 * K + ld * r;
                x[i] -= dot16<true>(blk(I, K), i, zI);
            }
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
    if (mask != nullptr) {
        // a masked launch mostly finds nothing to do: read this workgroup's
        // share of the mask at once (one entry per thread) instead of one
        // dependent load per matrix, and leave if none is selected
        const int64_t bt = blockIdx.x + (int64_t)threadIdx.x * gridDim.x;
        const int need = (bt < batch && mask[bt] < 0) ? 1 : 0;
        const bool covered = (int64_t)kThreads * gridDim.x >= batch;  // uniform
        if (!__syncthreads_or(need) && covered) return;
    }
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

// skip != nullptr: systems with skip[b] != 0 (a failed factorisation) are
// left as they are, as LAPACK dposv leaves B.
template <typename T>
__global__ __launch_bounds__(kThreads) void potrs_global_kernel(int n, int nrhs, const T *L,
                                                                int64_t lda, int64_t strideL, T *B,
                                                                int64_t ldb, int64_t strideB,
                                                                int *info, int64_t batch,
                                                                const int *skip = nullptr) {
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    if (skip != nullptr && skip[b] != 0) return;
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

// Copies the rows x cols matrices src_b (ld lds) to dst_b (ld ldd).
template <typename T>
__global__ __launch_bounds__(kThreads) void copy_batch_kernel(int rows, int cols, const T *__restrict__ src,
                                                              int64_t lds, int64_t strideS, T *__restrict__ dst,
                                                              int64_t ldd, int64_t strideD, int64_t batch) {
    const int64_t per = (int64_t)rows * cols;
    for (int64_t b = blockIdx.x; b < batch; b += gridDim.x)
        for (int64_t e = threadIdx.x; e < per; e += kThreads) {
            const int64_t i = e % rows, j = e / rows;
            dst[b * strideD + i + j * ldd] = src[b * strideS + i + j * lds];
        }
}

// ----------------------------------------------------------- dispatch --
template <typename T>
int potrf_dispatch(int64_t n, T *A, int64_t lda, int64_t strideA, int *info, int64_t batch,
                   hipStream_t st) {
    // factor only: the register-resident small kernel with no right-hand
    // sides (B and X are not read or written when nrhs = 0)
    if (n <= 32)
        return chols::launch_posv_small<T>((int)n, 0, A, lda, strideA, A, lda, strideA, A, lda, strideA, info, batch,
                                           nullptr, true, st);
    // n > 32 from here on (smaller n returned above)
    const dim3 grid((unsigned)batch), block(kThreads);
    if (n <= 48) hipLaunchKernelGGL((potrf_kernel<T, 3>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else if (n <= 64) hipLaunchKernelGGL((potrf_kernel<T, 4>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else if (n <= 128) hipLaunchKernelGGL((potrf_kernel<T, 8>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    else hipLaunchKernelGGL((potrf_global_kernel<T>), grid, block, 0, st, (int)n, A, lda, strideA, info, batch);
    return launch_check("potrf_kernel");
}

template <typename T>
int potrs_dispatch(int64_t n, int64_t nrhs, const T *L, int64_t lda, int64_t strideL, T *B,
                   int64_t ldb, int64_t strideB, int *info, int64_t batch, hipStream_t st) {
    if (n <= 32)
        return chols::launch_potrs_small<T>((int)n, (int)nrhs, L, lda, strideL, B, ldb, strideB, info, batch, st);
    const dim3 grid((unsigned)batch), block(kThreads);
#define GPUTILS_POTRS(NBV) \
    hipLaunchKernelGGL((potrs_kernel<T, NBV>), grid, block, 0, st, (int)n, (int)nrhs, L, lda, strideL, B, ldb, strideB, info, batch)
    // n > 32 from here on (smaller n returned above)
    if (n <= 48) GPUTILS_POTRS(3);
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
    if (n <= 32)
        return chols::launch_posv_small<T>((int)n, (int)nrhs, A, lda, strideA, B, ldb, strideB, X, ldx, strideX, info,
                                           batch, mask, WRITE_L, st);
    // masked (fallback) launches scan the mask with a grid-stride loop: two
    // workgroups per CU fill the GPU at this kernel's LDS size
    const int64_t g = (mask != nullptr) ? (batch < 512 ? batch : 512) : batch;
    const dim3 grid((unsigned)g), block(kThreads);
#define GPUTILS_POSV(NBV)                                                                     \
    hipLaunchKernelGGL((posv_kernel<T, NBV, WRITE_L>), grid, block, 0, st, (int)n, (int)nrhs, A, \
                       lda, strideA, B, ldb, strideB, X, ldx, strideX, info, batch, mask)
    // n > 32 from here on (smaller n returned above); n <= 128 (checked by
    // the callers)
    if (n <= 48) GPUTILS_POSV(3);
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
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 4, "lda < max(1,n)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 10, "batch out of range");
    if (n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(nrhs == 0 || B != nullptr, 6, "B is NULL");
    hipStream_t st = as_stream(stream);
    if (n > 128) {
        // global-memory factor, then the solve of the systems it factorised
        // (a failed one keeps its B); info holds the factor's status
        GPUTILS_ARG(info != nullptr, 9, "info is NULL (needed for n > 128)");
        const int rc = chol::potrf_dispatch<double>(n, A, lda, strideA, info, batch, st);
        if (rc != GPUTILS_OK || nrhs == 0) return rc;
        hipLaunchKernelGGL((chol::potrs_global_kernel<double>), dim3((unsigned)batch), dim3(chol::kThreads), 0, st,
                           (int)n, (int)nrhs, A, lda, strideA, B, ldb, strideB, (int *)nullptr, batch, info);
        return launch_check("potrs_global_kernel");
    }
    return chol::posv_dispatch<double, true>(n, nrhs, A, lda, strideA, B, ldb, strideB, B, ldb,
                                             strideB, info, batch, nullptr, st);
}

int64_t gputils_dposv_mixed_workspace_bytes(int64_t n, int64_t batch) {
    if (n < 0 || batch < 0) return -1;
    return n <= cholw::kMixedMaxN ? 0 : n * n * (int64_t)sizeof(double) * batch;
}

}  // extern "C"

namespace {

// Arguments of the mixed solve after their checks; work is the FP64 copy
// of A for n > kMixedMaxN (the only device memory besides the caller's).
int posv_mixed_run(int64_t n, int64_t nrhs, const double *A, int64_t lda, int64_t strideA, const double *B,
                   int64_t ldb, int64_t strideB, double *X, int64_t ldx, int64_t strideX, double tol, int max_iter,
                   int *iters, int *info, double *work, int64_t batch, void *stream) {
    hipStream_t st = as_stream(stream);
    if (n > cholw::kMixedMaxN) {
        // beyond the fused kernel: FP64 factor and solve of copies (A and B
        // stay read only), every system reported as solved by the fallback
        double *L = work;
        const unsigned blocks = (unsigned)(batch < 4096 ? batch : 4096);
        hipLaunchKernelGGL((chol::copy_batch_kernel<double>), dim3(blocks), dim3(chol::kThreads), 0, st, (int)n, (int)n, A,
                           lda, strideA, L, n, n * n, batch);
        hipLaunchKernelGGL((chol::copy_batch_kernel<double>), dim3(blocks), dim3(chol::kThreads), 0, st, (int)n, (int)nrhs,
                           B, ldb, strideB, X, ldx, strideX, batch);
        int rc = launch_check("copy_batch_kernel");
        if (rc != GPUTILS_OK) return rc;
        rc = gputils_fill_bytes_async(iters, 0xff, (size_t)batch * sizeof(int), stream);  // iters = -1
        if (rc != GPUTILS_OK) return rc;
        rc = chol::potrf_dispatch<double>(n, L, n, n * n, info, batch, st);
        if (rc != GPUTILS_OK) return rc;
        return chol::potrs_dispatch<double>(n, nrhs, L, n, n * n, X, ldx, strideX, nullptr, batch, st);
    }
    int rc = cholw::launch_posv_mixed((int)n, (int)nrhs, A, lda, strideA, B, ldb, strideB, X, ldx, strideX, tol,
                                      max_iter, iters, info, batch, st);
    if (rc != GPUTILS_OK) return rc;
    // FP64 fallback for matrices flagged with iters < 0 (A, B read only)
    return chol::posv_dispatch<double, false>(n, nrhs, const_cast<double *>(A), lda, strideA, B, ldb,
                                              strideB, X, ldx, strideX, info, batch, iters, st);
}

// The argument checks both mixed entry points share (their argument
// numbers agree up to info, 15).
int posv_mixed_check(int64_t n, int64_t nrhs, const double *A, int64_t lda, const double *B, int64_t ldb,
                     const double *X, int64_t ldx, double tol, int max_iter, const int *iters, const int *info,
                     int64_t batch, int batch_arg) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");
    GPUTILS_ARG(lda >= (n > 1 ? n : 1), 4, "lda < max(1,n)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(ldx >= (n > 1 ? n : 1), 10, "ldx < max(1,n)");
    GPUTILS_ARG(tol > 0.0, 12, "tol <= 0");
    GPUTILS_ARG(max_iter >= 0, 13, "max_iter < 0");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, batch_arg, "batch out of range");
    if (n == 0 || nrhs == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(B != nullptr, 6, "B is NULL");
    GPUTILS_ARG(X != nullptr, 9, "X is NULL");
    GPUTILS_ARG(iters != nullptr, 14, "iters is NULL");
    GPUTILS_ARG(info != nullptr, 15, "info is NULL");
    return GPUTILS_OK;
}

// Scratch of gputils_dposv_mixed_batched_noalloc: the FP64 copies of A that
// systems beyond the fused kernel (n > kMixedMaxN) are factorised in.  Each
// GPU holds its own copy for the life of the process; every call on one GPU
// uses the same array.
constexpr int64_t kNoallocScratchBytes = 16ll << 20;
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
double *g_scratch_addr[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on the current device, looked up once per
// device (the host name of the array is only a stand-in for it).
int scratch_on_current_device(double **addr) {
    int dev = -1;
    GPUTILS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d is beyond the %d devices the allocation-free entry point supports", dev, kMaxDevices);
        return 1;
    }
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (g_scratch_addr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return e != hipSuccess ? (int)e : 1;
        }
        g_scratch_addr[dev] = static_cast<double *>(p);
    }
    *addr = g_scratch_addr[dev];
    return GPUTILS_OK;
}

}  // namespace

extern "C" {

int gputils_dposv_mixed_batched(int64_t n, int64_t nrhs, const double *A, int64_t lda,
                                int64_t strideA, const double *B, int64_t ldb, int64_t strideB,
                                double *X, int64_t ldx, int64_t strideX, double tol, int max_iter,
                                int *iters, int *info, void *work, int64_t batch, void *stream) {
    int rc = posv_mixed_check(n, nrhs, A, lda, B, ldb, X, ldx, tol, max_iter, iters, info, batch, 17);
    if (rc != GPUTILS_OK || n == 0 || nrhs == 0 || batch == 0) return rc;
    GPUTILS_ARG(n <= cholw::kMixedMaxN || work != nullptr, 16,
                "work is NULL (query gputils_dposv_mixed_workspace_bytes)");
    return posv_mixed_run(n, nrhs, A, lda, strideA, B, ldb, strideB, X, ldx, strideX, tol, max_iter, iters, info,
                          static_cast<double *>(work), batch, stream);
}

int gputils_dposv_mixed_batched_noalloc(int64_t n, int64_t nrhs, const double *A, int64_t lda, int64_t strideA,
                                        const double *B, int64_t ldb, int64_t strideB, double *X, int64_t ldx,
                                        int64_t strideX, double tol, int max_iter, int *iters, int *info,
                                        int64_t batch, void *stream) {
    int rc = posv_mixed_check(n, nrhs, A, lda, B, ldb, X, ldx, tol, max_iter, iters, info, batch, 16);
    if (rc != GPUTILS_OK || n == 0 || nrhs == 0 || batch == 0) return rc;
    double *work = nullptr;
    if (n > cholw::kMixedMaxN) {
        // n <= 1448 keeps n*n*8*batch inside int64 for any batch <= 2^31
        GPUTILS_ARG(n <= 1448 && n * n * (int64_t)sizeof(double) * batch <= kNoallocScratchBytes, 16,
                    "n*n*batch*8 bytes exceed the 16 MiB scratch of gputils_dposv_mixed_batched_noalloc");
        rc = scratch_on_current_device(&work);
        if (rc != GPUTILS_OK) return rc;
    }
    return posv_mixed_run(n, nrhs, A, lda, strideA, B, ldb, strideB, X, ldx, strideX, tol, max_iter, iters, info,
                          work, batch, stream);
}

}  // extern "C"
