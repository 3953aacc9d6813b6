// chol_mixed.hip -- mixed-precision batched Cholesky solve for gfx950, one
// wavefront per matrix with the FP32 factor held in registers.
//
// A = U^T U (U upper, FP32) is computed and kept as 16x16 blocks in the
// MFMA accumulator layout: lane (g = lane/16, c = lane%16), item r holds
// element (4g + r, c) of a block.  For two blocks M, N in that layout the
// four 16x16x4 MFMA steps that take item r of M as the A operand and item
// r of N as the B operand sum over k = 4g + r and produce M^T N, again in
// the accumulator layout.  Every product the right-looking factorisation
// and the forward solve need has that form:
//   panel      U_JI  = W_J^T A_JI          (W_J = U_JJ^{-1})
//   update     A_IK -= U_JI^T U_JK
//   forward    y_I   = W_I^T (b_I - sum_J U_JI^T y_J)
// so they run without moving data between lanes.  Off-diagonal blocks live
// in registers (right-looking updates); the diagonal blocks live in LDS as
// FP32, receive their updates left-looking (A_II - sum_J U_JI^T U_JI) right
// before they are factored by rows with lane broadcasts, and are then
// replaced by W_I.  The backward solve (products with U, not U^T) reads
// W_I transposed from LDS and transposes the off-diagonal blocks one at a
// time through LDS.  Refinement residuals r = b - A x are formed in FP64 from
// the FP64 matrix (lower triangle) in HBM/L2; the iterate x stays FP64 in
// LDS.  Stopping test per right-hand side (LAPACK DSPOSV):
//   ||r||_inf <= tol (||A||_inf ||x||_inf + ||b||_inf).
// Right-hand sides are processed four at a time (columns 0..3 of a 16x16
// vector block; the other columns are zero).
#include <type_traits>

#include "chol_mixed.hpp"
#include "common.hpp"

namespace gputils {
namespace cholw {

using f4 = f32x4;

constexpr int kRhs = 4;

__host__ __device__ constexpr int uo(int P, int Q) { return (Q * (Q - 1)) / 2 + P; }  // block (P < Q)
__host__ __device__ constexpr int noff(int NB) { return NB > 1 ? NB * (NB - 1) / 2 : 1; }  // array size

__device__ __forceinline__ f4 zero4() { return f4{0.f, 0.f, 0.f, 0.f}; }

// f(std::integral_constant<int, i>) for i = B .. E-1: block indices are
// compile-time constants, so register-resident block arrays are never
// indexed dynamically (which would move them to scratch memory).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// The lane id as a value the compiler cannot treat as loop invariant:
// values derived from it are recomputed where they are used instead of
// being hoisted out of the refinement loop and held in registers.
__device__ __forceinline__ int opaque_lane() {
    int l;
    asm volatile("v_mov_b32 %0, %1" : "=v"(l) : "v"((int)(threadIdx.x & 63)));
    return l;
}

// Wave-uniform values moved to scalar registers.
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// C + M^T N (see the file comment).
__device__ __forceinline__ f4 mtn(const f4 &M, const f4 &N, f4 C) {
    C = __builtin_amdgcn_mfma_f32_16x16x4f32(M[0], N[0], C, 0, 0, 0);
    C = __builtin_amdgcn_mfma_f32_16x16x4f32(M[1], N[1], C, 0, 0, 0);
    C = __builtin_amdgcn_mfma_f32_16x16x4f32(M[2], N[2], C, 0, 0, 0);
    C = __builtin_amdgcn_mfma_f32_16x16x4f32(M[3], N[3], C, 0, 0, 0);
    return C;
}

// A column-major 16x16 LDS block in the accumulator layout, and its
// transpose.
__device__ __forceinline__ f4 lds_block(const float *blk, int lane) {
    const int g = lane >> 4, c = lane & 15;
    return *reinterpret_cast<const f4 *>(blk + 16 * c + 4 * g);
}
__device__ __forceinline__ f4 lds_block_t(const float *blk, int lane) {
    const int g = lane >> 4, c = lane & 15;
    const float *src = blk + c + 64 * g;
    f4 t;
#pragma unroll
    for (int r = 0; r < 4; ++r) t[r] = src[16 * r];
    return t;
}

// M^T in the accumulator layout, through 256 floats of LDS.  LDS accesses
// of one wavefront complete in program order, so no fence is needed.
__device__ __forceinline__ f4 transpose(const f4 &M, float *S, int lane) {
    const int g = lane >> 4, c = lane & 15;
    *reinterpret_cast<f4 *>(S + 16 * c + 4 * g) = M;  // column-major M
    return lds_block_t(S, lane);
}

// Factor the symmetric block D = Ut^T Ut (Ut upper), write W = Ut^{-1}
// column-major to `blk` and return it in the accumulator layout.  Lane c
// (and its copies in the other lane groups) works on row c of L = Ut^T.
// fail receives -1 or the index of the first non-positive pivot (uniform).
__device__ __forceinline__ f4 diag_factor(const f4 &D, float *blk, int &fail) {
    const int lane = opaque_lane(), g = lane >> 4, c = lane & 15;  // lane masks stay local
    *reinterpret_cast<f4 *>(blk + 16 * c + 4 * g) = D;
    float row[16], invd[16];
    const float *src = blk + c;
#pragma unroll
    for (int k = 0; k < 16; ++k) row[k] = src[16 * k];
    fail = -1;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float d = readlane(row[k], k);
        if (!(d > 0.f)) {
            fail = k;
            return D;
        }
        const float inv = uniform(__builtin_amdgcn_rsqf(d));
        invd[k] = inv;
        const float lk = (c > k) ? row[k] * inv : ((c == k) ? d * inv : 0.f);
        row[k] = lk;
#pragma unroll
        for (int j = k + 1; j < 16; ++j) {
            const float ljk = readlane(lk, j);
            if (c >= j) row[j] -= lk * ljk;
        }
    }
    // x[r] = (L^{-1})[r][c], column c of L^{-1} = row c of W
    float x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        float acc = (c == r) ? 1.f : 0.f;
#pragma unroll
        for (int p = 0; p < r; ++p) acc -= readlane(row[p], r) * x[p];
        x[r] = acc * invd[r];
    }
    if (lane < 16) {
        float *dst = blk + c;
#pragma unroll
        for (int r = 0; r < 16; ++r) dst[16 * r] = x[r];  // W column-major
    }
    return lds_block(blk, lane);
}

// Element (i, j) of the FP64 matrix from its lower triangle; `outside` is
// returned beyond n.  Branch-free: the load always reads a valid address.
__device__ __forceinline__ double a_low(const double *__restrict__ a, int64_t lda, int n, int i, int j,
                                        double outside) {
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const bool in = hi < n;
    const double v = a[in ? hi + (int64_t)lo * lda : 0];
    return in ? v : outside;
}

// Off-diagonal upper blocks to registers and diagonal blocks to LDS
// (column-major, FP32), identity beyond n.
template <int NB>
__device__ __forceinline__ void load_matrix(const double *__restrict__ a, int64_t lda, int n, f4 (&U)[noff(NB)],
                                            float *DS) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    static_for<0, NB>([&](auto Qc) {
        constexpr int Q = Qc;
        static_for<0, Q + 1>([&](auto Pc) {
            constexpr int P = Pc;
            f4 v;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = 16 * P + 4 * g + r, gc = 16 * Q + c;
                // identity padding only touches diagonal blocks
                const double outside = (P == Q && 4 * g + r == c) ? 1.0 : 0.0;
                v[r] = (float)a_low(a, lda, n, gr, gc, outside);
            }
            if constexpr (P < Q) U[uo(P, Q)] = v;
            else *reinterpret_cast<f4 *>(DS + 256 * P + 16 * c + 4 * g) = v;
        });
        // bound the loads in flight (and their registers) to one block column
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Blocked factorisation: off-diagonal blocks right-looking in registers,
// diagonal blocks left-looking from LDS; DS[J] ends up holding W_J.
// Returns 0 or the 1-based failing row.
template <int NB>
__device__ __forceinline__ int factor(f4 (&U)[noff(NB)], float *DS) {
    // A runtime loop over block columns, with the block-index-specific
    // register work selected by uniform branches: the diagonal
    // factorisation is emitted once, which keeps the kernel small enough for
    // the instruction cache.
#pragma unroll 1
    for (int J = 0; J < NB; ++J) {
        f4 d = lds_block(DS + 256 * J, opaque_lane());
        static_for<1, NB>([&](auto Jc) {
            if (J == Jc) static_for<0, Jc>([&](auto Ic) { d = mtn(-U[uo(Ic, Jc)], U[uo(Ic, Jc)], d); });
        });
        int fail;
        const f4 w = diag_factor(d, DS + 256 * J, fail);
        if (fail >= 0) return 16 * J + fail + 1;
        static_for<0, NB - 1>([&](auto Jc) {
            constexpr int JJ = Jc;
            if (J == JJ) {
                static_for<JJ + 1, NB>([&](auto Ic) { U[uo(JJ, Ic)] = mtn(w, U[uo(JJ, Ic)], zero4()); });
                static_for<JJ + 1, NB>([&](auto Ic) {
                    constexpr int I = Ic;
                    const f4 m = -U[uo(JJ, I)];
                    static_for<I + 1, NB>([&](auto Kc) { U[uo(I, Kc)] = mtn(m, U[uo(JJ, Kc)], U[uo(I, Kc)]); });
                });
            }
        });
    }
    return 0;
}

// Y <- U^{-T} Y, then Y <- U^{-1} Y (vector blocks).
template <int NB>
__device__ __forceinline__ void solve(const f4 (&U)[noff(NB)], const float *DS, f4 (&Y)[NB], float *S, int lane) {
    static_for<0, NB>([&](auto Ic) {
        constexpr int I = Ic;
        f4 s = zero4();
        static_for<0, I>([&](auto Jc) { s = mtn(U[uo(Jc, I)], Y[Jc], s); });
        Y[I] = mtn(lds_block(DS + 256 * I, lane), Y[I] - s, zero4());
    });
    static_for<0, NB>([&](auto tc) {
        constexpr int I = NB - 1 - tc;
        f4 s = zero4();
        static_for<I + 1, NB>([&](auto Kc) { s = mtn(transpose(U[uo(I, Kc)], S, lane), Y[Kc], s); });
        Y[I] = mtn(lds_block_t(DS + 256 * I, lane), Y[I] - s, zero4());
    });
}

// Residual task t (output block Out, inner block X), row-major over the
// NB x NB block grid.
__host__ __device__ constexpr int task_out(int t, int NB) { return t / NB; }
__host__ __device__ constexpr int task_x(int t, int NB) { return t % NB; }

// The four FP64 elements A(16 Out + c, 16 X + 4g + r), r = 0..3, of task t
// read from the lower triangle.  Below the diagonal the lanes of a lane
// group read consecutive rows of a column; above it each lane reads four
// consecutive rows of its own column.  n = 16 NB (no padding).
template <int NB, int T>
__device__ __forceinline__ void task_load(const double *__restrict__ a, int64_t lda, int lane, double (&v)[4]) {
    constexpr int Out = task_out(T, NB), X = task_x(T, NB);
    const int g = lane >> 4, c = lane & 15;
    const int o = 16 * Out + c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int xr = 16 * X + 4 * g + r;
        if constexpr (X < Out) v[r] = a[o + (int64_t)xr * lda];
        else if constexpr (X > Out) v[r] = a[xr + (int64_t)o * lda];
        else v[r] = a[(o > xr ? o : xr) + (int64_t)(o > xr ? xr : o) * lda];
    }
}

// RD <- b - A XD (FP64, all 16*NB padded rows, 4 columns, [row][4] layout;
// b from HBM/L2, zero beyond n and nr); when want_asum, ASUM[i] receives
// sum_j |A(i,j)|.  Lane (g, c) accumulates output row 16 Out + c over
// inner rows 16 X + 4g + r of block X; two xor shuffles finish each output
// block.
//
// n = 16 NB: the NB^2 block tasks are unrolled and the matrix loads run four
// tasks ahead of the arithmetic.
template <int NB>
__device__ __forceinline__ void residual_full(const double *__restrict__ a, int64_t lda, const double *XD,
                                              const double *__restrict__ bb, int64_t ldb, int nr, double *RD,
                                              double *ASUM, bool want_asum, int lane) {
    constexpr int NT = NB * NB, D = 4;
    const int g = lane >> 4, c = lane & 15;
    double ring[D][4];
    static_for<0, (D < NT ? D : NT)>([&](auto T) { task_load<NB, T>(a, lda, lane, ring[T]); });
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, as = 0.0;
    static_for<0, NT>([&](auto T) {
        constexpr int Out = task_out(T, NB), X = task_x(T, NB);
        double av[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) av[r] = ring[T % D][r];
        if constexpr (T + D < NT) task_load<NB, T + D>(a, lda, lane, ring[T % D]);
        __builtin_amdgcn_sched_barrier(0);
        const double *xb = XD + 4 * (16 * X + 4 * g);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double *x = xb + 4 * r;
            p0 += av[r] * x[0];
            p1 += av[r] * x[1];
            p2 += av[r] * x[2];
            p3 += av[r] * x[3];
            as += fabs(av[r]);
        }
        if constexpr (X == NB - 1) {
#pragma unroll
            for (int off = 16; off <= 32; off <<= 1) {
                p0 += __shfl_xor(p0, off, kWave);
                p1 += __shfl_xor(p1, off, kWave);
                p2 += __shfl_xor(p2, off, kWave);
                p3 += __shfl_xor(p3, off, kWave);
                as += __shfl_xor(as, off, kWave);
            }
            const int o = 16 * Out + c;
            const double pg = (g == 0) ? p0 : (g == 1) ? p1 : (g == 2) ? p2 : p3;
            const bool in = g < nr;
            const double bvv = bb[in ? o + (int64_t)g * ldb : 0];
            RD[4 * o + g] = (in ? bvv : 0.0) - pg;
            if (want_asum && g == 0) ASUM[o] = as;
            p0 = p1 = p2 = p3 = as = 0.0;
        }
    });
}

// Any n <= 16 NB (identity padding): runtime loops over the block grid.
template <int NB>
__device__ __forceinline__ void residual_padded(const double *__restrict__ a, int64_t lda, int n, const double *XD,
                                                const double *__restrict__ bb, int64_t ldb, int nr, double *RD,
                                                double *ASUM, bool want_asum, int lane) {
    const int g = lane >> 4, c = lane & 15;
#pragma unroll 1
    for (int Out = 0; Out < NB; ++Out) {
        const int o = 16 * Out + c;
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, as = 0.0;
#pragma unroll 1
        for (int X = 0; X < NB; ++X) {
            double av[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = a_low(a, lda, n, o, 16 * X + 4 * g + r, 0.0);
            const double *xb = XD + 4 * (16 * X + 4 * g);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double *x = xb + 4 * r;
                p0 += av[r] * x[0];
                p1 += av[r] * x[1];
                p2 += av[r] * x[2];
                p3 += av[r] * x[3];
                as += fabs(av[r]);
            }
        }
#pragma unroll
        for (int off = 16; off <= 32; off <<= 1) {
            p0 += __shfl_xor(p0, off, kWave);
            p1 += __shfl_xor(p1, off, kWave);
            p2 += __shfl_xor(p2, off, kWave);
            p3 += __shfl_xor(p3, off, kWave);
            as += __shfl_xor(as, off, kWave);
        }
        const double pg = (g == 0) ? p0 : (g == 1) ? p1 : (g == 2) ? p2 : p3;
        const bool in = o < n && g < nr;
        const double bvv = bb[in ? o + (int64_t)g * ldb : 0];
        RD[4 * o + g] = (in ? bvv : 0.0) - pg;
        if (want_asum && g == 0) ASUM[o] = as;
    }
}

template <int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void posv_mixed_wave_kernel(
    int n, int nrhs, const double *__restrict__ A, int64_t lda, int64_t strideA, const double *__restrict__ B,
    int64_t ldb, int64_t strideB, double *__restrict__ X, int64_t ldx, int64_t strideX, double tol, int max_iter,
    int *iters, int *info, int64_t batch) {
    constexpr int NP = 16 * NB;
    __shared__ __attribute__((aligned(16))) float S[256];
    __shared__ __attribute__((aligned(16))) float DS[256 * NB];
    __shared__ __attribute__((aligned(16))) double XD[NP * kRhs];
    __shared__ __attribute__((aligned(16))) double RD[NP * kRhs];
    __shared__ double ASUM[NP];
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int lane = threadIdx.x;
    const double *a = A + b * strideA;
    double *xx = X + b * strideX;
    if (lane == 0) info[b] = 0;

    f4 U[noff(NB)];
    load_matrix<NB>(a, lda, n, U, DS);
    if (factor<NB>(U, DS) != 0) {
        if (lane == 0) iters[b] = -1;
        return;
    }
    double anorm = 0.0;
    int used = 0;
    for (int r0 = 0; r0 < nrhs; r0 += kRhs) {
        const int nr = (nrhs - r0 < kRhs) ? (nrhs - r0) : kRhs;
        const double *bb = B + b * strideB + (int64_t)r0 * ldb;
        // b (zero-padded) into RD: the first right-hand side; x = 0
#pragma unroll 1
        for (int e = lane; e < NP * kRhs; e += 64) {
            const int row = e >> 2, q = e & 3;
            const bool in = row < n && q < nr;
            const double v = bb[in ? row + (int64_t)q * ldb : 0];
            RD[e] = in ? v : 0.0;
            XD[e] = 0.0;
        }
        double bn[kRhs];
#pragma unroll
        for (int q = 0; q < kRhs; ++q) {
            double m = 0.0;
#pragma unroll 1
            for (int row = lane; row < NP; row += 64) m = fmax(m, fabs(RD[4 * row + q]));
            bn[q] = uniform(wave_max(m));
        }
        int it = 0;
        bool converged = false;
        for (;;) {
            asm volatile("" ::: "memory");
            const int ln = opaque_lane();
            const bool col_ok = (ln & 15) < kRhs;
            f4 Y[NB];
            {
                const double *src = RD + 16 * (ln >> 4) + (ln & 3);
                static_for<0, NB>([&](auto Ic) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) Y[Ic][r] = col_ok ? (float)src[64 * Ic + 4 * r] : 0.f;
                });
            }
            solve<NB>(U, DS, Y, S, ln);
            if (col_ok) {
                double *dst = XD + 16 * (ln >> 4) + (ln & 3);
                static_for<0, NB>([&](auto Ic) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) dst[64 * Ic + 4 * r] += (double)Y[Ic][r];
                });
            }
            if (n == 16 * NB) residual_full<NB>(a, lda, XD, bb, ldb, nr, RD, ASUM, anorm == 0.0, ln);
            else residual_padded<NB>(a, lda, n, XD, bb, ldb, nr, RD, ASUM, anorm == 0.0, ln);
            if (anorm == 0.0) {
                double m = 0.0;
#pragma unroll 1
                for (int row = ln; row < NP; row += 64) m = fmax(m, ASUM[row]);
                anorm = uniform(wave_max(m));
                if (anorm == 0.0) anorm = 1e-300;  // A = 0 is not SPD; keep the test finite
            }
            bool ok = true;
#pragma unroll
            for (int q = 0; q < kRhs; ++q) {
                double rm = 0.0, xm = 0.0;
#pragma unroll 1
                for (int row = ln; row < NP; row += 64) {
                    rm = fmax(rm, fabs(RD[4 * row + q]));
                    xm = fmax(xm, fabs(XD[4 * row + q]));
                }
                rm = wave_max(rm);
                xm = wave_max(xm);
                if (q < nr) ok = ok && (rm <= tol * (anorm * xm + bn[q]));
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
#pragma unroll 1
        for (int e = lane; e < NP * kRhs; e += 64) {
            const int row = e >> 2, q = e & 3;
            if (row < n && q < nr) xx[row + (int64_t)(r0 + q) * ldx] = XD[e];
        }
    }
    if (lane == 0) iters[b] = used;
}

int launch_posv_mixed(int n, int nrhs, const double *A, int64_t lda, int64_t strideA, const double *B, int64_t ldb,
                      int64_t strideB, double *X, int64_t ldx, int64_t strideX, double tol, int max_iter, int *iters,
                      int *info, int64_t batch, hipStream_t st) {
    const dim3 grid((unsigned)batch), block(64);
#define GPUTILS_MIXED_WAVE(NBV)                                                                                    \
    hipLaunchKernelGGL((posv_mixed_wave_kernel<NBV>), grid, block, 0, st, n, nrhs, A, lda, strideA, B, ldb, strideB, \
                       X, ldx, strideX, tol, max_iter, iters, info, batch)
    switch ((n + 15) >> 4) {
        case 1: GPUTILS_MIXED_WAVE(1); break;
        case 2: GPUTILS_MIXED_WAVE(2); break;
        case 3: GPUTILS_MIXED_WAVE(3); break;
        case 4: GPUTILS_MIXED_WAVE(4); break;
        case 5: GPUTILS_MIXED_WAVE(5); break;
        case 6: GPUTILS_MIXED_WAVE(6); break;
        case 7: GPUTILS_MIXED_WAVE(7); break;
        default: GPUTILS_MIXED_WAVE(8); break;
    }
#undef GPUTILS_MIXED_WAVE
    return launch_check("posv_mixed_wave_kernel");
}

}  // namespace cholw
}  // namespace gputils
