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
// before they are factored (four-column panels with rank-4 MFMA updates),
// and are then replaced by W_I.  The backward solve (products with U, not
// U^T) reads W_I transposed from LDS and transposes the off-diagonal blocks
// one at a time through LDS.  Refinement residuals r = b - A x are formed in
// FP64 from the FP64 matrix in HBM, reading each 16x16 block of the lower
// triangle once per residual; the iterate x stays FP64 in LDS, one element
// per lane is read per block and the lanes that need it get it by DPP
// broadcast, and each block's four right-hand-side sums are reduced over
// the lane rows by one reduce-scatter.  Stopping
// test per right-hand side (LAPACK DSPOSV):
//   ||r||_inf <= tol (||A||_inf ||x||_inf + ||b||_inf).
// Right-hand sides are processed four at a time (columns 0..3 of a 16x16
// vector block; the other columns are zero).
#include <cfloat>
#include <type_traits>

#include "chol_mixed.hpp"
#include "common.hpp"

namespace gputils {
namespace cholw {

using f4 = f32x4;

constexpr int kRhs = 4;
// LDS tile of the residual: a 16 x 17 transpose tile followed by 16 sink
// words that absorb the stores of lanes with nothing to add
constexpr int kSinkOff = 16 * 17, kTileWords = 16 * 17 + 16;
// largest leading dimension for which in-block byte offsets fit 32 bits
constexpr int64_t kSymMaxLda = (int64_t)1 << 24;

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

// Wave-wide maximum of a double by DPP: row shifts 1, 2, 4, 8 give lane 15
// of each row the row maximum, row broadcasts 15 and 31 carry it to lane 63.
// Lanes whose DPP source is outside the row or masked keep their own value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_max_step(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = (int)(bits & 0xffffffffll), hi = (int)(bits >> 32);
    const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
    const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
    return fmax(v, __longlong_as_double(((long long)ohi << 32) | (unsigned)olo));
}
__device__ __forceinline__ double dpp_wave_max(double v) {
    v = dpp_max_step<0x111, 0xF>(v);  // row_shr:1
    v = dpp_max_step<0x112, 0xF>(v);  // row_shr:2
    v = dpp_max_step<0x114, 0xF>(v);  // row_shr:4
    v = dpp_max_step<0x118, 0xF>(v);  // row_shr:8
    v = dpp_max_step<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_max_step<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return readlane(v, 63);
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
// column-major to `blk` and return it in the accumulator layout.
//
// Blocked by four columns with rank-4 MFMA updates.  Item r of lane (g, c)
// holds element (4g + r, c) = (c, 4g + r) of the symmetric block, so lane
// group p holds columns 4p..4p+3 of every row c.  Panel p is factored there
// by rows (pivot and panel rows broadcast with readlane), moved through LDS
// into the MFMA operand layout (lane (g, c) <- L(c, 4p + g)), and one
// 16x16x4 MFMA subtracts L_p L_p^T from the whole block (the rows already
// factored are kept aside).  X = L^{-1} is formed the same way by block
// rows: X_p = L_pp^{-1} E_p on lane group p, then E -= L(:, 4p..4p+3) X_p.
// fail receives -1 or the index of the first non-positive pivot (uniform).
__device__ __forceinline__ f4 diag_factor(const f4 &D, float *blk, int &fail) {
    const int lane = opaque_lane(), g = lane >> 4, c = lane & 15;  // lane masks stay local
    f4 C = D;
    float Lop[4];  // MFMA operand of panel p: lane (g, c) holds L(c, 4p + g)
    float inv[16];       // 1 / L(k, k), uniform
    float lsub[4][4][4];  // L(4p + i, 4p + j), j < i, uniform
    fail = -1;
    static_for<0, 4>([&](auto Pc) {
        constexpr int p = Pc;
        f4 a = C;
        static_for<0, 4>([&](auto Kc) {
            constexpr int k = Kc;
            float d = readlane(a[k], 16 * p + 4 * p + k);  // pivot L(4p+k, 4p+k)^2
            if (!(d > 0.f)) {
                if (fail < 0) fail = 4 * p + k;
                d = 1.f;
            }
            const float iv = uniform(__builtin_amdgcn_rsqf(d));
            inv[4 * p + k] = iv;
            const float lk = (c > 4 * p + k) ? a[k] * iv : ((c == 4 * p + k) ? d * iv : 0.f);
            a[k] = lk;
            // rows above 4p + j only change entries that are never read again
            static_for<k + 1, 4>([&](auto Jc) {
                constexpr int j = Jc;
                a[j] = fmaf(-lk, readlane(lk, 16 * p + 4 * p + j), a[j]);
            });
        });
        static_for<1, 4>([&](auto Ic) {
            static_for<0, Ic>([&](auto Jc) {
                constexpr int i = Ic, j = Jc;
                lsub[p][i][j] = readlane(a[j], 16 * p + 4 * p + i);
            });
        });
        // lane (g, c) <- item g of lane (p, c)
        *reinterpret_cast<f4 *>(blk + 4 * lane) = a;
        Lop[p] = blk[64 * p + 4 * c + g];
        if constexpr (p < 3) C = __builtin_amdgcn_mfma_f32_16x16x4f32(-Lop[p], Lop[p], C, 0, 0, 0);
    });
    // E = I, then block rows of X = L^{-1}
    f4 E, Xkeep = zero4();
#pragma unroll
    for (int r = 0; r < 4; ++r) E[r] = (4 * g + r == c) ? 1.f : 0.f;
    static_for<0, 4>([&](auto Pc) {
        constexpr int p = Pc;
        f4 x;
        x[0] = E[0] * inv[4 * p];
        x[1] = (E[1] - lsub[p][1][0] * x[0]) * inv[4 * p + 1];
        x[2] = (E[2] - lsub[p][2][0] * x[0] - lsub[p][2][1] * x[1]) * inv[4 * p + 2];
        x[3] = (E[3] - lsub[p][3][0] * x[0] - lsub[p][3][1] * x[1] - lsub[p][3][2] * x[2]) * inv[4 * p + 3];
        Xkeep = (g == p) ? x : Xkeep;
        if constexpr (p < 3) {
            *reinterpret_cast<f4 *>(blk + 4 * lane) = x;
            const float bx = blk[64 * p + 4 * c + g];
            E = __builtin_amdgcn_mfma_f32_16x16x4f32(-Lop[p], bx, E, 0, 0, 0);
        }
    });
    // W(i, j) = X(j, i) column-major: element (4g + r, c) of X at 16 (4g + r) + c
#pragma unroll
    for (int r = 0; r < 4; ++r) blk[16 * (4 * g + r) + c] = Xkeep[r];
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
// diagonal blocks left-looking from LDS; DS[J] ends up holding W_J.  Each
// block product of an update is accumulated from zero and then added to
// its target: the 16x16x4 MFMA rounds after every product, so
// accumulating straight into the Schur complement would round 16 times
// per block product at the complement's magnitude, which doubles the
// factor's error and with it the refinement steps needed.
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
            if (J == Jc) static_for<0, Jc>([&](auto Ic) { d += mtn(-U[uo(Ic, Jc)], U[uo(Ic, Jc)], zero4()); });
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
                    static_for<I + 1, NB>([&](auto Kc) { U[uo(I, Kc)] += mtn(m, U[uo(JJ, Kc)], zero4()); });
                });
            }
        });
    }
    return 0;
}

// Y <- U^{-T} Y, then Y <- U^{-1} Y (vector blocks).
// The block products of one block row accumulate alternately into two
// sums (two independent chains of MFMAs).
template <int NB>
__device__ __forceinline__ void solve(const f4 (&U)[noff(NB)], const float *DS, f4 (&Y)[NB], float *S, int lane) {
    static_for<0, NB>([&](auto Ic) {
        constexpr int I = Ic;
        f4 s0 = zero4(), s1 = zero4();
        static_for<0, I>([&](auto Jc) {
            if constexpr (Jc % 2 == 0) s0 = mtn(U[uo(Jc, I)], Y[Jc], s0);
            else s1 = mtn(U[uo(Jc, I)], Y[Jc], s1);
        });
        Y[I] = mtn(lds_block(DS + 256 * I, lane), Y[I] - (s0 + s1), zero4());
    });
    static_for<0, NB>([&](auto tc) {
        constexpr int I = NB - 1 - tc;
        f4 s0 = zero4(), s1 = zero4();
        static_for<I + 1, NB>([&](auto Kc) {
            if constexpr ((Kc - I) % 2 == 1) s0 = mtn(transpose(U[uo(I, Kc)], S, lane), Y[Kc], s0);
            else s1 = mtn(transpose(U[uo(I, Kc)], S, lane), Y[Kc], s1);
        });
        Y[I] = mtn(lds_block_t(DS + 256 * I, lane), Y[I] - (s0 + s1), zero4());
    });
}

// Sum of v over the four 16-lane rows of the wavefront (every lane receives
// the sum of its column): row pairs (0,1), (2,3) exchanged by
// v_permlane16_swap, then row halves by v_permlane32_swap.
__device__ __forceinline__ double row_group_sum(double v) {
    const long long bits = __double_as_longlong(v);
    unsigned lo = (unsigned)(bits & 0xffffffffll), hi = (unsigned)(bits >> 32);
    auto join = [](unsigned l, unsigned h) { return __longlong_as_double(((long long)h << 32) | l); };
    const auto l16 = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h16 = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double t = join(l16[0], h16[0]) + join(l16[1], h16[1]);
    const long long tb = __double_as_longlong(t);
    lo = (unsigned)(tb & 0xffffffffll);
    hi = (unsigned)(tb >> 32);
    const auto l32 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h32 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return join(l32[0], h32[0]) + join(l32[1], h32[1]);
}

// Reduce-scatter of four values over the four 16-lane rows: row g
// receives the sum over all rows of s[rs_rhs(g)], rs_rhs(g) = 2 (g & 1) +
// (g >> 1).  v_permlane16_swap(s_k, s_k+2) leaves the pair sum of s_k in
// the even rows and of s_k+2 in the odd rows (k = 0, 1); v_permlane32_swap
// of those two then adds the row halves.  Six swaps and three adds, where
// four full row sums take sixteen swaps and eight adds.
__device__ __forceinline__ int rs_rhs(int g) { return ((g & 1) << 1) | (g >> 1); }
__device__ __forceinline__ double row_group_reduce_scatter(const double (&s)[4]) {
    auto lo = [](double v) { return (unsigned)(__double_as_longlong(v) & 0xffffffffll); };
    auto hi = [](double v) { return (unsigned)(__double_as_longlong(v) >> 32); };
    auto join = [](unsigned l, unsigned h) { return __longlong_as_double(((long long)h << 32) | l); };
    double t[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const auto l16 = __builtin_amdgcn_permlane16_swap(lo(s[k]), lo(s[k + 2]), false, false);
        const auto h16 = __builtin_amdgcn_permlane16_swap(hi(s[k]), hi(s[k + 2]), false, false);
        t[k] = join(l16[0], h16[0]) + join(l16[1], h16[1]);
    }
    const auto l32 = __builtin_amdgcn_permlane32_swap(lo(t[0]), lo(t[1]), false, false);
    const auto h32 = __builtin_amdgcn_permlane32_swap(hi(t[0]), hi(t[1]), false, false);
    return join(l32[0], h32[0]) + join(l32[1], h32[1]);
}

// s[q] += x(lane 16 (lane / 16) + 4 r + q) a[r] for r, q < 4: sixteen FP64
// FMAs whose first operand is x broadcast within the 16-lane row by DPP
// (row_newbcast, the DPP control 64-bit arithmetic takes), so no separate
// broadcast moves are issued (the compiler does not fold them into the
// FMA).  Two wait states before (x may have just been written by the VALU:
// a DPP source needs two) and after (the sums are read next by the lane
// swaps of the reduction, which need two after a VALU write).
__device__ __forceinline__ void fma_bcast16(double (&s)[4], const double (&a)[4], double x) {
    asm("s_nop 1\n"
        "v_fmac_f64_dpp %0, %4, %5 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %1, %4, %5 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %2, %4, %5 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %3, %4, %5 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %4, %6 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %1, %4, %6 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %2, %4, %6 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %3, %4, %6 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %4, %7 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %1, %4, %7 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %2, %4, %7 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %3, %4, %7 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %0, %4, %8 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %1, %4, %8 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %2, %4, %8 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
        "v_fmac_f64_dpp %3, %4, %8 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1"
        : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3])
        : "v"(x), "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]));
}

// Lower-triangle block t of the column-ordered sequence (0,0), (1,0), ...,
// (NB-1,0), (1,1), ..., (NB-1,NB-1): (row block, column block).
__host__ __device__ constexpr int sym_col(int t, int NB, int J = 0) {
    return t < NB - J ? J : sym_col(t - (NB - J), NB, J + 1);
}
__host__ __device__ constexpr int sym_row(int t, int NB, int J = 0) {
    return t < NB - J ? J + t : sym_row(t - (NB - J), NB, J + 1);
}

// Lane (g, c) item r of lower block (I, J): element (16I + c, 16J + 4g + r),
// read from the lower triangle (the diagonal block is symmetrised).  The
// block base is uniform; ob / od are the lane's byte offsets inside a
// block below / on the diagonal.  Below the diagonal each lane group reads
// 16 consecutive rows of a column.
template <int NB, int T>
__device__ __forceinline__ void sym_load(const double *__restrict__ a, int64_t lda, const uint32_t (&ob)[4],
                                         const uint32_t (&od)[4], double (&v)[4]) {
    constexpr int I = sym_row(T, NB), J = sym_col(T, NB);
    const char *base = reinterpret_cast<const char *>(a + (16 * I + 16 * J * lda));
#pragma unroll
    for (int r = 0; r < 4; ++r)
        v[r] = *reinterpret_cast<const double *>(base + ((I > J) ? ob[r] : od[r]));
}

// RD <- b - A XD for n = 16 NB, reading each block of the lower triangle
// once (FP64, [row][4] layout, zero beyond nr right-hand sides); when
// want_asum, ASUM[i] receives sum_j |A(i,j)|.
//
// Column J of blocks is swept top to bottom.  The diagonal block and, after
// a transpose through the LDS tile TT, every block (I, J) below it add
// A(I,J)^T x_I into the output rows of block J, accumulated per lane and
// reduced over the lane groups once per column.  The block itself adds
// A(I,J) x_J into the rows of block I, reduced at once and added to RD.
template <int NB, bool WANT_ASUM>
__device__ __forceinline__ void residual_sym(const double *__restrict__ a, int64_t lda, const double *XD,
                                             const double *__restrict__ bb, int64_t ldb, int nr, double *RD,
                                             double *ASUM, double *TT, int lane) {
    constexpr bool want_asum = WANT_ASUM;
    constexpr int NP = 16 * NB, NT = NB * (NB + 1) / 2, D = 4;
    const int g = lane >> 4, c = lane & 15;
    uint32_t ob[4], od[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t j = 4 * g + r, l = (uint32_t)lda;
        ob[r] = 8u * (c + j * l);
        od[r] = 8u * ((c > (int)j) ? c + j * l : j + c * l);
    }
    double ring[D][4];
    static_for<0, (D < NT ? D : NT)>([&](auto T) { sym_load<NB, T>(a, lda, ob, od, ring[T]); });
    // RD <- b (rows of this lane: e = lane + 64 m, row e / 4, column e % 4)
    {
        double bv[NP * kRhs / 64];
#pragma unroll
        for (int m = 0; m < NP * kRhs / 64; ++m) {
            const int e = lane + 64 * m, row = e >> 2, q = e & 3;
            bv[m] = bb[row + (int64_t)(q < nr ? q : 0) * ldb];
        }
#pragma unroll
        for (int m = 0; m < NP * kRhs / 64; ++m) {
            const int e = lane + 64 * m;
            RD[e] = ((e & 3) < nr) ? bv[m] : 0.0;
        }
        if (want_asum) {
#pragma unroll
            for (int m = 0; m < (NP + 63) / 64; ++m)
                if (m * 64 + 64 <= NP || lane + 64 * m < NP) ASUM[lane + 64 * m] = 0.0;
        }
    }
    // x_K is read as one element per lane, XD[64 K + lane] (row 16 K + lane / 4,
    // right-hand side lane % 4); lane (g, c) takes x(16 K + 4 g + r, q) from
    // lane 16 g + 4 r + q of its row by DPP broadcast
    const int qg = rs_rhs(g);
    double p[4] = {0.0, 0.0, 0.0, 0.0}, as = 0.0;
    static_for<0, NT>([&](auto T) {
        constexpr int I = sym_row(T, NB), J = sym_col(T, NB);
        double av[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) av[r] = ring[T % D][r];
        if constexpr (T + D < NT) sym_load<NB, T + D>(a, lda, ob, od, ring[T % D]);
        __builtin_amdgcn_sched_barrier(0);
        const double xiv = XD[64 * I + lane];
        double at[4];
        if constexpr (I > J) {
            // at[r] = element (16I + 4g + r, 16J + c): the block transposed
#pragma unroll
            for (int r = 0; r < 4; ++r) TT[c + 17 * (4 * g + r)] = av[r];
#pragma unroll
            for (int r = 0; r < 4; ++r) at[r] = TT[4 * g + r + 17 * c];
            // A(I,J) x_J into rows 16I + c
            const double xjv = XD[64 * J + lane];
            double s[4] = {0.0, 0.0, 0.0, 0.0};
            fma_bcast16(s, av, xjv);
            const int o = 16 * I + c;
            RD[4 * o + qg] -= row_group_reduce_scatter(s);
            // |A| row sums (first residual only): lane group 0 adds to
            // ASUM, the other groups to a sink
            if constexpr (WANT_ASUM) {
                double sa = fabs(av[0]) + fabs(av[1]) + fabs(av[2]) + fabs(av[3]);
                sa = row_group_sum(sa);
                double *dst = (g == 0) ? ASUM + o : TT + kSinkOff + c;
                *dst = *dst + sa;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) at[r] = av[r];
        }
        // A(I,J)^T x_I (the diagonal block itself when I = J) into rows 16J + c
        fma_bcast16(p, at, xiv);
        if constexpr (WANT_ASUM) as += fabs(at[0]) + fabs(at[1]) + fabs(at[2]) + fabs(at[3]);
        if constexpr (I == NB - 1) {
            const int o = 16 * J + c;
            RD[4 * o + qg] -= row_group_reduce_scatter(p);
            if constexpr (WANT_ASUM) {
                as = row_group_sum(as);
                double *dst = (g == 0) ? ASUM + o : TT + kSinkOff + c;
                *dst = *dst + as;
            }
            p[0] = p[1] = p[2] = p[3] = as = 0.0;
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

// m[q] = max over rows of |V(row, q)| for a [row][4] LDS array of NP rows
// (uniform result).
// Column maxima of |V| over the wavefront.  fmax drops NaN operands, so
// the return value says separately whether every entry was finite (one
// ballot over the lanes' flags, uniform across the wavefront).
template <int NP>
__device__ __forceinline__ bool col_absmax(const double *V, int ln, double (&m)[kRhs]) {
    bool bad = false;
#pragma unroll
    for (int q = 0; q < kRhs; ++q) m[q] = 0.0;
#pragma unroll
    for (int row0 = 0; row0 < NP; row0 += 64) {
        const int row = row0 + ln;
        if (row0 + 64 <= NP || row < NP) {
            const double *v = V + 4 * row;
#pragma unroll
            for (int q = 0; q < kRhs; ++q) {
                const double a = fabs(v[q]);
                bad = bad || !(a <= DBL_MAX);  // NaN or infinity
                m[q] = fmax(m[q], a);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kRhs; ++q) m[q] = dpp_wave_max(m[q]);
    return __ballot(bad) == 0;
}

template <int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void posv_mixed_wave_kernel(
    int n, int nrhs, const double *__restrict__ A, int64_t lda, int64_t strideA, const double *__restrict__ B,
    int64_t ldb, int64_t strideB, double *__restrict__ X, int64_t ldx, int64_t strideX, double tol, int max_iter,
    int *iters, int *info, int64_t batch) {
    constexpr int NP = 16 * NB;
    // S: transposes of the solve (256 floats); TS as doubles: the transpose
    // tile of the residual (the two uses never overlap in time)
    __shared__ __attribute__((aligned(16))) double TS[kTileWords];
    float *S = reinterpret_cast<float *>(TS);
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
        // b (zero-padded) into RD: the first right-hand side; x = 0.  Four
        // loads per lane in flight at a time (one at a time, the eight
        // round trips cost as much as a solve)
        static_assert(NP * kRhs == 64 * NB, "b staging: NB entries per lane");
#pragma unroll 1
        for (int t0 = 0; t0 < NB; t0 += 4) {
            double v[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int e = lane + 64 * (t0 + g), row = e >> 2, q = e & 3;
                const bool in = t0 + g < NB && row < n && q < nr;
                v[g] = bb[in ? row + (int64_t)q * ldb : 0];
            }
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int e = lane + 64 * (t0 + g), row = e >> 2, q = e & 3;
                if (t0 + g < NB) {
                    RD[e] = (row < n && q < nr) ? v[g] : 0.0;
                    XD[e] = 0.0;
                }
            }
        }
        double bn[kRhs];
        col_absmax<NP>(RD, lane, bn);
        int it = 0;
        bool converged = false;
        for (;;) {
            asm volatile("" ::: "memory");
            const int ln = opaque_lane();
            // r into the vector-block layout: each lane reads one element
            // per block, RD[64 I + lane], and item r of lane (g, c < 4),
            // RD[64 I + 16 g + 4 r + c], comes from lane 16 g + 4 r + c of
            // its row by a row shift of 4 r lanes written into the row's
            // first four lanes only (DPP bank mask; the other lanes are the
            // zero padding of the right-hand sides)
            f4 Y[NB];
            static_for<0, NB>([&](auto Ic) {
                const int v = __float_as_int((float)RD[64 * Ic + ln]);
                Y[Ic][0] = __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0xE4, 0xF, 0x1, false));   // own lane
                Y[Ic][1] = __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0x1, false));  // row_shl:4
                Y[Ic][2] = __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x108, 0xF, 0x1, false));  // row_shl:8
                Y[Ic][3] = __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x10C, 0xF, 0x1, false));  // row_shl:12
            });
            solve<NB>(U, DS, Y, S, ln);
            // x += d back in the element-per-lane layout: lane 16 g + 4 r + c
            // takes item r of lane (g, c) by a row shift right of 4 r lanes
            // written into bank r of the row
            static_for<0, NB>([&](auto Ic) {
                int y = __builtin_amdgcn_update_dpp(0, __float_as_int(Y[Ic][0]), 0xE4, 0xF, 0x1, false);
                y = __builtin_amdgcn_update_dpp(y, __float_as_int(Y[Ic][1]), 0x114, 0xF, 0x2, false);  // row_shr:4
                y = __builtin_amdgcn_update_dpp(y, __float_as_int(Y[Ic][2]), 0x118, 0xF, 0x4, false);  // row_shr:8
                y = __builtin_amdgcn_update_dpp(y, __float_as_int(Y[Ic][3]), 0x11C, 0xF, 0x8, false);  // row_shr:12
                XD[64 * Ic + ln] += (double)__int_as_float(y);
            });
            // orders the float accesses of the solve before the double
            // accesses of the residual to the same LDS words
            asm volatile("" ::: "memory");
            if (n == 16 * NB && lda <= kSymMaxLda) {
                // the |A| row sums are formed by the first residual only
                if (anorm == 0.0) residual_sym<NB, true>(a, lda, XD, bb, ldb, nr, RD, ASUM, TS, ln);
                else residual_sym<NB, false>(a, lda, XD, bb, ldb, nr, RD, ASUM, TS, ln);
            } else {
                residual_padded<NB>(a, lda, n, XD, bb, ldb, nr, RD, ASUM, anorm == 0.0, ln);
            }
            if (anorm == 0.0) {
                double m = 0.0;
#pragma unroll
                for (int row0 = 0; row0 < NP; row0 += 64)
                    if (row0 + 64 <= NP || row0 + ln < NP) m = fmax(m, ASUM[row0 + ln]);
                anorm = dpp_wave_max(m);
                if (anorm == 0.0) anorm = 1e-300;  // A = 0 is not SPD; keep the test finite
            }
            double rm[kRhs], xm[kRhs];
            // a residual or iterate that is not finite (overflow of the FP32
            // solve, NaN or infinity in A or b) never passes the test: the
            // system goes to the FP64 fallback, as LAPACK DSPOSV does
            const bool finite_r = col_absmax<NP>(RD, ln, rm);
            const bool finite_x = col_absmax<NP>(XD, ln, xm);
            const bool finite = finite_r && finite_x;
            bool ok = finite && anorm <= DBL_MAX;
#pragma unroll
            for (int q = 0; q < kRhs; ++q)
                if (q < nr) ok = ok && (rm[q] <= tol * (anorm * xm[q] + bn[q]));
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
