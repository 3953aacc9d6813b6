// householder.hpp -- wavefront-level Householder building blocks shared by
// the QR / least-squares and SVD kernels (LAPACK dlarfg / dlarf conventions:
// beta = -sign(alpha)*||[alpha; x]||, tau = (beta - alpha)/beta,
// v = [1; x / (alpha - beta)]).  The factorisation spreads the columns over
// the workgroup's wavefronts (four unless a kernel asks for another count).
#pragma once

#include "common.hpp"

namespace gputils {
namespace hh {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;

// Order one wavefront's earlier memory writes before its later reads by
// other lanes (LDS or global): a wavefront-scope fence plus a scheduling
// barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The scalars of a reflector from alpha and the squared norm sq > 0 of the
// rest of the column: beta = -sign(alpha) sqrt(alpha^2 + sq),
// tau = (beta - alpha) / beta, scale = 1 / (alpha - beta).  When the
// factors are stored in FP32 but the scalars are formed in FP64, float
// accuracy is all that survives: the square root is the hardware reciprocal
// square root with one Newton step, and one reciprocal of
// d beta (d = alpha - beta, |d| = |alpha| + norm, so no cancellation; the
// reciprocal with one Newton step) gives tau = -d^2 / (d beta) and
// 1 / d = beta / (d beta).  Otherwise the square root and the divisions
// are the exact ones.
template <typename T, typename ACC>
__device__ __forceinline__ void reflector_scalars(ACC alpha, ACC sq, ACC &beta, ACC &tau, ACC &scale) {
    if constexpr (sizeof(T) < sizeof(ACC)) {
        const ACC s2 = alpha * alpha + sq;
        ACC r = __builtin_amdgcn_rsq(s2);
        r = r * fma(ACC(-0.5) * s2, r * r, ACC(1.5));
        const ACC norm = s2 * r;
        beta = (alpha >= ACC(0)) ? -norm : norm;
        const ACC d = alpha - beta;
        const ACC db = d * beta;
        ACC inv = __builtin_amdgcn_rcp(db);
        inv = inv * fma(-db, inv, ACC(2));
        tau = -d * d * inv;
        scale = beta * inv;
    } else {
        const ACC norm = sqrt(alpha * alpha + sq);
        beta = (alpha >= ACC(0)) ? -norm : norm;
        tau = (beta - alpha) / beta;
        scale = ACC(1) / (alpha - beta);
    }
}

// Reflector of column j (one wavefront).  Scales v in place, writes beta to
// A[j][j] and returns tau (identical in every lane).
template <typename T, typename ACC>
__device__ __forceinline__ T make_reflector(T *A, int64_t lda, int m, int j) {
    const int lane = threadIdx.x & 63;
    T *col = A + (int64_t)j * lda;
    ACC s = ACC(0);
    for (int i = j + 1 + lane; i < m; i += kWave) {
        const ACC x = (ACC)col[i];
        s += x * x;
    }
    s = wave_sum(s);
    const ACC alpha = (ACC)col[j];
    if (s == ACC(0)) return T(0);  // H = I (dlarfg with x = 0)
    ACC beta, tau, sc;
    reflector_scalars<T, ACC>(alpha, s, beta, tau, sc);
    const T scale = (T)sc;
    for (int i = j + 1 + lane; i < m; i += kWave) col[i] *= scale;
    wave_sync();
    if (lane == 0) col[j] = (T)beta;
    return (T)tau;
}

// The lookahead step of the owner of column j + 1, fused: reflector j
// (tau) is applied to column j + 1 and reflector j + 1 is formed from the
// result while the column stays in registers (rows j + 1 + lane + 64 u,
// u < R, so m - j - 1 <= 64 R), reading the column and v once and writing
// the column once.  The arithmetic is apply_reflector_col followed by
// make_reflector.  Returns tau_{j+1}.
template <typename T, typename ACC, typename DOT, int R>
__device__ __forceinline__ T update_and_reflect(T *A, int64_t lda, int m, int j, T tau) {
    const int lane = threadIdx.x & 63;
    const T *v = A + (int64_t)j * lda;
    T *x = A + (int64_t)(j + 1) * lda;
    T vr[R], xr[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int i = j + 1 + lane + kWave * u;
        const int ic = i < m ? i : m - 1;  // clamped: every load is valid
        vr[u] = v[ic];
        xr[u] = x[ic];
        if (i >= m) vr[u] = xr[u] = T(0);
    }
    const T xj = x[j];
    DOT w = DOT(0);
#pragma unroll
    for (int u = 0; u < R; ++u) w += (DOT)vr[u] * (DOT)xr[u];
    w = wave_sum(w) + (DOT)xj;
    const T tw = (T)((DOT)tau * w);
#pragma unroll
    for (int u = 0; u < R; ++u) xr[u] -= tw * vr[u];
    // reflector j + 1: alpha = row j + 1 (lane 0, u = 0), x = rows j + 2 ..
    ACC sq = ACC(0);
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const ACC e = (ACC)xr[u];
        sq += (u == 0 && lane == 0) ? ACC(0) : e * e;
    }
    sq = wave_sum(sq);
    const ACC alpha = (ACC)readlane(xr[0], 0);
    T t1 = T(0);
    if (sq != ACC(0)) {  // else H = I (dlarfg with x = 0): the column stays as it is
        ACC beta, tau, sc;
        reflector_scalars<T, ACC>(alpha, sq, beta, tau, sc);
        t1 = (T)tau;
        const T scale = (T)sc;
#pragma unroll
        for (int u = 0; u < R; ++u) xr[u] = (u == 0 && lane == 0) ? (T)beta : xr[u] * scale;
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int i = j + 1 + lane + kWave * u;
        if (i < m) x[i] = xr[u];
    }
    if (lane == 0) x[j] = xj - tw;
    wave_sync();
    return t1;
}

// x <- (I - tau v v^T) x on rows j..m-1 of column c, v = [1; A[j+1:m, j]]
// (one wavefront).
template <typename T, typename DOT>
__device__ __forceinline__ void apply_reflector_col(T *A, int64_t lda, int m, int j, T tau, int c) {
    const int lane = threadIdx.x & 63;
    const T *v = A + (int64_t)j * lda;
    T *x = A + (int64_t)c * lda;
    DOT w = DOT(0);
    for (int i = j + 1 + lane; i < m; i += kWave) w += (DOT)v[i] * (DOT)x[i];
    w = wave_sum(w) + (DOT)x[j];
    const T tw = (T)((DOT)tau * w);
    for (int i = j + 1 + lane; i < m; i += kWave) x[i] -= tw * v[i];
    wave_sync();
    if (lane == 0) x[j] -= tw;
}

// x + (x of the lane that DPP control CTRL selects), within 16-lane rows
template <int CTRL>
__device__ __forceinline__ float row16_plus(float x) {
    return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double row16_plus(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return x + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over each aligned group of 16 lanes, in every lane of the group:
// quad permutations [1,0,3,2] and [2,3,0,1], then the row half mirror and
// the row mirror.  Register moves only.
template <typename T>
__device__ __forceinline__ T row16_sum(T x) {
    x = row16_plus<0xB1>(x);   // quad_perm [1,0,3,2]
    x = row16_plus<0x4E>(x);   // quad_perm [2,3,0,1]
    x = row16_plus<0x141>(x);  // row_half_mirror
    return row16_plus<0x140>(x);  // row_mirror
}

// x_c <- (I - tau v v^T) x_c for the columns c = c0, c0 + NW, ... < n (the
// columns one of NW wavefronts owns), v = [1; A[j+1:m, j]], for
// m - j - 1 <= 16 KR: each 16-lane group of the wavefront takes one column
// (four at a time), lane l of a group rows j + 1 + l + 16 k (k < KR, v held
// in registers for all the columns), and reduces its dot product within
// its 16 lanes.  Four columns then cost one set of row loads and one
// four-step reduction, where the 64-lane form below pays a six-step
// reduction per column.
// KR < 0: any m, nothing held in registers (two passes over the rows, v
// read with each; for kernels that cannot spare the registers).
template <typename T, typename DOT, int NW, int KR>
__device__ __forceinline__ void apply_reflector_cols16(T *A, int64_t lda, int m, int j, T tau, int c0, int n) {
    const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
    const T *v = A + (int64_t)j * lda;
    if constexpr (KR < 0) {
        for (int cq = c0; cq < n; cq += 4 * NW) {
            const int c = cq + NW * g;
            const bool ok = c < n;
            T *x = A + (int64_t)(ok ? c : cq) * lda;
            DOT w = DOT(0);
#pragma unroll 4
            for (int i = j + 1 + l16; i < m; i += 16) w += (DOT)v[i] * (DOT)x[i];
            const T xj = x[j];
            w = row16_sum(w) + (DOT)xj;
            const T tw = ok ? (T)((DOT)tau * w) : T(0);
            if (ok) {
#pragma unroll 4
                for (int i = j + 1 + l16; i < m; i += 16) x[i] -= tw * v[i];
                if (l16 == 0) x[j] = xj - tw;
            }
        }
        return;
    }
    constexpr int KRH = KR > 0 ? KR : 1;
    const int kr = (m - j - 1 + 15) >> 4;  // row groups in use (uniform)
    T vr[KRH];
#pragma unroll
    for (int k = 0; k < KRH; ++k) {
        const int i = j + 1 + l16 + 16 * k;
        vr[k] = T(0);
        if (k < kr) {
            const T e = v[i < m ? i : m - 1];  // clamped: every load is valid
            vr[k] = (i < m) ? e : T(0);
        }
    }
    for (int cq = c0; cq < n; cq += 4 * NW) {
        const int c = cq + NW * g;
        const bool ok = c < n;
        T *x = A + (int64_t)(ok ? c : cq) * lda;
        T xr[KRH];
        DOT w = DOT(0);
#pragma unroll
        for (int k = 0; k < KRH; ++k) {
            xr[k] = T(0);
            if (k < kr) {
                const int i = j + 1 + l16 + 16 * k;
                xr[k] = x[i < m ? i : m - 1];
                w += (DOT)vr[k] * (DOT)xr[k];  // vr is zero past m
            }
        }
        const T xj = x[j];
        w = row16_sum(w) + (DOT)xj;
        const T tw = ok ? (T)((DOT)tau * w) : T(0);
#pragma unroll
        for (int k = 0; k < KRH; ++k) {
            if (k < kr) {
                const int i = j + 1 + l16 + 16 * k;
                if (ok && i < m) x[i] = xr[k] - tw * vr[k];
            }
        }
        if (ok && l16 == 0) x[j] = xj - tw;
    }
    // the caller's barrier orders these writes before other wavefronts' reads
}

// x_c <- (I - tau v v^T) x_c for the columns c = c0, c0 + NW, ... < n (the
// columns one of NW wavefronts owns), v = [1; A[j+1:m, j]], four columns at a
// time: their dot products are accumulated together and their reductions
// overlap, so the wavefront waits for one set of reductions per four
// columns instead of one per column.  Past the last column the slots work
// on the first owned column with tau w = 0.
template <typename T, typename DOT, int NW = kWaves>
__device__ __forceinline__ void apply_reflector_cols(T *A, int64_t lda, int m, int j, T tau, int c0, int n) {
    const int lane = threadIdx.x & 63;
    const T *v = A + (int64_t)j * lda;
    for (int cq = c0; cq < n; cq += 4 * NW) {
        T *x[4];
        bool ok[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cq + NW * q;
            ok[q] = c < n;
            x[q] = A + (int64_t)(ok[q] ? c : cq) * lda;
        }
        DOT w[4] = {DOT(0), DOT(0), DOT(0), DOT(0)};
        for (int i = j + 1 + lane; i < m; i += kWave) {
            const DOT vi = (DOT)v[i];
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] += vi * (DOT)x[q][i];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = wave_sum(w[q]);
        T tw[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) tw[q] = ok[q] ? (T)((DOT)tau * (w[q] + (DOT)x[q][j])) : T(0);
        for (int i = j + 1 + lane; i < m; i += kWave) {
            const T vi = v[i];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (ok[q]) x[q][i] -= tw[q] * vi;
        }
        wave_sync();
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (ok[q]) x[q][j] -= tw[q];
        }
        wave_sync();
    }
}

// Householder QR of the m x n (m >= n) matrix at A (LDS or global), tau[j]
// for j < n, column norms accumulated in ACC and the reflectors' dot
// products in DOT.  s_tau: two LDS slots.  Ends with a barrier.  Columns n ..
// ntot-1 (stored after the factorised ones) receive every reflector
// without being factorised: a right-hand side stored there ends as Q^T b.
// NW (a power of two) wavefronts: column c is owned by wavefront c % NW.
// Each column's arithmetic is the same for every NW.  KR16 != 0 applies the
// reflectors with 16-lane groups (apply_reflector_cols16; KR16 > 0 requires
// m <= 16 KR16 + 1).
template <typename T, typename ACC, typename DOT = ACC, int NW = kWaves, int KR16 = 0>
__device__ __forceinline__ void hqr_factor(T *A, int64_t lda, int m, int n, T *tau, T *s_tau, int ntot = 0) {
    static_assert(NW >= 1 && (NW & (NW - 1)) == 0, "NW must be a power of two");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (n <= 0) return;
    if (ntot < n) ntot = n;
    if (wave == 0) {
        const T t = make_reflector<T, ACC>(A, lda, m, 0);
        if (lane == 0) {
            s_tau[0] = t;
            tau[0] = t;
        }
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const T t = s_tau[j & 1];
        if (j + 1 < ntot && wave == ((j + 1) & (NW - 1))) {
            if (j + 1 < n && m <= 8 * kWave) {
                // as many row groups as the column has (the terms past m
                // are zeros either way)
                const T t1 = (m <= 2 * kWave)   ? update_and_reflect<T, ACC, DOT, 2>(A, lda, m, j, t)
                             : (m <= 4 * kWave) ? update_and_reflect<T, ACC, DOT, 4>(A, lda, m, j, t)
                                                : update_and_reflect<T, ACC, DOT, 8>(A, lda, m, j, t);
                if (lane == 0) {
                    s_tau[(j + 1) & 1] = t1;
                    tau[j + 1] = t1;
                }
            } else {
                apply_reflector_col<T, DOT>(A, lda, m, j, t, j + 1);
                if (j + 1 < n) {
                    const T t1 = make_reflector<T, ACC>(A, lda, m, j + 1);
                    if (lane == 0) {
                        s_tau[(j + 1) & 1] = t1;
                        tau[j + 1] = t1;
                    }
                }
            }
        }
        if (t != T(0)) {
            const int c0 = j + 2 + ((wave - (j + 2)) & (NW - 1));
            if constexpr (KR16 != 0) apply_reflector_cols16<T, DOT, NW, KR16>(A, lda, m, j, t, c0, ntot);
            else apply_reflector_cols<T, DOT, NW>(A, lda, m, j, t, c0, ntot);
        }
        __syncthreads();
    }
}

// c <- Q^T c (trans) or Q c for one m-vector in LDS, one wavefront; Q from
// the first k reflectors stored in A / tau.
template <typename T, typename ACC>
__device__ __forceinline__ void apply_q_vec(const T *A, int64_t lda, int m, int k, const T *tau, T *c,
                                            bool trans) {
    const int lane = threadIdx.x & 63;
    for (int t = 0; t < k; ++t) {
        const int j = trans ? t : (k - 1 - t);
        const T tj = tau[j];
        if (tj == T(0)) continue;
        const T *v = A + (int64_t)j * lda;
        ACC w = ACC(0);
        for (int i = j + 1 + lane; i < m; i += kWave) w += (ACC)v[i] * (ACC)c[i];
        w = wave_sum(w) + (ACC)c[j];
        const T tw = (T)((ACC)tj * w);
        for (int i = j + 1 + lane; i < m; i += kWave) c[i] -= tw * v[i];
        wave_sync();
        if (lane == 0) c[j] -= tw;
        wave_sync();
    }
}

// Solve R y = c in place (R upper triangular n x n at A, c in LDS), one
// wavefront, column-oriented back substitution.
template <typename T>
__device__ __forceinline__ void trsv_upper(const T *A, int64_t lda, int n, T *c) {
    const int lane = threadIdx.x & 63;
    for (int k = n - 1; k >= 0; --k) {
        const T ck = c[k] / A[k + (int64_t)k * lda];
        for (int i = lane; i < k; i += kWave) c[i] -= A[i + (int64_t)k * lda] * ck;
        wave_sync();
        if (lane == 0) c[k] = ck;
        wave_sync();
    }
}

}  // namespace hh
}  // namespace gputils
