// qr.hip -- batched Householder QR, Q^T application, triangular solve and
// least squares for gfx950.
//
// Replaces cublas?gelsBatched behind DTensor::leastSquaresBatched
// (include/tensor.cuh:1340-1394) and cusolverDn?geqrf / ?ormqr / cublas?trsm
// behind QRFactoriser (tensor.cuh:1796-1995), and adds a ragged
// mixed-precision least-squares solver (uneven m_i x n_i systems in one
// plan: FP32 Householder QR with FP64 norms and dot products plus one FP64
// refinement step).
//
// One 256-thread workgroup per matrix.  The matrix lives in LDS when it
// fits (otherwise the kernel works on it in global memory).  Column c is
// owned by wavefront c % 4.  Step j: every wave applies reflector j to the
// columns it owns; the owner of column j+1 does that column first and then
// forms reflector j+1 (dlarfg convention: beta = -sign(alpha)*||x||,
// tau = (beta - alpha)/beta, v = x / (alpha - beta)), so one barrier per
// column suffices.
#include <algorithm>
#include <queue>
#include <utility>
#include <vector>

#include "common.hpp"
#include "householder.hpp"

namespace gputils {
namespace qr {

using namespace hh;

template <typename T>
struct Acc;
template <>
struct Acc<float> {
    using type = float;
};
template <>
struct Acc<double> {
    using type = double;
};

// ------------------------------------------------------------- kernels --
// Uniform batched least squares (gelsBatched): A_b <- QR factors,
// B_b[0:n, :] <- x.  MAXE > 0: matrix staged in LDS (m*n <= MAXE).
template <typename T, int MAXE>
__global__ __launch_bounds__(kThreads) void gels_kernel(int m, int n, int nrhs, T *A, int64_t lda,
                                                        int64_t strideA, T *B, int64_t ldb,
                                                        int64_t strideB, T *tauws, int *info,
                                                        int64_t batch) {
    using ACC = typename Acc<T>::type;
    constexpr int kLds = MAXE > 0 ? MAXE : 1;
    __shared__ T S[kLds];
    __shared__ T s_tau[2];
    extern __shared__ unsigned char dyn[];  // tau (n) and one right-hand side (m)
    T *tau = reinterpret_cast<T *>(dyn);
    T *c = tau + n;
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    T *a = A + b * strideA;
    T *W = (MAXE > 0) ? S : a;
    const int64_t ldw = (MAXE > 0) ? m : lda;
    if (MAXE > 0) {
        for (int e = threadIdx.x; e < m * n; e += kThreads) S[e] = a[(e % m) + (int64_t)(e / m) * lda];
        __syncthreads();
    }
    hqr_factor<T, ACC>(W, ldw, m, n, tau, s_tau);
    if (MAXE > 0) {
        for (int e = threadIdx.x; e < m * n; e += kThreads) a[(e % m) + (int64_t)(e / m) * lda] = S[e];
    }
    const int wave = threadIdx.x >> 6;
    T *bb = B + b * strideB;
    for (int r = 0; r < nrhs; ++r) {
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += kThreads) c[i] = bb[i + (int64_t)r * ldb];
        __syncthreads();
        if (wave == 0) {
            apply_q_vec<T, ACC>(W, ldw, m, n, tau, c, true);
            trsv_upper<T>(W, ldw, n, c);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kThreads) bb[i + (int64_t)r * ldb] = c[i];
    }
    if (tauws != nullptr)
        for (int i = threadIdx.x; i < n; i += kThreads) tauws[b * n + i] = tau[i];
    if (threadIdx.x == 0 && info != nullptr) info[b] = 0;
}

// Batched QR factorisation only (geqrf): A_b <- QR factors, tau_b.
template <typename T, int MAXE>
__global__ __launch_bounds__(kThreads) void geqrf_kernel(int m, int n, T *A, int64_t lda, int64_t strideA,
                                                         T *tau, int64_t strideTau, int *info,
                                                         int64_t batch) {
    using ACC = typename Acc<T>::type;
    constexpr int kLds = MAXE > 0 ? MAXE : 1;
    __shared__ T S[kLds];
    __shared__ T s_tau[2];
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    T *a = A + b * strideA;
    T *W = (MAXE > 0) ? S : a;
    const int64_t ldw = (MAXE > 0) ? m : lda;
    if (MAXE > 0) {
        for (int e = threadIdx.x; e < m * n; e += kThreads) S[e] = a[(e % m) + (int64_t)(e / m) * lda];
        __syncthreads();
    }
    hqr_factor<T, ACC>(W, ldw, m, n, tau + b * strideTau, s_tau);
    if (MAXE > 0)
        for (int e = threadIdx.x; e < m * n; e += kThreads) a[(e % m) + (int64_t)(e / m) * lda] = S[e];
    if (threadIdx.x == 0 && info != nullptr) info[b] = 0;
}

// C_b <- Q_b^T C_b (trans) or Q_b C_b; each wavefront takes columns of C.
template <typename T>
__global__ __launch_bounds__(kThreads) void ormqr_kernel(int trans, int m, int ncols, int k, const T *A,
                                                         int64_t lda, int64_t strideA, const T *tau,
                                                         int64_t strideTau, T *C, int64_t ldc,
                                                         int64_t strideC, int64_t batch) {
    using ACC = typename Acc<T>::type;
    extern __shared__ unsigned char dyn[];
    T *cw = reinterpret_cast<T *>(dyn) + (threadIdx.x >> 6) * m;  // one column per wave
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const T *a = A + b * strideA;
    const T *t = tau + b * strideTau;
    T *cb = C + b * strideC;
    for (int col = wave; col < ncols; col += kWaves) {
        for (int i = lane; i < m; i += kWave) cw[i] = cb[i + (int64_t)col * ldc];
        wave_sync();
        apply_q_vec<T, ACC>(a, lda, m, k, t, cw, trans != 0);
        for (int i = lane; i < m; i += kWave) cb[i + (int64_t)col * ldc] = cw[i];
        wave_sync();
    }
}

// B_b <- R_b^{-1} B_b, R upper triangular n x n (non-unit).
template <typename T>
__global__ __launch_bounds__(kThreads) void trsm_upper_kernel(int n, int nrhs, const T *R, int64_t ldr,
                                                              int64_t strideR, T *B, int64_t ldb,
                                                              int64_t strideB, int64_t batch) {
    extern __shared__ unsigned char dyn[];
    T *cw = reinterpret_cast<T *>(dyn) + (threadIdx.x >> 6) * n;
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const T *r = R + b * strideR;
    T *bb = B + b * strideB;
    for (int col = wave; col < nrhs; col += kWaves) {
        for (int i = lane; i < n; i += kWave) cw[i] = bb[i + (int64_t)col * ldb];
        wave_sync();
        trsv_upper<T>(r, ldr, n, cw);
        for (int i = lane; i < n; i += kWave) bb[i + (int64_t)col * ldb] = cw[i];
        wave_sync();
    }
}

// -------------------------------------------------- ragged, mixed --
struct RaggedItem {
    int m, n;
    int64_t offA, offB;
};

// FP64 r_i = base_i - sum_k A[i][k] x[k] for the rows i of this thread
// (row-major over the NT threads, coalesced over rows); the loads of eight
// columns are issued together so that they are in flight at once.
template <int NT>
__device__ __forceinline__ void rows_minus_ax(const double *__restrict__ a, int m, int n, const double *x,
                                              const double *base0, const double *base1, double *out) {
    for (int i = threadIdx.x; i < m; i += NT) {
        double r = base0[i] - (base1 ? base1[i] : 0.0);
        int k = 0;
        for (; k + 8 <= n; k += 8) {
            double av[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) av[u] = a[i + (int64_t)(k + u) * m];
#pragma unroll
            for (int u = 0; u < 8; ++u) r = fma(-av[u], x[k + u], r);
        }
        for (; k < n; ++k) r = fma(-a[i + (int64_t)k * m], x[k], r);
        out[i] = r;
    }
}

// Panel-blocked, register-resident FP32 Householder QR of the m x (n + 1)
// matrix [A b] for the ragged solver (A FP64 column-major with ld m, b an
// FP64 m-vector; m <= 512, n <= 64), dlarfg conventions, column norms in
// FP64 and reflector dot products in FP32.
//
// Columns are grouped in panels of NB = 64 / NW; panel P is held by
// wavefront P % NW in registers (rows lane + 64 u, u < 8) and b by the
// wavefront after the last panel's.  Iteration p: every wavefront applies
// the reflectors of panel p (read from S, where their owner stored them) to
// its columns right of the panel, one reflector at a time: w = v . x,
// x -= (tau w) v, with v = [0 .. 0, 1, v_{j+1} .. v_{m-1}] and zeros past
// m; then the owner of panel p + 1 factorises it in its registers alone
// (no barrier, no LDS traffic inside the panel) and stores it to S.  One
// barrier per panel.  Each column's arithmetic is the same for every NW.
// At the end S (column-major, ld m) holds R on and above the diagonal and
// v below it, column n holds Q^T b, and tau[0:n] the scalars.
template <int NW>
__device__ __forceinline__ void hqr_panels(const double *__restrict__ a, const double *__restrict__ bcol, float *S,
                                           int m, int n, float *tau) {
    constexpr int NB = 64 / NW;  // columns per panel; one panel per wavefront
    constexpr int CM = NB + 1;   // register columns: the panel and maybe b
    constexpr int R = 8;         // row groups (m <= 512)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ru = (m + 63) >> 6;
    const int np = (n + NB - 1) / NB;
    const bool own_b = wave == (np % NW);
    // column of register slot q (q < NB: panel `wave`; q == NB: b)
    auto col_of = [&](int q) { return (q < NB) ? wave * NB + q : n; };
    auto valid = [&](int q) { return (q < NB) ? (wave * NB + q < n) : own_b; };
    // every lane loads a valid element (row and column clamped into the
    // system) so that a column's eight loads are unconditional and in
    // flight together; the padding is zeroed afterwards.  The compiler
    // barrier per column keeps the doubles in flight to one column's worth
    // of registers.
    float x[CM][R];
#pragma unroll
    for (int q = 0; q < CM; ++q) {
        const bool ok = valid(q);
        const double *src = (q < NB) ? a + (int64_t)(ok ? col_of(q) : 0) * m : bcol;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int i = lane + 64 * u;
            x[q][u] = (float)src[i < m ? i : m - 1];
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int i = lane + 64 * u;
            if (!(ok && i < m)) x[q][u] = 0.0f;
        }
        asm volatile("" ::: "memory");
    }
    // reflector of column cc from its registers; col becomes [R part;
    // beta; v], vn the full reflector vector; returns tau
    auto reflect = [&](float (&col)[R], int cc, float (&vn)[R]) -> float {
        const int uc = cc >> 6;
        float asel = col[0];
#pragma unroll
        for (int u = 1; u < R; ++u)
            if (u == uc) asel = col[u];
        const double alpha = (double)readlane(asel, cc & 63);
        double sq = 0.0;  // rows past m hold zeros
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int i = lane + 64 * u;
            const double e = (double)col[u];
            sq += (i > cc) ? e * e : 0.0;
        }
        sq = wave_sum(sq);
        float t1 = 0.0f;
        if (sq != 0.0) {
            double beta, tau_d, sc;
            reflector_scalars<float, double>(alpha, sq, beta, tau_d, sc);
            t1 = (float)tau_d;
            const float scale = (float)sc;
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int i = lane + 64 * u;
                col[u] = (i > cc) ? col[u] * scale : ((i == cc) ? (float)beta : col[u]);
                vn[u] = (i > cc) ? col[u] : ((i == cc) ? 1.0f : 0.0f);
            }
        } else {  // H = I (dlarfg with x = 0): the column stays as it is
#pragma unroll
            for (int u = 0; u < R; ++u) vn[u] = (lane + 64 * u == cc) ? 1.0f : 0.0f;
        }
        return t1;
    };
    // x_q -= (t v . x_q) v for the slots with sel(q)
    auto apply = [&](const float (&vr)[R], float t, auto sel) {
        float w[CM];
#pragma unroll
        for (int q = 0; q < CM; ++q) {
            w[q] = 0.0f;
            if (sel(q)) {
#pragma unroll
                for (int u = 0; u < R; ++u)
                    if (u < ru) w[q] = fmaf(vr[u], x[q][u], w[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < CM; ++q)
            if (sel(q)) w[q] = wave_sum(w[q]);
#pragma unroll
        for (int q = 0; q < CM; ++q) {
            if (sel(q)) {
                const float tw = t * w[q];
#pragma unroll
                for (int u = 0; u < R; ++u)
                    if (u < ru) x[q][u] = fmaf(-tw, vr[u], x[q][u]);
            }
        }
    };
    // factorise this wavefront's panel (columns wave * NB ..) and store it
    auto factor_panel = [&]() {
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int c = wave * NB + k;
            if (c < n) {
                float vn[R];
                const float t1 = reflect(x[k], c, vn);
                if (lane == 0) tau[c] = t1;
                apply(vn, t1, [&](int q) { return q > k && q < NB && wave * NB + q < n; });
            }
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int c = wave * NB + k;
            if (c < n) {
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    const int i = lane + 64 * u;
                    if (i < m) S[i + c * m] = x[k][u];
                }
            }
        }
    };
    if (wave == 0) factor_panel();
    __syncthreads();
    for (int p = 0; p < np; ++p) {
        const int j0 = p * NB, j1 = min(j0 + NB, n);
        // panel p's reflectors on every column right of it (rows past m of
        // a column of S lie inside S: j m + 512 <= m n + 512)
        if (wave > p || own_b) {
            for (int j = j0; j < j1; ++j) {
                const float *colv = S + j * m;
                float vr[R];
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    const int i = lane + 64 * u;
                    const float e = (u < ru) ? colv[i] : 0.0f;
                    vr[u] = (i > j && i < m) ? e : ((i == j) ? 1.0f : 0.0f);
                }
                apply(vr, tau[j], [&](int q) { return valid(q) && col_of(q) >= j1; });
            }
        }
        if (wave == p + 1 && p + 1 < np) factor_panel();
        __syncthreads();
    }
    if (own_b) {
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int i = lane + 64 * u;
            if (i < m) S[i + n * m] = x[NB][u];
        }
    }
    __syncthreads();
}

// The one-wavefront vector pieces of the ragged solver keep their vector
// in registers, so each step costs one batch of LDS reads of the factors
// plus one reduction or one pivot broadcast instead of a chain of LDS
// round trips.

// c <- Q^T c (trans) or Q c for one m-vector in LDS (m <= 512), one
// wavefront; Q from the n reflectors stored below the diagonal of S (ld m)
// and tau.  Rows lane + 64 u of c are held in registers; reflector j is
// applied as w = v . c, c -= (tau_j w) v with v = [0 .. 0, 1, S[j+1:m, j]]
// (its eight loads per lane issued together; reading them a step ahead
// would cost the kernel 28 more VGPRs and a wavefront per SIMD).
__device__ __forceinline__ void apply_q_regs(const float *S, int m, int n, const float *tau, float *c, bool trans) {
    constexpr int R = 8;
    const int lane = threadIdx.x & 63;
    const int ru = (m + 63) >> 6;
    float cr[R], vc[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int i = lane + 64 * u;
        cr[u] = (u < ru && i < m) ? c[i] : 0.0f;
    }
    // rows past m of column j lie inside S (j m + 512 <= m n + 512), so the
    // loads need no per-lane guard; the select masks them
    auto load_v = [&](int j, float *vr) {
        const float *col = S + j * m;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int i = lane + 64 * u;
            const float e = (u < ru) ? col[i] : 0.0f;
            vr[u] = (i > j && i < m) ? e : ((i == j) ? 1.0f : 0.0f);
        }
    };
    const int dj = trans ? 1 : -1;
    int j = trans ? 0 : n - 1;
#pragma unroll 1
    for (int t = 0; t < n; ++t, j += dj) {
        load_v(j, vc);
        float w = 0.0f;
#pragma unroll
        for (int u = 0; u < R; ++u)
            if (u < ru) w = fmaf(vc[u], cr[u], w);
        w = wave_sum(w);
        const float tw = tau[j] * w;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            if (u < ru) cr[u] = fmaf(-tw, vc[u], cr[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int i = lane + 64 * u;
        if (u < ru && i < m) c[i] = cr[u];
    }
}

// Solve R y = c in place (R the n x n upper triangle of S, ld m; c in LDS,
// n <= 64), one wavefront: lane i holds c_i and 1 / R(i, i) (formed for all
// rows at once, so that a step multiplies instead of dividing on the
// chain), column k - 1 of R is read while step k runs.
__device__ __forceinline__ void trsv_upper_regs(const float *S, int m, int n, float *c) {
    const int lane = threadIdx.x & 63;
    float cv = (lane < n) ? c[lane] : 0.0f;
    const float dinv = (lane < n) ? 1.0f / S[lane + (int64_t)lane * m] : 0.0f;
    float rk = (lane < n) ? S[lane + (int64_t)(n - 1) * m] : 0.0f;
#pragma unroll 1
    for (int k = n - 1; k >= 0; --k) {
        const float rn = (lane < k) ? S[lane + (int64_t)(k - 1) * m] : 0.0f;  // rows 0..k-1 of column k-1
        const float ck = readlane(cv, k) * readlane(dinv, k);
        cv = (lane == k) ? ck : ((lane < k) ? fmaf(-rk, ck, cv) : cv);
        rk = rn;
    }
    if (lane < n) c[lane] = cv;
}

// Solve R^T y = c in place (forward substitution), one wavefront: lane i
// holds c_i and 1 / R(i, i), row k + 1 of R is read while step k runs.
__device__ __forceinline__ void trsv_upper_trans_regs(const float *S, int m, int n, float *c) {
    const int lane = threadIdx.x & 63;
    float cv = (lane < n) ? c[lane] : 0.0f;
    const float dinv = (lane < n) ? 1.0f / S[lane + (int64_t)lane * m] : 0.0f;
    float rk = (lane < n) ? S[(int64_t)lane * m] : 0.0f;  // row 0
#pragma unroll 1
    for (int k = 0; k < n; ++k) {
        const float rn = (lane > k && lane < n) ? S[(k + 1) + (int64_t)lane * m] : 0.0f;  // row k+1, columns > k
        const float ck = readlane(cv, k) * readlane(dinv, k);
        cv = (lane == k) ? ck : ((lane > k) ? fmaf(-rk, ck, cv) : cv);
        rk = rn;
    }
    if (lane < n) c[lane] = cv;
}

// One bucket of the ragged plan: workgroup w solves the items
// wstart[w] .. wstart[w+1]-1, a range the plan balanced by estimated cost
// (longest-processing-time assignment).  Per system: FP32 Householder QR
// with FP64 column norms of FP64 A (b rides along as an extra column, so
// the factorisation leaves Q^T b), x0 = R^{-1} Q^T b, r0 = b - A x0 in
// FP64, then `steps` refinements of the augmented system
// [I A; A^T 0][r; x] = [b; 0] (Bjorck 1967): f = b - r - A x, g = -A^T r
// in FP64, and with the FP32 factors h = R^{-T} g, d = Q^T f,
// dx = R^{-1}(d[0:n] - h), dr = Q [h; d[n:m]].  In the first step f is
// b - r0 - A x0 = 0 up to rounding and is not formed.  This recovers FP64
// accuracy also for inconsistent systems, where refining x alone stalls at
// the FP32 level.  x overwrites b[0:n].  Global loads are batched (eight
// per thread in flight) because one wavefront per SIMD leaves little else
// to hide their latency.  METHOD selects the factorisation: 0 stages A in
// LDS and factorises column by column (hqr_factor, lookahead column in
// registers, one barrier per column; NW = 2, 4, 8 or 16), 1 keeps panels in
// registers (hqr_panels, one barrier per panel; NW = 8 or 16).  The plan
// picks METHOD, NW and the grid per bucket.  For one METHOD every system's
// arithmetic is the same for every NW and every item-to-workgroup
// assignment, so the results are too; the two methods round the FP32
// factorisation differently.
template <int MAXE, int NW, int METHOD>
__global__ __launch_bounds__(NW * kWave) void gels_ragged_kernel(const RaggedItem *__restrict__ items,
                                                               const int *__restrict__ wstart,
                                                               const double *__restrict__ A, double *B,
                                                               int steps, int *queue) {
    __shared__ float S[MAXE + 512];  // A (m x n, m*n <= MAXE), then b (m <= 512)
    __shared__ int s_next;
    __shared__ float s_tau[2];
    __shared__ float tau[64];
    __shared__ float c[512];
    __shared__ float h[64];
    __shared__ float dx[64];
    __shared__ double xd[64];
    __shared__ double rd[512];
    __shared__ double bd[512];
    constexpr int NT = NW * kWave;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // static: the workgroup's range wstart[w] .. wstart[w + 1] - 1.  Queue:
    // the bucket's items wstart[0] .. wstart[grid] - 1, most expensive
    // first, each taken by the next free workgroup (queue[0] counts them
    // out; queue[1] counts finished workgroups, and the last one resets
    // both, so the next launch finds them at zero)
    const int it_end = queue ? wstart[gridDim.x] : wstart[blockIdx.x + 1];
    int it = queue ? 0 : wstart[blockIdx.x];
    for (;; ++it) {
        if (queue) {
            __syncthreads();  // the previous system's last reads of s_next are done
            if (threadIdx.x == 0) s_next = wstart[0] + atomicAdd(queue, 1);
            __syncthreads();
            it = s_next;
        }
        if (it >= it_end) break;
        const RaggedItem w = items[it];
        const int m = w.m, n = w.n;
        const double *a = A + w.offA;
        double *bb = B + w.offB;
        __syncthreads();
        float *qb = S + m * n;  // b as column n: the factorisation turns it into Q^T b
        if constexpr (METHOD == 1) {
            for (int i = threadIdx.x; i < m; i += NT) bd[i] = bb[i];
            hqr_panels<NW>(a, bb, S, m, n, tau);  // FP64 norms, FP32 dots
        } else {
            const int mn = m * n;
            int e0 = threadIdx.x;
            for (; e0 + 7 * NT < mn; e0 += 8 * NT) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = a[e0 + u * NT];
#pragma unroll
                for (int u = 0; u < 8; ++u) S[e0 + u * NT] = (float)v[u];
            }
            for (; e0 < mn; e0 += NT) S[e0] = (float)a[e0];
            for (int i = threadIdx.x; i < m; i += NT) {
                const double v = bb[i];
                bd[i] = v;
                qb[i] = (float)v;
            }
            __syncthreads();
            // FP64 norms, FP32 dots; trailing columns updated by 16-lane groups
            hqr_factor<float, double, float, NW, -1>(S, m, m, n, tau, s_tau, n + 1);
        }
        if (wave == 0) trsv_upper_regs(S, m, n, qb);  // x0 = R^{-1} (Q^T b)[0:n]
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += NT) xd[i] = (double)qb[i];
        __syncthreads();
        // r0 = b - A x0 (FP64, A from global memory)
        rows_minus_ax<NT>(a, m, n, xd, bd, nullptr, rd);
        for (int s = 0; s < steps; ++s) {
            __syncthreads();
            // f = b - r - A x  -> c (FP32).  In the first step r = r0 was just
            // formed as b - A x0, so f is zero up to FP64 rounding: it is
            // taken as zero (no pass over A, and Q^T f = 0 below).
            const bool f_zero = (s == 0);
            for (int i = threadIdx.x; i < m; i += NT) {
                double f = 0.0;
                if (!f_zero) {
                    f = bd[i] - rd[i];
                    int k = 0;
                    for (; k + 8 <= n; k += 8) {
                        double av[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) av[u] = a[i + (int64_t)(k + u) * m];
#pragma unroll
                        for (int u = 0; u < 8; ++u) f = fma(-av[u], xd[k + u], f);
                    }
                    for (; k < n; ++k) f = fma(-a[i + (int64_t)k * m], xd[k], f);
                }
                c[i] = (float)f;
            }
            // g = -A^T r -> h (FP32); one wavefront per column, eight rows
            // per lane in flight
            for (int j = wave; j < n; j += NW) {
                const double *col = a + (int64_t)j * m;
                double g = 0.0;
                int i0 = lane;
                for (; i0 + 7 * kWave < m; i0 += 8 * kWave) {
                    double av[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) av[u] = col[i0 + u * kWave];
#pragma unroll
                    for (int u = 0; u < 8; ++u) g = fma(av[u], rd[i0 + u * kWave], g);
                }
                for (; i0 < m; i0 += kWave) g = fma(col[i0], rd[i0], g);
                g = wave_sum(g);
                if (lane == 0) h[j] = (float)(-g);
            }
            __syncthreads();
            // the two independent serial pieces of each half run on two
            // wavefronts at once
            if (wave == 0) trsv_upper_trans_regs(S, m, n, h);                 // h = R^{-T} g
            if (wave == 1 && !f_zero) apply_q_regs(S, m, n, tau, c, true);   // c = Q^T f
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += NT) {
                dx[i] = c[i] - h[i];
                c[i] = h[i];
            }
            __syncthreads();
            if (wave == 0) {
                trsv_upper_regs(S, m, n, dx);                                 // dx = R^{-1} (d1 - h)
                for (int i = lane; i < n; i += kWave) xd[i] += (double)dx[i];
            }
            if (wave == 1) apply_q_regs(S, m, n, tau, c, false);  // c = dr = Q [h; d2]
            __syncthreads();
            for (int i = threadIdx.x; i < m; i += NT) rd[i] += (double)c[i];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += NT) bb[i] = xd[i];
    }
    if (queue && threadIdx.x == 0) {
        __threadfence();  // this workgroup's last take from queue[0] precedes its count
        if (atomicAdd(queue + 1, 1) == (int)gridDim.x - 1) {
            atomicExch(queue, 0);
            atomicExch(queue + 1, 0);
        }
    }
}

// Launch one bucket of a ragged plan: the LDS capacity MAXE, the
// factorisation method and the wavefronts per workgroup pick the
// instantiation.
template <int MAXE, int METHOD, int NW>
void launch_ragged_one(int grid, hipStream_t st, const RaggedItem *items, const int *wstart, const double *A,
                       double *B, int steps, int *queue) {
    hipLaunchKernelGGL((gels_ragged_kernel<MAXE, NW, METHOD>), dim3(grid), dim3(NW * kWave), 0, st, items, wstart, A,
                       B, steps, queue);
}

template <int MAXE>
void launch_ragged_maxe(int method, int waves, int grid, hipStream_t st, const RaggedItem *items, const int *wstart,
                        const double *A, double *B, int steps, int *queue) {
    if (method == 1) {
        if (waves == 16) launch_ragged_one<MAXE, 1, 16>(grid, st, items, wstart, A, B, steps, queue);
        else launch_ragged_one<MAXE, 1, 8>(grid, st, items, wstart, A, B, steps, queue);
    } else {
        if (waves == 2) launch_ragged_one<MAXE, 0, 2>(grid, st, items, wstart, A, B, steps, queue);
        else if (waves == 4) launch_ragged_one<MAXE, 0, 4>(grid, st, items, wstart, A, B, steps, queue);
        else if (waves == 16) launch_ragged_one<MAXE, 0, 16>(grid, st, items, wstart, A, B, steps, queue);
        else launch_ragged_one<MAXE, 0, 8>(grid, st, items, wstart, A, B, steps, queue);
    }
}

// queue: NULL for the static ranges, else the bucket's two counters
int launch_ragged(int maxe, int method, int waves, int grid, hipStream_t st, const RaggedItem *items,
                  const int *wstart, const double *A, double *B, int steps, int *queue) {
    switch (maxe) {
    case 2048: launch_ragged_maxe<2048>(method, waves, grid, st, items, wstart, A, B, steps, queue); break;
    case 4096: launch_ragged_maxe<4096>(method, waves, grid, st, items, wstart, A, B, steps, queue); break;
    case 8192: launch_ragged_maxe<8192>(method, waves, grid, st, items, wstart, A, B, steps, queue); break;
    case 16384: launch_ragged_maxe<16384>(method, waves, grid, st, items, wstart, A, B, steps, queue); break;
    default: launch_ragged_maxe<32768>(method, waves, grid, st, items, wstart, A, B, steps, queue); break;
    }
    return launch_check("gels_ragged_kernel");
}

template <typename T>
int gels_dispatch(int64_t m, int64_t n, int64_t nrhs, T *A, int64_t lda, int64_t strideA, T *B,
                  int64_t ldb, int64_t strideB, T *tauws, int *info, int64_t batch, hipStream_t st) {
    const size_t dyn = sizeof(T) * (size_t)(n + m);
    if (dyn > 48 * 1024) {
        set_error("gels: m + n too large for the vector workspace");
        return -1;
    }
    const dim3 grid((unsigned)batch), block(kThreads);
    const int64_t mn = m * n;
#define GPUTILS_GELS(ME)                                                                                  \
    hipLaunchKernelGGL((gels_kernel<T, ME>), grid, block, dyn, st, (int)m, (int)n, (int)nrhs, A, lda, strideA, \
                       B, ldb, strideB, tauws, info, batch)
    constexpr int kBig = (int)(56 * 1024 / sizeof(T));
    if (mn <= 1024) GPUTILS_GELS(1024);
    else if (mn <= 4096) GPUTILS_GELS(4096);
    else if (mn <= kBig) GPUTILS_GELS(kBig);
    else GPUTILS_GELS(0);
#undef GPUTILS_GELS
    return launch_check("gels_kernel");
}

template <typename T>
int geqrf_dispatch(int64_t m, int64_t n, T *A, int64_t lda, int64_t strideA, T *tau, int64_t strideTau,
                   int *info, int64_t batch, hipStream_t st) {
    const dim3 grid((unsigned)batch), block(kThreads);
    const int64_t mn = m * n;
#define GPUTILS_GEQRF(ME) \
    hipLaunchKernelGGL((geqrf_kernel<T, ME>), grid, block, 0, st, (int)m, (int)n, A, lda, strideA, tau, strideTau, info, batch)
    constexpr int kBig = (int)(56 * 1024 / sizeof(T));
    if (mn <= 1024) GPUTILS_GEQRF(1024);
    else if (mn <= 4096) GPUTILS_GEQRF(4096);
    else if (mn <= kBig) GPUTILS_GEQRF(kBig);
    else GPUTILS_GEQRF(0);
#undef GPUTILS_GEQRF
    return launch_check("geqrf_kernel");
}

}  // namespace qr
}  // namespace gputils

using namespace gputils;

// ---------------------------------------------------------- ragged plan --
// The systems are grouped into buckets by LDS need (FP32 m*n elements);
// each bucket is one launch whose launch shape (wavefronts per workgroup,
// workgroups per CU) the plan holds and the tuner can change.  Workgroup w
// of a bucket solves a contiguous range of the bucket's items, filled by a
// longest-processing-time assignment of the cost-sorted systems.
struct gputils_gels_plan {
    struct Bucket {
        int maxe;
        int begin, count;  // items begin .. begin + count - 1 of d_items
        int wbase;         // first of this bucket's count + 1 slots in d_wstart
        int method;        // factorisation: 0 column steps in LDS, 1 register panels
        int waves;         // wavefronts per workgroup: 2, 4, 8, 16 (method 0) or 8, 16 (method 1)
        int per_cu;        // workgroups per CU the grid is sized for
        int grid;          // min(count, CUs * per_cu)
        bool queue;        // systems taken from a work queue, else static LPT ranges
        std::vector<int> order;  // the bucket's systems, most expensive first
    };
    int64_t count = 0;
    int cus = 256;
    qr::RaggedItem *d_items = nullptr;
    int *d_wstart = nullptr;  // per bucket: grid + 1 absolute item offsets
    int *d_queue = nullptr;   // per bucket: the work queue's two counters (zero between launches)
    std::vector<int32_t> m, n;
    std::vector<int64_t> offA, offB;
    std::vector<Bucket> buckets;
    int64_t extentB = 0;  // max over systems of offB + m
    // bucket b > 0 runs on streams[b - 1]: forked from the caller's stream
    // by `fork`, joined back by done[b - 1]
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;
    hipEvent_t fork = nullptr;
    // buckets concurrently (default) or one after another on the caller's
    // stream; the tuner keeps whichever runs the whole plan faster
    bool concurrent = true;
};

namespace {

constexpr int kRaggedBucketMax[] = {2048, 4096, 8192, 16384, 32768};
constexpr int kRaggedBuckets = 5;

// LDS of gels_ragged_kernel<MAXE, *>: the FP32 matrix and b column plus
// the fixed per-system vectors (s_next 4 and padding 4, s_tau 8, tau 256,
// c 2048, h 256, dx 256, xd 512, rd 4096, bd 4096 bytes)
int ragged_lds_bytes(int maxe) { return (maxe + 512) * 4 + 11536; }

// most workgroups of one bucket that a CU can hold at once: LDS (160 KiB)
// and 32 wavefronts per CU
int ragged_max_per_cu(int maxe, int waves) {
    return std::max(1, std::min(160 * 1024 / ragged_lds_bytes(maxe), 32 / waves));
}

// estimated cost of one system: the factorisation's m n^2 work spread over
// a wavefront plus the n-step serial chains of the solves
double ragged_cost(int m, int n) { return (double)m * n * n / 64.0 + 2000.0 * n; }

// Assign the bucket's items to its grid (longest processing time first:
// each system, most expensive first, goes to the least-loaded workgroup)
// and upload the items and range boundaries.
int ragged_upload_bucket(gputils_gels_plan *p, const gputils_gels_plan::Bucket &bk) {
    if (bk.queue) {
        // the bucket's systems most expensive first; only the first and the
        // last range boundary are read
        std::vector<qr::RaggedItem> items;
        items.reserve((size_t)bk.count);
        for (int o : bk.order) items.push_back({p->m[o], p->n[o], p->offA[o], p->offB[o]});
        std::vector<int> wstart((size_t)bk.grid + 1, bk.begin);
        wstart.back() = bk.begin + bk.count;
        GPUTILS_HIP(hipMemcpy(p->d_items + bk.begin, items.data(), sizeof(qr::RaggedItem) * items.size(),
                              hipMemcpyHostToDevice));
        GPUTILS_HIP(hipMemcpy(p->d_wstart + bk.wbase, wstart.data(), sizeof(int) * wstart.size(),
                              hipMemcpyHostToDevice));
        return GPUTILS_OK;
    }
    std::vector<std::vector<int>> bins((size_t)bk.grid);
    std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>,
                        std::greater<std::pair<double, int>>> heap;
    for (int w = 0; w < bk.grid; ++w) heap.push({0.0, w});
    for (int o : bk.order) {
        auto top = heap.top();
        heap.pop();
        bins[(size_t)top.second].push_back(o);
        heap.push({top.first + ragged_cost(p->m[o], p->n[o]), top.second});
    }
    std::vector<qr::RaggedItem> items;
    std::vector<int> wstart;
    items.reserve((size_t)bk.count);
    wstart.reserve((size_t)bk.grid + 1);
    for (int w = 0; w < bk.grid; ++w) {
        wstart.push_back(bk.begin + (int)items.size());
        for (int o : bins[(size_t)w]) items.push_back({p->m[o], p->n[o], p->offA[o], p->offB[o]});
    }
    wstart.push_back(bk.begin + (int)items.size());
    GPUTILS_HIP(hipMemcpy(p->d_items + bk.begin, items.data(), sizeof(qr::RaggedItem) * items.size(),
                          hipMemcpyHostToDevice));
    GPUTILS_HIP(hipMemcpy(p->d_wstart + bk.wbase, wstart.data(), sizeof(int) * wstart.size(),
                          hipMemcpyHostToDevice));
    return GPUTILS_OK;
}

// One execution of the plan on st.  Concurrent: buckets after the first
// wait for the work before this call on st (fork), run on the plan's
// streams, and st waits for all of them (join), so the call stays ordered
// on st; otherwise every bucket runs on st in turn.
int ragged_execute(gputils_gels_plan *plan, const double *A, double *B, int steps, hipStream_t st) {
    const size_t nb = plan->buckets.size();
    const bool conc = plan->concurrent && nb > 1;
    if (conc) GPUTILS_HIP(hipEventRecord(plan->fork, st));
    for (size_t b = 0; b < nb; ++b) {
        const auto &bk = plan->buckets[b];
        hipStream_t sb = (conc && b > 0) ? plan->streams[b - 1] : st;
        if (conc && b > 0) GPUTILS_HIP(hipStreamWaitEvent(sb, plan->fork, 0));
        int rc = qr::launch_ragged(bk.maxe, bk.method, bk.waves, bk.grid, sb, plan->d_items,
                                   plan->d_wstart + bk.wbase, A, B, steps, bk.queue ? plan->d_queue + 2 * b : nullptr);
        if (rc) return rc;
        if (conc && b > 0) GPUTILS_HIP(hipEventRecord(plan->done[b - 1], sb));
    }
    if (conc)
        for (size_t b = 1; b < nb; ++b) GPUTILS_HIP(hipStreamWaitEvent(st, plan->done[b - 1], 0));
    return GPUTILS_OK;
}

void ragged_set_shape(const gputils_gels_plan *p, gputils_gels_plan::Bucket &bk, int method, int waves, int per_cu,
                      bool queue) {
    bk.queue = queue;
    bk.method = method;
    bk.waves = waves;
    bk.per_cu = per_cu;
    bk.grid = (int)std::min<int64_t>(bk.count, (int64_t)p->cus * per_cu);
}

}  // namespace

extern "C" {

int gputils_sgels_batched(int64_t m, int64_t n, int64_t nrhs, float *A, int64_t lda, int64_t strideA,
                          float *B, int64_t ldb, int64_t strideB, int *info, int64_t batch, void *stream) {
    GPUTILS_ARG(m >= 0, 1, "m < 0");
    GPUTILS_ARG(n >= 0 && n <= m, 2, "need 0 <= n <= m");
    GPUTILS_ARG(nrhs >= 0, 3, "nrhs < 0");
    GPUTILS_ARG(lda >= (m > 1 ? m : 1), 5, "lda < max(1,m)");
    GPUTILS_ARG(ldb >= (m > 1 ? m : 1), 8, "ldb < max(1,m)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 11, "batch out of range");
    if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 4, "A is NULL");
    GPUTILS_ARG(nrhs == 0 || B != nullptr, 7, "B is NULL");
    return qr::gels_dispatch<float>(m, n, nrhs, A, lda, strideA, B, ldb, strideB, nullptr, info, batch,
                                    as_stream(stream));
}

int gputils_dgels_batched(int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t strideA,
                          double *B, int64_t ldb, int64_t strideB, int *info, int64_t batch, void *stream) {
    GPUTILS_ARG(m >= 0, 1, "m < 0");
    GPUTILS_ARG(n >= 0 && n <= m, 2, "need 0 <= n <= m");
    GPUTILS_ARG(nrhs >= 0, 3, "nrhs < 0");
    GPUTILS_ARG(lda >= (m > 1 ? m : 1), 5, "lda < max(1,m)");
    GPUTILS_ARG(ldb >= (m > 1 ? m : 1), 8, "ldb < max(1,m)");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 11, "batch out of range");
    if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 4, "A is NULL");
    GPUTILS_ARG(nrhs == 0 || B != nullptr, 7, "B is NULL");
    return qr::gels_dispatch<double>(m, n, nrhs, A, lda, strideA, B, ldb, strideB, nullptr, info, batch,
                                     as_stream(stream));
}

#define GPUTILS_GEQRF_ENTRY(NAME, T)                                                                    \
    int NAME(int64_t m, int64_t n, T *A, int64_t lda, int64_t strideA, T *tau, int64_t strideTau,       \
             int *info, int64_t batch, void *stream) {                                                  \
        GPUTILS_ARG(m >= 0, 1, "m < 0");                                                                \
        GPUTILS_ARG(n >= 0 && n <= m, 2, "need 0 <= n <= m");                                           \
        GPUTILS_ARG(lda >= (m > 1 ? m : 1), 4, "lda < max(1,m)");                                       \
        GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 9, "batch out of range");                        \
        if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;                                          \
        GPUTILS_ARG(A != nullptr, 3, "A is NULL");                                                      \
        GPUTILS_ARG(tau != nullptr, 6, "tau is NULL");                                                  \
        return qr::geqrf_dispatch<T>(m, n, A, lda, strideA, tau, strideTau, info, batch, as_stream(stream)); \
    }
GPUTILS_GEQRF_ENTRY(gputils_sgeqrf_batched, float)
GPUTILS_GEQRF_ENTRY(gputils_dgeqrf_batched, double)

#define GPUTILS_ORMQR_ENTRY(NAME, T)                                                                     \
    int NAME(int trans, int64_t m, int64_t ncols, int64_t k, const T *A, int64_t lda, int64_t strideA,  \
             const T *tau, int64_t strideTau, T *C, int64_t ldc, int64_t strideC, int64_t batch,        \
             void *stream) {                                                                             \
        GPUTILS_ARG(trans == 0 || trans == 1, 1, "trans must be 0 or 1");                               \
        GPUTILS_ARG(m >= 0, 2, "m < 0");                                                                 \
        GPUTILS_ARG(ncols >= 0, 3, "ncols < 0");                                                         \
        GPUTILS_ARG(k >= 0 && k <= m, 4, "need 0 <= k <= m");                                            \
        GPUTILS_ARG(lda >= (m > 1 ? m : 1), 6, "lda < max(1,m)");                                        \
        GPUTILS_ARG(ldc >= (m > 1 ? m : 1), 11, "ldc < max(1,m)");                                       \
        GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 13, "batch out of range");                        \
        GPUTILS_ARG(sizeof(T) * (size_t)m * 4 <= 64 * 1024, 2, "m too large");                           \
        if (m == 0 || ncols == 0 || k == 0 || batch == 0) return GPUTILS_OK;                             \
        hipLaunchKernelGGL((qr::ormqr_kernel<T>), dim3((unsigned)batch), dim3(qr::kThreads),             \
                           sizeof(T) * (size_t)m * qr::kWaves, as_stream(stream), trans, (int)m,         \
                           (int)ncols, (int)k, A, lda, strideA, tau, strideTau, C, ldc, strideC, batch); \
        return launch_check("ormqr_kernel");                                                             \
    }
GPUTILS_ORMQR_ENTRY(gputils_sormqr_batched, float)
GPUTILS_ORMQR_ENTRY(gputils_dormqr_batched, double)

#define GPUTILS_TRSM_ENTRY(NAME, T)                                                                      \
    int NAME(int64_t n, int64_t nrhs, const T *R, int64_t ldr, int64_t strideR, T *B, int64_t ldb,      \
             int64_t strideB, int64_t batch, void *stream) {                                             \
        GPUTILS_ARG(n >= 0, 1, "n < 0");                                                                 \
        GPUTILS_ARG(nrhs >= 0, 2, "nrhs < 0");                                                           \
        GPUTILS_ARG(ldr >= (n > 1 ? n : 1), 4, "ldr < max(1,n)");                                        \
        GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");                                        \
        GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 9, "batch out of range");                         \
        GPUTILS_ARG(sizeof(T) * (size_t)n * 4 <= 64 * 1024, 1, "n too large");                           \
        if (n == 0 || nrhs == 0 || batch == 0) return GPUTILS_OK;                                        \
        hipLaunchKernelGGL((qr::trsm_upper_kernel<T>), dim3((unsigned)batch), dim3(qr::kThreads),        \
                           sizeof(T) * (size_t)n * qr::kWaves, as_stream(stream), (int)n, (int)nrhs, R,  \
                           ldr, strideR, B, ldb, strideB, batch);                                        \
        return launch_check("trsm_upper_kernel");                                                        \
    }
GPUTILS_TRSM_ENTRY(gputils_strsm_upper_batched, float)
GPUTILS_TRSM_ENTRY(gputils_dtrsm_upper_batched, double)

int gputils_gels_ragged_plan_create(gputils_gels_plan_t *plan, int64_t count, const int32_t *m,
                                    const int32_t *n, const int64_t *offA, const int64_t *offB) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(count >= 0 && count <= 0x7fffffff, 2, "count out of range");
    GPUTILS_ARG(count == 0 || (m && n && offA && offB), 3, "size/offset arrays are NULL");
    *plan = nullptr;
    for (int64_t i = 0; i < count; ++i) {
        if (m[i] < 1 || n[i] < 1 || n[i] > m[i] || n[i] > 64 || m[i] > 512) {
            set_error("invalid argument 3: system %lld has m=%d n=%d (need 1 <= n <= m, n <= 64, m <= 512)",
                      (long long)i, m[i], n[i]);
            return -3;
        }
    }
    auto *p = new gputils_gels_plan();
    p->count = count;
    p->m.assign(m, m + count);
    p->n.assign(n, n + count);
    p->offA.assign(offA, offA + count);
    p->offB.assign(offB, offB + count);
    for (int64_t i = 0; i < count; ++i) p->extentB = std::max(p->extentB, offB[i] + m[i]);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        p->cus = prop.multiProcessorCount;
    // buckets by LDS need, largest first; within a bucket by cost, largest first
    std::vector<std::vector<int>> members(kRaggedBuckets);
    for (int i = 0; i < (int)count; ++i) {
        const int64_t e = (int64_t)m[i] * n[i];
        int bk = 0;
        while (bk + 1 < kRaggedBuckets && e > kRaggedBucketMax[bk]) ++bk;
        members[(size_t)bk].push_back(i);
    }
    int begin = 0;
    for (int bk = kRaggedBuckets - 1; bk >= 0; --bk) {
        auto &mem = members[(size_t)bk];
        if (mem.empty()) continue;
        std::stable_sort(mem.begin(), mem.end(),
                         [&](int x, int y) { return ragged_cost(m[x], n[x]) > ragged_cost(m[y], n[y]); });
        gputils_gels_plan::Bucket b;
        b.maxe = kRaggedBucketMax[bk];
        b.begin = begin;
        b.count = (int)mem.size();
        b.wbase = begin + (int)p->buckets.size();
        b.order = std::move(mem);
        // untuned shape: column steps, eight wavefronts, as many workgroups
        // per CU as fit
        ragged_set_shape(p, b, 0, 8, ragged_max_per_cu(b.maxe, 8), false);
        begin += b.count;
        p->buckets.push_back(std::move(b));
    }
    // the buckets' launches run concurrently (each one's tail overlaps the
    // others'): one stream and one event per bucket after the first
    if (p->buckets.size() > 1) {
        hipError_t e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
        for (size_t b = 1; e == hipSuccess && b < p->buckets.size(); ++b) {
            hipStream_t sb = nullptr;
            hipEvent_t eb = nullptr;
            e = hipStreamCreateWithFlags(&sb, hipStreamNonBlocking);
            if (e == hipSuccess) p->streams.push_back(sb);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&eb, hipEventDisableTiming);
            if (e == hipSuccess) p->done.push_back(eb);
        }
        if (e != hipSuccess) {
            gputils_gels_ragged_plan_destroy(p);
            return hip_fail(e, "plan streams");
        }
    }
    if (count > 0) {
        hipError_t e = hipMalloc(&p->d_items, sizeof(qr::RaggedItem) * (size_t)count);
        if (e == hipSuccess) e = hipMalloc(&p->d_wstart, sizeof(int) * ((size_t)count + p->buckets.size()));
        if (e == hipSuccess) e = hipMalloc(&p->d_queue, sizeof(int) * 2 * p->buckets.size());
        if (e == hipSuccess) e = hipMemset(p->d_queue, 0, sizeof(int) * 2 * p->buckets.size());
        int rc = e == hipSuccess ? GPUTILS_OK : hip_fail(e, "plan allocation");
        for (size_t i = 0; rc == GPUTILS_OK && i < p->buckets.size(); ++i) rc = ragged_upload_bucket(p, p->buckets[i]);
        if (rc != GPUTILS_OK) {
            gputils_gels_ragged_plan_destroy(p);
            return rc;
        }
    }
    *plan = p;
    return GPUTILS_OK;
}

int gputils_gels_ragged_plan_destroy(gputils_gels_plan_t plan) {
    if (plan == nullptr) return GPUTILS_OK;
    hipError_t e = hipSuccess;
    auto keep = [&e](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    if (plan->d_items) keep(hipFree(plan->d_items));
    if (plan->d_wstart) keep(hipFree(plan->d_wstart));
    if (plan->d_queue) keep(hipFree(plan->d_queue));
    for (hipStream_t sb : plan->streams) keep(hipStreamDestroy(sb));
    for (hipEvent_t eb : plan->done) keep(hipEventDestroy(eb));
    if (plan->fork) keep(hipEventDestroy(plan->fork));
    delete plan;
    return e == hipSuccess ? GPUTILS_OK : hip_fail(e, "plan destroy");
}

int gputils_gels_ragged_plan_describe(gputils_gels_plan_t plan, int32_t *desc, int capacity, int *nbuckets) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(capacity >= 0, 3, "capacity < 0");
    GPUTILS_ARG(capacity == 0 || desc != nullptr, 2, "desc is NULL");
    if (nbuckets) *nbuckets = (int)plan->buckets.size();
    for (int i = 0; i < capacity && i < (int)plan->buckets.size(); ++i) {
        const auto &bk = plan->buckets[(size_t)i];
        const int32_t v[7] = {bk.maxe, bk.count, bk.method, bk.waves, bk.per_cu, bk.grid, bk.queue ? 1 : 0};
        for (int q = 0; q < 7; ++q) desc[7 * i + q] = v[q];
    }
    return GPUTILS_OK;
}

int gputils_gels_ragged_plan_configure(gputils_gels_plan_t plan, int bucket, int method, int waves, int per_cu,
                                       int queue) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(bucket >= 0 && bucket < (int)plan->buckets.size(), 2, "bucket out of range");
    GPUTILS_ARG(method == 0 || method == 1, 3, "method must be 0 or 1");
    GPUTILS_ARG(method == 0 ? (waves == 2 || waves == 4 || waves == 8 || waves == 16) : (waves == 8 || waves == 16), 4,
                "waves must be 2, 4, 8 or 16 (method 0) or 8 or 16 (method 1)");
    GPUTILS_ARG(per_cu >= 1 && per_cu <= 64, 5, "per_cu must be in [1, 64]");
    GPUTILS_ARG(queue == 0 || queue == 1, 6, "queue must be 0 or 1");
    // the device tables change: no execution of the plan may be pending
    GPUTILS_HIP(hipDeviceSynchronize());
    auto &bk = plan->buckets[(size_t)bucket];
    ragged_set_shape(plan, bk, method, waves, per_cu, queue != 0);
    return ragged_upload_bucket(plan, bk);
}

int gputils_gels_ragged_plan_tune(gputils_gels_plan_t plan, const double *A, const double *B, int refine_steps,
                                  void *stream) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(refine_steps >= 0 && refine_steps <= 16, 4, "refine_steps must be in [0, 16]");
    if (plan->count == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 2, "A is NULL");
    GPUTILS_ARG(B != nullptr, 3, "B is NULL");
    hipStream_t st = as_stream(stream);
    GPUTILS_HIP(hipDeviceSynchronize());
    // candidates run on a copy of B (refreshed before every run) so the
    // caller's right-hand sides stay as they are
    double *scratch = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipError_t e = hipMalloc(&scratch, sizeof(double) * (size_t)plan->extentB);
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    int rc = e == hipSuccess ? GPUTILS_OK : hip_fail(e, "tune set-up");
    constexpr int kRepeats = 3;
    for (size_t bi = 0; rc == GPUTILS_OK && bi < plan->buckets.size(); ++bi) {
        auto &bk = plan->buckets[bi];
        int best_m = bk.method, best_w = bk.waves, best_pc = bk.per_cu;
        bool best_q = bk.queue;
        float best_ms = -1.0f;
        std::vector<std::pair<int, int>> seen;  // (queue * 128 + method * 64 + waves, grid) already timed
        static const int kShapes[][2] = {{0, 2}, {0, 4}, {0, 8}, {0, 16}, {1, 8}, {1, 16}};
        for (int qmode = 0; qmode < 2; ++qmode)
        for (const auto &sh : kShapes) {
            const int method = sh[0], waves = sh[1];
            const int cap = ragged_max_per_cu(bk.maxe, waves);
            for (int pc = 1; pc <= cap && rc == GPUTILS_OK; ++pc) {
                ragged_set_shape(plan, bk, method, waves, pc, qmode == 1);
                const auto key = std::make_pair(qmode * 128 + method * 64 + waves, bk.grid);
                if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
                seen.push_back(key);
                rc = ragged_upload_bucket(plan, bk);
                float t_min = -1.0f;
                for (int r = 0; rc == GPUTILS_OK && r <= kRepeats; ++r) {  // run 0 warms up
                    e = hipMemcpyAsync(scratch, B, sizeof(double) * (size_t)plan->extentB, hipMemcpyDeviceToDevice, st);
                    if (e == hipSuccess) e = hipEventRecord(ev0, st);
                    if (e != hipSuccess) {
                        rc = hip_fail(e, "tune");
                        break;
                    }
                    rc = qr::launch_ragged(bk.maxe, bk.method, bk.waves, bk.grid, st, plan->d_items,
                                           plan->d_wstart + bk.wbase, A, scratch, refine_steps,
                                           bk.queue ? plan->d_queue + 2 * bi : nullptr);
                    if (rc != GPUTILS_OK) break;
                    e = hipEventRecord(ev1, st);
                    if (e == hipSuccess) e = hipEventSynchronize(ev1);
                    float ms = 0.0f;
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
                    if (e != hipSuccess) {
                        rc = hip_fail(e, "tune timing");
                        break;
                    }
                    if (r > 0 && (t_min < 0.0f || ms < t_min)) t_min = ms;
                }
                if (rc == GPUTILS_OK && (best_ms < 0.0f || t_min < best_ms)) {
                    best_ms = t_min;
                    best_m = method;
                    best_w = waves;
                    best_pc = pc;
                    best_q = qmode == 1;
                }
            }
        }
        ragged_set_shape(plan, bk, best_m, best_w, best_pc, best_q);
        if (rc == GPUTILS_OK) rc = ragged_upload_bucket(plan, bk);
    }
    // with the shapes fixed, the whole plan concurrently and in turn
    if (rc == GPUTILS_OK && plan->buckets.size() > 1) {
        float best[2] = {-1.0f, -1.0f};
        for (int mode = 0; mode < 2 && rc == GPUTILS_OK; ++mode) {
            plan->concurrent = mode == 1;
            for (int r = 0; r <= kRepeats && rc == GPUTILS_OK; ++r) {
                e = hipMemcpyAsync(scratch, B, sizeof(double) * (size_t)plan->extentB, hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess) e = hipEventRecord(ev0, st);
                if (e != hipSuccess) {
                    rc = hip_fail(e, "tune");
                    break;
                }
                rc = ragged_execute(plan, A, scratch, refine_steps, st);
                if (rc != GPUTILS_OK) break;
                e = hipEventRecord(ev1, st);
                if (e == hipSuccess) e = hipEventSynchronize(ev1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
                if (e != hipSuccess) {
                    rc = hip_fail(e, "tune timing");
                    break;
                }
                if (r > 0 && (best[mode] < 0.0f || ms < best[mode])) best[mode] = ms;
            }
        }
        plan->concurrent = rc == GPUTILS_OK && best[1] < best[0];
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (scratch) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(scratch);
    }
    return rc;
}

int gputils_dgels_ragged(gputils_gels_plan_t plan, const double *A, double *B, int refine_steps,
                         void *stream) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(refine_steps >= 0 && refine_steps <= 16, 4, "refine_steps must be in [0, 16]");
    if (plan->count == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 2, "A is NULL");
    GPUTILS_ARG(B != nullptr, 3, "B is NULL");
    return ragged_execute(plan, A, B, refine_steps, as_stream(stream));
}

int gputils_gels_ragged_plan_set_concurrency(gputils_gels_plan_t plan, int concurrent) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(concurrent == 0 || concurrent == 1, 2, "concurrent must be 0 or 1");
    plan->concurrent = concurrent != 0;
    return GPUTILS_OK;
}

int gputils_gels_ragged_plan_get_concurrency(gputils_gels_plan_t plan, int *concurrent) {
    GPUTILS_ARG(plan != nullptr, 1, "plan is NULL");
    GPUTILS_ARG(concurrent != nullptr, 2, "concurrent is NULL");
    *concurrent = plan->concurrent ? 1 : 0;
    return GPUTILS_OK;
}

}  // extern "C"
