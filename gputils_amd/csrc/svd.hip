// svd.hip -- batched SVD, rank count and nullspace construction for gfx950.
//
// Replaces the host loop of cusolverDn?gesvd calls behind Svd<T>
// (include/tensor.cuh:1624-1676, jobu 'A' or 'N', jobvt 'A'), the rank
// kernel k_countNonzeroSingularValues (tensor.cuh:1484-1491, 1600-1609) and
// the per-matrix slice/copy/GEMM loop of the Nullspace constructor
// (tensor.cuh:2046-2079).
//
// One 256-thread workgroup per matrix, FP64 arithmetic for both float and
// double data except the first Jacobi sweeps of float data without VT
// (step 2):
//   1. G = A (m x n, m >= n) is staged in LDS and factorised G = Q [R; 0]
//      (Householder, LAPACK conventions).  LAPACK dgesvd takes the same
//      QR-first path for m >= 1.6 n, so Q[:, n:m] is the completion of U it
//      returns there.
//   2. One-sided (Hestenes) Jacobi on the columns of W = R (or R^T when VT
//      is not wanted), V accumulated.  Parallel tournament ordering: n/2
//      disjoint column pairs per round, 8 lanes per pair, one barrier per
//      round; sweeps repeat until no pair needs a rotation,
//      |w_p . w_q| <= n eps ||w_p|| ||w_q||.  Float data without VT
//      (n <= 64) sweeps first in FP32 until the largest rotation left is
//      below 1e-2, then one Newton-Schulz step and W = W0 V in FP64
//      (fp32_first_sweeps), and the FP64 sweeps finish from there.
//   3. sigma_j = ||w_j||, sorted descending; for nonzero sigma_j, v_j is
//      signed so that its largest-magnitude entry is negative and
//      u_j = w_j / sigma_j.  For a numerically zero sigma_j
//      (<= m eps sigma_max) v_j keeps its sign and u_j is the first unit
//      vector e_k whose residual against the columns already placed exceeds
//      1/2 (two Gram-Schmidt passes), as LAPACK's bidiagonal QR leaves it.
//   4. U = Q diag(U_R, I_{m-n}): for m <= 128 all columns at once in
//      registers (eight per 16-lane group, eight rows per lane); otherwise up
//      to 16 columns at a time with one 16-lane group per column.
//
// Where the matrices live (the kernel's MODE): kResident keeps G, W and V
// in LDS; kParked factorises G in LDS, parks the reflectors in the global
// workspace and lets W and V reuse G's LDS (half the LDS of kResident, so
// twice the workgroups per CU); kGlobalG moves G to the global workspace
// when it does not fit; kWide (n > 64, or m too large for the others) keeps
// G, W and V in the global workspace and LDS holds only the per-column
// scalars.  All but kWide hold each Jacobi pair's column entries in
// registers (n <= 64); kWide loops over them.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <type_traits>

#include "common.hpp"
#include "householder.hpp"

namespace gputils {
namespace svd {

using hh::kThreads;
using hh::wave_sync;

constexpr int kMaxN = 64;  // largest n whose Jacobi pairs are held in registers
constexpr int kMaxSweeps = 40;
constexpr int kMaxLds = 160 * 1024;

constexpr int kResident = 0;  // G, W, V in LDS
constexpr int kGlobalG = 1;   // G in the workspace, W, V in LDS
constexpr int kWide = 2;      // G, W, V in the workspace
constexpr int kParked = 3;    // G in LDS for the QR, then in the workspace; W, V over G's LDS

// Storage layout, in doubles unless noted: offsets into LDS, or (kWide, and
// G of kGlobalG) into the matrix's slice of the workspace.
struct Layout {
    int64_t g;       // G (m x n)
    int64_t w;       // W (n x n) during Jacobi, then a chunk of U (m x ch)
    int64_t ws;      // size of the W region
    int64_t x;       // V (n x n) during Jacobi, then U_R
    int sig;         // n
    int tau;         // n + 2 (two lookahead slots)
    int sgn;         // n
    int ints;        // pos (n), ord (n), nonzero count (1) as ints
    int ch;          // U columns per chunk
    int bytes;       // LDS bytes
    int64_t wstride; // workspace doubles per matrix
};

__host__ __device__ inline Layout layout(int m, int n, int mode) {
    Layout L;
    const int64_t mn = (int64_t)m * n, nn = (int64_t)n * n;
    L.ws = (nn > 4 * (int64_t)m) ? nn : 4 * (int64_t)m;
    int64_t o = 0, gofs = 0;  // LDS and workspace cursors
    if (mode == kParked) {
        // one LDS region holds G during the QR, then W and V; the reflectors
        // are parked in the workspace for the U phase
        L.g = 0;
        L.w = 0;
        L.x = L.ws;
        o = (mn > L.ws + nn) ? mn : L.ws + nn;
        gofs = mn;
    } else if (mode == kResident) {
        L.g = o;
        o += mn;
    } else {
        L.g = gofs;
        gofs += mn;
    }
    if (mode == kParked) {
    } else if (mode == kWide) {
        L.w = gofs;
        gofs += L.ws;
        L.x = gofs;
        gofs += nn;
    } else {
        L.w = o;
        o += L.ws;
        L.x = o;
        o += nn;
    }
    L.sig = (int)o;
    o += n;
    L.tau = (int)o;
    o += n + 2;
    L.sgn = (int)o;
    o += n;
    L.ints = (int)o;
    o += (2 * n + 2) / 2 + 1;
    L.ch = L.ws / m < 16 ? (int)(L.ws / m) : 16;
    L.bytes = (int)(o * 8 < (int64_t)1 << 30 ? o * 8 : (int64_t)1 << 30);
    L.wstride = gofs;
    return L;
}

// Rotation that makes columns p and q orthogonal, from al = |w_p|^2,
// be = |w_q|^2 and ga = w_p . w_q (ga != 0):
// t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)) with zeta = (be - al) / (2 ga),
// multiplied through by 2 |ga|: one square root and one division, both
// from the hardware estimates with one Newton step each (t only sets the
// angle).  cs = 1 / sqrt(1 + t^2) takes two Newton steps, so that
// cs^2 + sn^2 = 1 to FP64 rounding and the rotation itself loses no
// orthogonality above the convergence tolerance n eps.
__device__ __forceinline__ void jacobi_angle(double al, double be, double ga, double &cs, double &sn) {
    const double d = be - al, g2 = 2.0 * fabs(ga);
    const double sg = (d == 0.0 || ((d > 0.0) == (ga > 0.0))) ? 1.0 : -1.0;
    const double h = fma(d, d, g2 * g2);
    double rh = __builtin_amdgcn_rsq(h);
    rh = rh * fma(-0.5 * h, rh * rh, 1.5);
    const double den = fma(h, rh, fabs(d));  // |d| + sqrt(h)
    double inv = __builtin_amdgcn_rcp(den);
    inv = inv * fma(-den, inv, 2.0);
    const double t = sg * g2 * inv;
    const double t2p1 = fma(t, t, 1.0);
    cs = __builtin_amdgcn_rsq(t2p1);
    cs = cs * fma(-0.5 * t2p1, cs * cs, 1.5);
    cs = cs * fma(-0.5 * t2p1, cs * cs, 1.5);
    sn = cs * t;
}

// The same rotation in FP32 for the first sweeps of float data: the
// hardware square root and reciprocal estimates (1 ulp) without Newton
// steps.  The caller scales al, be, ga to max(al, be) = 1 and calls it only
// when ga^2 > tol^2 al be, so that h >= min(1/4, 2 tol^2) is a normal
// number (|d| >= 1/2 unless both sums exceed 1/2).
__device__ __forceinline__ void jacobi_angle(float al, float be, float ga, float &cs, float &sn) {
    const float d = be - al, g2 = 2.0f * fabsf(ga);
    const float sg = (d == 0.0f || ((d > 0.0f) == (ga > 0.0f))) ? 1.0f : -1.0f;
    const float h = fmaf(d, d, g2 * g2);
    const float den = fmaf(h, __builtin_amdgcn_rsqf(h), fabsf(d));  // |d| + sqrt(h)
    const float t = sg * g2 * __builtin_amdgcn_rcpf(den);
    cs = __builtin_amdgcn_rsqf(fmaf(t, t, 1.0f));
    sn = cs * t;
}

// x + (x of the lane that DPP control CTRL selects), within 16-lane rows
template <int CTRL>
__device__ __forceinline__ double dpp_plus(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return x + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sums over aligned groups of 8 / 16 lanes, result in every lane of the
// group: quad permutations [1,0,3,2] and [2,3,0,1] sum each quad, the row
// half mirror (lane i <-> 7 - i) adds the other quad of the 8, the row
// mirror (lane i <-> 15 - i) the other 8 of the 16.  Register moves only.
__device__ __forceinline__ double group8_sum(double x) {
    x = dpp_plus<0xB1>(x);      // quad_perm [1,0,3,2]
    x = dpp_plus<0x4E>(x);      // quad_perm [2,3,0,1]
    return dpp_plus<0x141>(x);  // row_half_mirror
}
__device__ __forceinline__ double group16_sum(double x) {
    return dpp_plus<0x140>(group8_sum(x));  // row_mirror
}
template <int CTRL>
__device__ __forceinline__ float dpp_plus(float x) {
    return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float group8_sum(float x) {
    x = dpp_plus<0xB1>(x);
    x = dpp_plus<0x4E>(x);
    return dpp_plus<0x141>(x);
}

// Pair i of round r in a round-robin tournament over n2 (even) players:
// player 0 is fixed, players 1..n2-1 rotate.  Every unordered pair meets
// exactly once in the n2 - 1 rounds.
__device__ __forceinline__ void tournament_pair(int i, int r, int n1, int &p, int &q) {
    int a, b;
    if (i == 0) {
        a = 0;
        b = 1 + r;
    } else {
        a = 1 + (r + i) % n1;
        b = 1 + (r - i + n1) % n1;
    }
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// One Jacobi rotation of the column pair (p, q) by the 8 lanes gl = 0..7
// of a group, for any n: entries i = gl + 8 u are read, summed by
// group8_sum and rotated in place (W and V columns).  Returns whether the
// pair needed a rotation (uniform over the group).
__device__ __forceinline__ bool jacobi_pair_strided(double *wp, double *wq, double *vp, double *vq, int n, int gl,
                                                    double tol2) {
    double al = 0.0, be = 0.0, ga = 0.0;
    for (int i = gl; i < n; i += 8) {
        const double x = wp[i], y = wq[i];
        al = fma(x, x, al);
        be = fma(y, y, be);
        ga = fma(x, y, ga);
    }
    al = group8_sum(al);
    be = group8_sum(be);
    ga = group8_sum(ga);
    if (!(ga * ga > tol2 * al * be)) return false;
    double cs, sn;
    jacobi_angle(al, be, ga, cs, sn);
    for (int i = gl; i < n; i += 8) {
        const double x = wp[i], y = wq[i], xv = vp[i], yv = vq[i];
        wp[i] = cs * x - sn * y;
        wq[i] = sn * x + cs * y;
        vp[i] = cs * xv - sn * yv;
        vq[i] = sn * xv + cs * yv;
    }
    return true;
}

// Completion of X's columns nz..n-1 (numerically zero sigma) by one
// wavefront, for any n: column col is the first unit vector e_k (k from
// `next` on) whose residual against columns 0..col-1 (two Gram-Schmidt
// passes, worked in place in column col) exceeds 1/2, normalised; if none
// does, the one with the largest residual.
__device__ void complete_basis_strided(double *X, int n, int nz, int lane) {
    int next = 0;
    for (int col = nz; col < n; ++col) {
        double *x = X + (int64_t)col * n;
        auto residual_of = [&](int k) {
            for (int i = lane; i < n; i += 64) x[i] = (i == k) ? 1.0 : 0.0;
            wave_sync();
            for (int pass = 0; pass < 2; ++pass)
                for (int c = 0; c < col; ++c) {
                    const double *uc = X + (int64_t)c * n;
                    double part = 0.0;
                    for (int i = lane; i < n; i += 64) part = fma(uc[i], x[i], part);
                    const double coef = wave_sum(part);
                    for (int i = lane; i < n; i += 64) x[i] -= coef * uc[i];
                    wave_sync();
                }
            double sq = 0.0;
            for (int i = lane; i < n; i += 64) sq = fma(x[i], x[i], sq);
            return sqrt(wave_sum(sq));
        };
        if (next >= n) next = 0;  // every unit vector used: search them all again
        double best_norm = -1.0, nrm = -1.0;
        int best_k = next;
        for (int k = next; k < n; ++k) {
            nrm = residual_of(k);
            if (nrm > best_norm) {
                best_norm = nrm;
                best_k = k;
            }
            if (nrm > 0.5) break;
        }
        if (!(nrm > 0.5) && best_k != n - 1) best_norm = residual_of(best_k);  // the best was overwritten
        for (int i = lane; i < n; i += 64) x[i] /= best_norm;
        wave_sync();
        if (best_norm > 0.5) next = best_k + 1;
    }
}

// Whether no column pair of W (n x n, n <= 64, column-major, ld n, FP64 in
// LDS) needs a rotation: g_pq^2 <= tol2 g_pp g_qq for all p < q, with
// G = W^T W.  This is the test a sweep that rotates nothing applies pair by
// pair in n - 1 rounds; here G's upper tile triangle comes from FP64 MFMA
// 16x16x4 products (wavefront w takes tiles w, w + 4, ..., the operands
// read straight from W), its diagonal goes through gdiag (n doubles of
// LDS), and one barrier-or gives the answer.
__device__ bool gram_converged(const double *W, int n, double tol2, double *gdiag) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int T = (n + 15) >> 4, ntiles = T * (T + 1) / 2;
    constexpr int kTilesPerWave = 3;  // 10 tiles for n = 64 over 4 wavefronts
    f64x4 acc[kTilesPerWave];
    int ta[kTilesPerWave], tb[kTilesPerWave];
#pragma unroll
    for (int u = 0; u < kTilesPerWave; ++u) {
        const int t = wave + 4 * u;
        acc[u] = Mfma16<double>::zero();
        ta[u] = tb[u] = -1;
        if (t < ntiles) {  // uniform per wavefront
            int a = 0, rem = t;
            while (rem >= T - a) {
                rem -= T - a;
                ++a;
            }
            const int bt = a + rem;
            ta[u] = a;
            tb[u] = bt;
            const int p = 16 * a + (lane & 15), q = 16 * bt + (lane & 15);
            const int pc = p < n ? p : n - 1, qc = q < n ? q : n - 1;
            for (int k0 = 0; k0 < n; k0 += 4) {
                const int i = k0 + (lane >> 4), ic = i < n ? i : n - 1;  // reads clamped into W
                const double av = (i < n && p < n) ? W[ic + pc * n] : 0.0;
                const double bv = (i < n && q < n) ? W[ic + qc * n] : 0.0;
                acc[u] = Mfma16<double>::fma(av, bv, acc[u]);  // += W^T[p][i] W[i][q]
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kTilesPerWave; ++u) {
        if (ta[u] >= 0 && ta[u] == tb[u]) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ta[u] + Mfma16<double>::drow(lane, r), col = 16 * tb[u] + (lane & 15);
                if (row == col && row < n) gdiag[row] = acc[u][r];
            }
        }
    }
    __syncthreads();
    int bad = 0;
#pragma unroll
    for (int u = 0; u < kTilesPerWave; ++u) {
        if (ta[u] >= 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ta[u] + Mfma16<double>::drow(lane, r), col = 16 * tb[u] + (lane & 15);
                if (row < col && col < n) {
                    const double g = acc[u][r];
                    if (g * g > tol2 * gdiag[row] * gdiag[col]) bad = 1;
                }
            }
        }
    }
    return !__syncthreads_or(bad);
}

// Jacobi sweeps in the arithmetic of R on W (n x n, column-major, ld n),
// the rotations accumulated in X, until a sweep has no pair with
// |w_p . w_q| > stop |w_p| |w_q| or max_sweeps have run; returns whether the
// last sweep had none.  Pair (p, q) is rotated when |w_p . w_q| > tol |w_p|
// |w_q| (squared: no square roots on a round's critical path).  stop = tol
// runs to convergence; a larger stop ends once the largest rotation left is
// below it (the first sweeps of a mixed-precision SVD: the FP64 sweeps that
// follow finish the work).  With gdiag (FP64, W in LDS) a sweep whose
// rotations were all below 1e-4 is followed by gram_converged, which ends
// the sweeps when a further sweep would rotate nothing.  Parallel tournament ordering:
// n2/2 disjoint pairs per round, 8 lanes per pair, one barrier per round.
// All threads of the block call it.
// KPER: column entries per lane (n <= 8 KPER) in the register path.
template <typename R, int MODE, int KPER>
__device__ __forceinline__ bool jacobi_sweeps_k(R *W, R *X, int n, R tol2, R stop2, int max_sweeps, int &sweeps,
                                                double *gdiag) {
    const int tid = threadIdx.x;
    const int n2 = n + (n & 1), n1 = n2 - 1, npairs = n2 >> 1;
    const int grp = tid >> 3, gl = tid & 7;
    sweeps = 0;
    bool converged = (n1 == 0);
    while (!converged && sweeps < max_sweeps) {
        ++sweeps;
        int rot = 0, far = 0;
        // players of pair grp in round 0 (tournament_pair), advanced by one
        // place per round: no division on the round's path
        int pa = (grp == 0) ? 0 : 1 + grp, pb = (grp == 0) ? 1 : 1 + n1 - grp;
        for (int r = 0; r < n1; ++r) {
            if constexpr (MODE == kWide) {
                for (int pr = grp; pr < npairs; pr += kThreads / 8) {
                    int p, q;
                    tournament_pair(pr, r, n1, p, q);
                    if (q < n && jacobi_pair_strided(W + (int64_t)p * n, W + (int64_t)q * n, X + (int64_t)p * n,
                                                     X + (int64_t)q * n, n, gl, tol2))
                        rot = 1;
                }
            } else if (grp < npairs) {
                const int p = pa < pb ? pa : pb, q = pa < pb ? pb : pa;
                if (grp != 0) pa = (pa == n1) ? 1 : pa + 1;
                pb = (pb == n1) ? 1 : pb + 1;
                if (q < n) {
                    R *wp = W + p * n, *wq = W + q * n;
                    R *vp = X + p * n, *vq = X + q * n;
                    // the lane's entries i = gl + 8 u of both columns, loaded
                    // together (indices clamped into the column, zero past n,
                    // which adds nothing below) and kept for the rotation.
                    // FP32 loads the X columns with them (registers to
                    // spare), FP64 only once the pair is known to rotate.
                    constexpr int kPer = KPER;
                    constexpr bool kEarlyX = sizeof(R) == 4;
                    R xw[kPer], yw[kPer], xv[kPer], yv[kPer];
#pragma unroll
                    for (int u = 0; u < kPer; ++u) {
                        const int i = gl + 8 * u, ic = i < n ? i : n - 1;
                        xw[u] = wp[ic];
                        yw[u] = wq[ic];
                        if (kEarlyX) {
                            xv[u] = vp[ic];
                            yv[u] = vq[ic];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kPer; ++u) {
                        if (gl + 8 * u >= n) xw[u] = yw[u] = R(0);
                    }
                    R al = R(0), be = R(0), ga = R(0);
#pragma unroll
                    for (int u = 0; u < kPer; ++u) {
                        al = fma(xw[u], xw[u], al);
                        be = fma(yw[u], yw[u], be);
                        ga = fma(xw[u], yw[u], ga);
                    }
                    al = group8_sum(al);
                    be = group8_sum(be);
                    ga = group8_sum(ga);
                    if constexpr (sizeof(R) == 4) {
                        // scale the pair's sums to max(al, be) = 1 (the angle
                        // and the test are scale-free), so that neither the
                        // test nor h in the angle underflows for the tiny
                        // columns of a rank-deficient matrix; pairs of two
                        // columns below the FP32 normal range are left to FP64
                        const R mab = al > be ? al : be;
                        const R sc = mab >= FLT_MIN ? __builtin_amdgcn_rcpf(mab) : R(0);
                        al *= sc;
                        be *= sc;
                        ga *= sc;
                    }
                    if (ga * ga > stop2 * al * be) rot = 1;
                    if (ga * ga > R(1e-8) * al * be) far = 1;  // a rotation above 1e-4
                    if (ga * ga > tol2 * al * be) {
                        R cs, sn;
                        jacobi_angle(al, be, ga, cs, sn);
                        if (!kEarlyX) {
#pragma unroll
                            for (int u = 0; u < kPer; ++u) {
                                const int i = gl + 8 * u, ic = i < n ? i : n - 1;
                                xv[u] = vp[ic];
                                yv[u] = vq[ic];
                            }
                        }
#pragma unroll
                        for (int u = 0; u < kPer; ++u) {
                            const int i = gl + 8 * u;
                            if (i < n) {
                                wp[i] = cs * xw[u] - sn * yw[u];
                                wq[i] = sn * xw[u] + cs * yw[u];
                                vp[i] = cs * xv[u] - sn * yv[u];
                                vq[i] = sn * xv[u] + cs * yv[u];
                            }
                        }
                    }
                }
            }
            if (r + 1 < n1) __syncthreads();
        }
        converged = !__syncthreads_or(rot);
        if constexpr (sizeof(R) == 8 && MODE != kWide) {
            // once the largest rotation of a sweep is below 1e-4 the next
            // sweep is likely to rotate nothing (quadratic convergence):
            // test that on the Gram matrix instead of sweeping
            if (!converged && gdiag != nullptr && !__syncthreads_or(far))
                converged = gram_converged(W, n, tol2, gdiag);
        }
    }
    return converged;
}

template <typename R, int MODE>
__device__ __forceinline__ bool jacobi_sweeps(R *W, R *X, int n, R tol2, R stop2, int max_sweeps, int &sweeps,
                                              double *gdiag = nullptr) {
    if constexpr (MODE != kWide) {
        if (n <= 32) return jacobi_sweeps_k<R, MODE, 4>(W, X, n, tol2, stop2, max_sweeps, sweeps, gdiag);
        if (n <= 48) return jacobi_sweeps_k<R, MODE, 6>(W, X, n, tol2, stop2, max_sweeps, sweeps, gdiag);
    }
    return jacobi_sweeps_k<R, MODE, kMaxN / 8>(W, X, n, tol2, stop2, max_sweeps, sweeps, gdiag);
}

// FP32 first sweeps for float data (n <= 64 held in LDS): on entry W holds
// the FP64 n x n matrix W0 to be swept and X (n x n doubles) is free; on
// return W = W0 X and X holds the FP32 sweeps' rotations, made orthogonal
// to FP64 level, ready for the FP64 sweeps.
//   * W0 is scaled by a power of two to max |w| in [1, 2) and rounded to
//     FP32 (W32) in the first half of X's space, X32 = I in the second;
//     the FP32 sweeps rotate the pairs above the FP32 tolerance n u32
//     until a sweep's largest rotation measure is below 1e-2 (rotations
//     do not depend on the scale).
//   * X32 is orthogonal to about n u32.  One Newton-Schulz step
//     X = X32 - F X32 / 2 with F = X32 X32^T - I (FP64 sums; F ~ 1e-6 is
//     kept in FP32 over W32, which adds ~1e-13) brings that to ~1e-11.
//   * W = W0 X in FP64, so that W = W0 X holds to FP64 rounding: the FP64
//     sweeps then finish in about two sweeps and the Gram test.
// Lanes run over the row index i of each product, so one operand is read
// along a column (consecutive addresses) and the other is the same for the
// whole wave (broadcast).  Returns the FP32 sweep count (0 when max |W0| is
// zero or not finite: the FP64 sweeps then start from W0 and X = I).
template <int MODE>
__device__ int fp32_first_sweeps(double *W, double *X, int n, bool lower, double *red) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nn = n * n;
    double mx = 0.0;
    for (int e = tid; e < nn; e += kThreads) mx = fmax(mx, fabs(W[e]));
    mx = wave_max(mx);
    if (lane == 0) red[wave] = mx;
    __syncthreads();
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    const bool run = mx > 0.0 && mx <= DBL_MAX;  // uniform
    float *W32 = reinterpret_cast<float *>(X), *X32 = W32 + nn;
    int s32 = 0;
    if (run) {
        const double sc = ldexp(1.0, -ilogb(mx));
        for (int e = tid; e < nn; e += kThreads) {
            W32[e] = (float)(W[e] * sc);
            X32[e] = (e % n == e / n) ? 1.0f : 0.0f;
        }
        __syncthreads();
        // FP32 sweeps until the largest rotation measure |w_p . w_q| /
        // (|w_p| |w_q|) of a sweep is below 1e-2: the FP64 sweeps that follow
        // take it from there in as many sweeps as from the FP32-converged
        // state (the last FP32 sweeps before convergence are saved)
        const float tol = (float)(n > 2 ? n : 2) * FLT_EPSILON;
        jacobi_sweeps<float, MODE>(W32, X32, n, tol * tol, 1e-4f, kMaxSweeps, s32);
        float *F = W32;  // W32 is not needed any more
        for (int e = tid; e < nn; e += kThreads) {
            const int i = e % n, j = e / n;
            double acc = (i == j) ? -1.0 : 0.0;
            for (int k = 0; k < n; ++k) acc = fma((double)X32[i + k * n], (double)X32[j + k * n], acc);
            F[e] = (float)acc;
        }
        __syncthreads();
    }
    constexpr int kR = (kMaxN * kMaxN + kThreads - 1) / kThreads;
    double xr[kR];
#pragma unroll
    for (int t = 0; t < kR; ++t) {
        const int e = tid + kThreads * t, i = e % n, j = e / n;
        if (e < nn) {
            if (run) {
                double acc = (double)X32[e];
                for (int k = 0; k < n; ++k) acc = fma(-0.5 * (double)W32[i + k * n], (double)X32[k + j * n], acc);
                xr[t] = acc;
            } else {
                xr[t] = (i == j) ? 1.0 : 0.0;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kR; ++t) {
        const int e = tid + kThreads * t;
        if (e < nn) X[e] = xr[t];
    }
    __syncthreads();
    if (run) {
        // W0 is triangular: lower (R^T) or upper (R)
#pragma unroll
        for (int t = 0; t < kR; ++t) {
            const int e = tid + kThreads * t, i = e % n, j = e / n;
            if (e < nn) {
                const int k0 = lower ? 0 : i, k1 = lower ? i + 1 : n;
                double acc = 0.0;
                for (int k = k0; k < k1; ++k) acc = fma(W[i + k * n], X[k + j * n], acc);
                xr[t] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kR; ++t) {
            const int e = tid + kThreads * t;
            if (e < nn) W[e] = xr[t];
        }
        __syncthreads();
    }
    return s32;
}

template <typename T, int MODE, bool MIXED>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void gesvd_kernel(int m, int n, const T *__restrict__ A, int64_t lda,
                                                         int64_t strideA, T *S, int64_t strideS, T *U,
                                                         int64_t ldu, int64_t strideU, T *VT, int64_t ldvt,
                                                         int64_t strideVT, double *work, int *info,
                                                         int *sweeps_out, int64_t batch) {
    extern __shared__ double lds[];
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const Layout L = layout(m, n, MODE);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *ws = work + b * L.wstride;  // this matrix's workspace (unused by kResident)
    double *G = (MODE == kResident || MODE == kParked) ? lds + L.g : ws + L.g;
    double *W = MODE == kWide ? ws + L.w : lds + L.w;
    double *X = MODE == kWide ? ws + L.x : lds + L.x;
    double *sig = lds + L.sig;
    double *tau = lds + L.tau;
    double *s_tau = tau + n;
    double *sgnv = lds + L.sgn;
    int *pos = reinterpret_cast<int *>(lds + L.ints);
    int *ord = pos + n;
    int *s_nz = ord + n;

    // 1. stage and factorise
    const T *a = A + b * strideA;
    for (int e = tid; e < m * n; e += kThreads) {
        const int i = e % m, j = e / m;
        G[i + (int64_t)j * m] = (double)a[i + (int64_t)j * lda];
    }
    __syncthreads();
    if (m <= 16 * 8 + 1) hh::hqr_factor<double, double, double, hh::kWaves, 8>(G, m, m, n, tau, s_tau);
    else hh::hqr_factor<double, double>(G, m, m, n, tau, s_tau);
    // VT == NULL (no right singular vectors): the sweeps run on W = R^T, so
    // that the accumulated rotations X are the left singular vectors of R
    // (R^T Z = V_R Sigma  =>  R = Z Sigma V_R^T): no completion is needed for
    // zero sigma, and on this matrix type (the nullspace builds' A^T, rank
    // deficient) they converge in fewer sweeps.
    const bool on_rt = (VT == nullptr);
    if constexpr (MODE == kParked) {
        // W and V take over G's LDS: R goes through registers and, when U
        // is wanted, the reflectors to the workspace (read again for U only)
        constexpr int kR = (kMaxN * kMaxN + kThreads - 1) / kThreads;
        double rv[kR];
#pragma unroll
        for (int t = 0; t < kR; ++t) {
            const int e = tid + kThreads * t, i = e % n, j = e / n;
            rv[t] = 0.0;
            if (e < n * n) {
                if (on_rt) rv[t] = (j <= i) ? G[j + (int64_t)i * m] : 0.0;
                else rv[t] = (i <= j) ? G[i + (int64_t)j * m] : 0.0;
            }
        }
        if (U != nullptr)
            for (int e = tid; e < m * n; e += kThreads) ws[e] = G[e];
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kR; ++t) {
            const int e = tid + kThreads * t, i = e % n, j = e / n;
            if (e < n * n) {
                W[e] = rv[t];
                X[e] = (i == j) ? 1.0 : 0.0;
            }
        }
        G = ws;
    } else {
        for (int e = tid; e < n * n; e += kThreads) {
            const int i = e % n, j = e / n;
            if (on_rt) W[e] = (j <= i) ? G[j + (int64_t)i * m] : 0.0;
            else W[e] = (i <= j) ? G[i + (int64_t)j * m] : 0.0;
            X[e] = (i == j) ? 1.0 : 0.0;
        }
    }
    __syncthreads();

    // 2. one-sided Jacobi on W, V accumulated in X: for float data first in
    // FP32 (MIXED), then in FP64 to the FP64 tolerance
    int s32 = 0, s64 = 0;
    if constexpr (MIXED) {
        if (n >= 4) s32 = fp32_first_sweeps<MODE>(W, X, n, on_rt, sig);  // sig[0..3]: wave maxima
    }
    const double tol = (double)(n > 2 ? n : 2) * DBL_EPSILON;
    // sig (n doubles, free until the singular values) carries the Gram
    // diagonal of the convergence test
    const bool converged = jacobi_sweeps<double, MODE>(W, X, n, tol * tol, tol * tol, kMaxSweeps, s64, sig);
    __syncthreads();

    // 3. singular values, order, signs
    const int grp = tid >> 3, gl = tid & 7;
    for (int j = grp; j < n; j += kThreads / 8) {
        double s = 0.0;
        for (int i = gl; i < n; i += 8) s = fma(W[i + j * n], W[i + j * n], s);
        s = group8_sum(s);
        if (gl == 0) sig[j] = sqrt(s);
    }
    __syncthreads();
    for (int j = tid; j < n; j += kThreads) {
        const double sj = sig[j];
        double smax = 0.0;
        int ps = 0;
        for (int k = 0; k < n; ++k) {
            const double sk = sig[k];
            smax = sk > smax ? sk : smax;
            ps += (sk > sj || (sk == sj && k < j)) ? 1 : 0;
        }
        pos[j] = ps;
        ord[ps] = j;
        int imax = 0;
        double vmax = fabs(X[j * n]);
        for (int i = 1; i < n; ++i) {
            const double v = fabs(X[i + j * n]);
            if (v > vmax) {
                vmax = v;
                imax = i;
            }
        }
        // sigma_j <= zt counts as zero: its v_j keeps the accumulated sign
        // and its u_j comes from the completion below
        const double zt = smax * (double)m * DBL_EPSILON;
        sgnv[j] = (sj > zt && X[imax + j * n] > 0.0) ? -1.0 : 1.0;
        if (j == 0) {
            int nz = 0;
            for (int k = 0; k < n; ++k) nz += (sig[k] > zt) ? 1 : 0;
            *s_nz = nz;
        }
    }
    __syncthreads();
    T *s_out = S + b * strideS;
    for (int c = tid; c < n; c += kThreads) s_out[c] = (T)sig[ord[c]];
    if (!on_rt) {
        T *vt = VT + b * strideVT;
        for (int e = tid; e < n * n; e += kThreads) {
            const int rr = e % n, i = e / n, j = ord[rr];
            vt[rr + (int64_t)i * ldvt] = (T)(sgnv[j] * X[i + j * n]);
        }
    }
    if (tid == 0) {
        info[b] = converged ? 0 : 1;
        if (sweeps_out) sweeps_out[b] = s64 | (s32 << 16);
    }
    if (U == nullptr) return;
    __syncthreads();

    // U_R in output order (X is free once VT is written)
    const int nz = on_rt ? n : *s_nz;
    if (on_rt) {
        // U_R = the rotations, columns reordered (through W, which is free)
        for (int e = tid; e < n * n; e += kThreads) {
            const int i = e % n, c = e / n, j = ord[c];
            W[e] = sgnv[j] * X[i + j * n];
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += kThreads) X[e] = W[e];
    } else {
        for (int e = tid; e < n * n; e += kThreads) {
            const int i = e % n, c = e / n, j = ord[c];
            X[e] = (c < nz) ? sgnv[j] * W[i + j * n] / sig[j] : 0.0;
        }
    }
    __syncthreads();
    if (MODE == kWide) {
        if (wave == 0 && nz < n) complete_basis_strided(X, n, nz, lane);
    } else if (wave == 0 && nz < n) {
        int next = 0;
        for (int col = nz; col < n; ++col) {
            double best_x = 0.0, best_norm = -1.0;
            int best_k = next;
            for (int k = next; k < n; ++k) {
                double x = (lane == k) ? 1.0 : 0.0;
                for (int pass = 0; pass < 2; ++pass) {
                    for (int c = 0; c < col; ++c) {
                        const double uc = (lane < n) ? X[lane + c * n] : 0.0;
                        const double coef = wave_sum(uc * x);
                        x -= coef * uc;
                    }
                }
                const double nrm = sqrt(wave_sum(x * x));
                if (nrm > best_norm) {
                    best_norm = nrm;
                    best_x = x;
                    best_k = k;
                }
                if (nrm > 0.5) break;
            }
            if (lane < n) X[lane + col * n] = best_x / best_norm;
            wave_sync();
            if (best_norm > 0.5) next = best_k + 1;
        }
    }
    __syncthreads();

    // 4. U = Q diag(U_R, I)
    const int g16 = tid >> 4, l16 = tid & 15;
    T *u = U + b * strideU;
    if (m <= 128) {
        // 64 columns at a time in registers (one pass for m <= 64, two up
        // to 128): 16-lane group g16 holds columns c0 + g16 + 16 q, lane l16
        // rows l16 + 16 k; each reflector is read once per pass (rows
        // clamped into the column, masked to [0..0, 1, v]) and its four
        // column updates reduce together
        constexpr int kQ = 4, kK = 8;
        for (int c0 = 0; c0 < m; c0 += 16 * kQ) {
        double xr[kQ][kK];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int cc = c0 + g16 + 16 * q;
#pragma unroll
            for (int k = 0; k < kK; ++k) {
                const int i = l16 + 16 * k;
                const int xi = (i < n && cc < n) ? i + cc * n : 0;
                const double xv = X[xi];
                xr[q][k] = (cc < n) ? ((i < n) ? xv : 0.0) : ((i == cc) ? 1.0 : 0.0);
            }
        }
        for (int j = n - 1; j >= 0; --j) {
            const double t = tau[j];
            if (t == 0.0) continue;  // uniform
            const double *vcol = G + (int64_t)j * m;
            double vr[kK];
#pragma unroll
            for (int k = 0; k < kK; ++k) {
                const int i = l16 + 16 * k;
                const double e = vcol[i < m ? i : m - 1];
                vr[k] = (i > j && i < m) ? e : ((i == j) ? 1.0 : 0.0);
            }
            double w[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                w[q] = 0.0;
#pragma unroll
                for (int k = 0; k < kK; ++k) w[q] = fma(vr[k], xr[q][k], w[q]);
            }
#pragma unroll
            for (int q = 0; q < kQ; ++q) w[q] = group16_sum(w[q]);
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const double tw = t * w[q];
#pragma unroll
                for (int k = 0; k < kK; ++k) xr[q][k] = fma(-tw, vr[k], xr[q][k]);
            }
        }
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int cc = c0 + g16 + 16 * q;
#pragma unroll
            for (int k = 0; k < kK; ++k) {
                const int i = l16 + 16 * k;
                if (cc < m && i < m) u[i + (int64_t)cc * ldu] = (T)xr[q][k];
            }
        }
        }
        return;
    }
    // larger m: chunks of columns in W, one 16-lane group each
    for (int c0 = 0; c0 < m; c0 += L.ch) {
        const int cc = c0 + g16;
        if (g16 < L.ch && cc < m) {
            double *xc = W + g16 * m;
            for (int i = l16; i < m; i += 16)
                xc[i] = (cc < n) ? (i < n ? X[i + cc * n] : 0.0) : (i == cc ? 1.0 : 0.0);
            wave_sync();
            for (int j = n - 1; j >= 0; --j) {
                const double t = tau[j];
                if (t == 0.0) continue;
                const double *v = G + (int64_t)j * m;
                double w = 0.0;
                for (int i = j + 1 + l16; i < m; i += 16) w = fma(v[i], xc[i], w);
                w = group16_sum(w) + xc[j];
                const double tw = t * w;
                for (int i = j + 1 + l16; i < m; i += 16) xc[i] -= tw * v[i];
                wave_sync();
                if (l16 == 0) xc[j] -= tw;
                wave_sync();
            }
            for (int i = l16; i < m; i += 16) u[i + (int64_t)cc * ldu] = (T)xc[i];
        }
    }
}

template <typename T>
__global__ void rank_kernel(int n, const T *__restrict__ S, int64_t strideS, T eps, unsigned *rank,
                            int64_t batch) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    unsigned r = 0;
    for (int i = 0; i < n; ++i) r += (S[b * strideS + i] > eps) ? 1u : 0u;
    rank[b] = r;
}

// N_b[:, 0:nul] = U_b[:, r:n] (other columns zero), P_b = N_b N_b^T with
// nul = n - rank_b.  P is computed on the lower tile triangle and mirrored.
constexpr int kNbKc = 16;

template <typename T>
__global__ __launch_bounds__(kThreads) void nullspace_build_kernel(int n, const T *__restrict__ U, int64_t ldu,
                                                                   int64_t strideU, const unsigned *rank, T *N,
                                                                   int64_t ldn, int64_t strideN, T *P,
                                                                   int64_t ldp, int64_t strideP, int64_t batch) {
    extern __shared__ unsigned char dyn_raw[];
    T *Ns = reinterpret_cast<T *>(dyn_raw);  // kNbKc columns of N, ld n
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int tid = threadIdx.x;
    const unsigned rk = rank[b];
    const int r = (int)(rk < (unsigned)n ? rk : (unsigned)n);
    const int nul = n - r;
    const T *ub = U + b * strideU;
    T *nb = N + b * strideN;
    T *pb = P + b * strideP;
    for (int e = tid; e < n * n; e += kThreads) {
        const int i = e % n, c = e / n;
        nb[i + (int64_t)c * ldn] = (c < nul) ? ub[i + (int64_t)(r + c) * ldu] : T(0);
    }
    const int nt = (n + 3) / 4;
    const int ntiles = nt * (nt + 1) / 2;
    for (int t0 = 0; t0 < ntiles; t0 += kThreads) {
        const int t = t0 + tid;
        // tile t -> (ti, tj) with ti >= tj
        int ti = 0, tj = 0;
        if (t < ntiles) {
            ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            while (ti * (ti + 1) / 2 > t) --ti;
            tj = t - ti * (ti + 1) / 2;
        }
        T acc[4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = T(0);
        for (int k0 = 0; k0 < nul; k0 += kNbKc) {
            const int kc = (nul - k0) < kNbKc ? (nul - k0) : kNbKc;
            __syncthreads();
            for (int e = tid; e < kc * n; e += kThreads) {
                const int i = e % n, c = e / n;
                Ns[e] = ub[i + (int64_t)(r + k0 + c) * ldu];
            }
            __syncthreads();
            if (t < ntiles) {
                for (int k = 0; k < kc; ++k) {
                    T ri[4], rj[4];
#pragma unroll
                    for (int x = 0; x < 4; ++x) {
                        const int i = 4 * ti + x, j = 4 * tj + x;
                        ri[x] = (i < n) ? Ns[i + k * n] : T(0);
                        rj[x] = (j < n) ? Ns[j + k * n] : T(0);
                    }
#pragma unroll
                    for (int x = 0; x < 4; ++x)
#pragma unroll
                        for (int y = 0; y < 4; ++y) acc[x][y] = fma(ri[x], rj[y], acc[x][y]);
                }
            }
        }
        if (t < ntiles) {
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int i = 4 * ti + x, j = 4 * tj + y;
                    if (i < n && j < n) {
                        pb[i + (int64_t)j * ldp] = acc[x][y];
                        pb[j + (int64_t)i * ldp] = acc[x][y];
                    }
                }
        }
    }
}

// The same outputs with P = N N^T on the matrix cores: one workgroup per
// matrix, the nullity columns of U staged once in LDS (np = 16 ceil(n / 16)
// rows and kp = 4 ceil(q / 4) columns, the padding zero, so every operand
// read is unconditional) while N is written; wavefront w then computes the
// 16 x 16 tiles w, w + 4, ... of P, each by q / 4 16x16x4 MFMAs whose A and B
// operands are rows of the staged block (A[i][k] = N[16 ti + i][k],
// B[k][j] = N[16 tj + j][k]).  Every tile of P is formed (the MFMA work is
// small next to the writes of N and P), stored with one 16-byte store per
// lane when the rows of a lane are consecutive and aligned (FP32, ldp and
// strideP multiples of 4), else element by element.
template <typename T>
__global__ __launch_bounds__(kThreads) void nullspace_build_mfma_kernel(int n, const T *__restrict__ U, int64_t ldu,
                                                                        int64_t strideU, const unsigned *rank, T *N,
                                                                        int64_t ldn, int64_t strideN, T *P,
                                                                        int64_t ldp, int64_t strideP, int64_t batch,
                                                                        int vec4) {
    extern __shared__ unsigned char dyn_raw[];
    T *Ns = reinterpret_cast<T *>(dyn_raw);  // kp columns of np rows, ld np
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned rk = rank[b];
    const int r = (int)(rk < (unsigned)n ? rk : (unsigned)n);
    const int nul = n - r;
    const int nt = (n + 15) >> 4, np = 16 * nt, kp = (nul + 3) & ~3;
    const int ncol = n > kp ? n : kp;
    const T *ub = U + b * strideU;
    T *nb = N + b * strideN;
    T *pb = P + b * strideP;
    // eight elements per thread at a time, their loads issued together.
    // Column of element e: floor((e + 1/2) / np) in FP32, exact for
    // e < 2^16 and np a multiple of 16 up to 512 (the LDS bound keeps
    // np ncol below 2^16): (e + 1/2) / np lies at least 1 / (2 np) from an
    // integer and the float quotient within 2^-17 of it.
    constexpr int kU = 8;
    const int total = np * ncol;
    const float inv_np = 1.0f / (float)np;
    auto col_of = [&](int e) { return (int)(((float)e + 0.5f) * inv_np); };
    for (int e0 = tid; e0 < total; e0 += kU * kThreads) {
        T v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = e0 + u * kThreads, c = col_of(e), i = e - c * np;
            v[u] = (e < total && i < n && c < nul) ? ub[i + (int64_t)(r + c) * ldu] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = e0 + u * kThreads, c = col_of(e), i = e - c * np;
            if (e < total) {
                if (i < n && c < n) nb[i + (int64_t)c * ldn] = v[u];
                if (c < kp) Ns[e] = v[u];
            }
        }
    }
    __syncthreads();
    using M = Mfma16<T>;
    const int col_in_tile = lane & 15, kq = lane >> 4;
    for (int t = wave; t < nt * nt; t += kThreads / kWave) {
        const int ti = t / nt, tj = t % nt;
        const T *as = Ns + 16 * ti + col_in_tile + kq * np;
        const T *bs = Ns + 16 * tj + col_in_tile + kq * np;
        typename M::acc_t acc = M::zero();
        for (int k0 = 0; k0 < kp; k0 += 4) acc = M::fma(as[(int64_t)k0 * np], bs[(int64_t)k0 * np], acc);
        const int col = 16 * tj + col_in_tile;
        if constexpr (sizeof(T) == 4) {
            const int row0 = 16 * ti + 4 * kq;  // rows row0 .. row0 + 3 (f32 layout)
            if (vec4 && row0 + 3 < n && col < n) {
                *reinterpret_cast<f32x4 *>(pb + row0 + (int64_t)col * ldp) = acc;
                continue;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * ti + M::drow(lane, q);
            if (row < n && col < n) pb[row + (int64_t)col * ldp] = acc[q];
        }
    }
}

// Where a matrix of m x n lives: in LDS when it fits, G in the workspace
// next, everything in the workspace for n > kMaxN or very tall matrices.
inline int gesvd_mode(int m, int n) {
    if (n <= kMaxN) {
        // the most workgroups per CU (LDS-bound; resident on ties)
        const int res = layout(m, n, kResident).bytes, park = layout(m, n, kParked).bytes;
        const int per_res = res <= kMaxLds ? kMaxLds / res : 0, per_park = park <= kMaxLds ? kMaxLds / park : 0;
        if (per_res > 0 && per_res >= per_park) return kResident;
        if (per_park > 0) return kParked;
        if (layout(m, n, kGlobalG).bytes <= kMaxLds) return kGlobalG;
    }
    return kWide;
}

// FP32 first sweeps for float data (gputils_gesvd_set_fp32_sweeps)
std::atomic<int> g_fp32_sweeps{1};

template <typename T, int MODE>
int launch_gesvd(int m, int n, const T *A, int64_t lda, int64_t strideA, T *S, int64_t strideS, T *U,
                 int64_t ldu, int64_t strideU, T *VT, int64_t ldvt, int64_t strideVT, double *work, int *info,
                 int *sweeps, int64_t batch, hipStream_t st) {
    const Layout L = layout(m, n, MODE);
    // FP32 first sweeps: float data, W and V in LDS, and the sweeps on R^T
    // (VT == NULL).  On R (VT wanted) the FP64 sweeps would have to
    // orthogonalise the null cluster of a rank-deficient matrix from scratch
    // after the FP32 ones, which costs more than it saves.
    constexpr bool kCanMix = std::is_same<T, float>::value && MODE != kWide;
    const bool mix = kCanMix && VT == nullptr && g_fp32_sweeps.load(std::memory_order_relaxed) != 0;
    auto kern = mix ? gesvd_kernel<T, MODE, kCanMix> : gesvd_kernel<T, MODE, false>;
    if (L.bytes > 64 * 1024)
        GPUTILS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, L.bytes));
    hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kThreads), L.bytes, st, m, n, A, lda, strideA, S,
                       strideS, U, ldu, strideU, VT, ldvt, strideVT, work, info, sweeps, batch);
    return launch_check("gesvd_kernel");
}

template <typename T>
int gesvd_batched(int m, int n, const T *A, int64_t lda, int64_t strideA, T *S, int64_t strideS, T *U,
                  int64_t ldu, int64_t strideU, T *VT, int64_t ldvt, int64_t strideVT, void *work, int *info,
                  int *sweeps, int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 1, 2, "n must be >= 1");
    GPUTILS_ARG(m >= n, 1, "m must be >= n (tall matrices; transpose a fat one first)");
    GPUTILS_ARG(lda >= m, 4, "lda < m");
    GPUTILS_ARG(strideA >= 0, 5, "negative strideA");
    GPUTILS_ARG(strideS >= 0, 7, "negative strideS");
    GPUTILS_ARG(U == nullptr || ldu >= m, 9, "ldu < m");
    GPUTILS_ARG(strideU >= 0, 10, "negative strideU");
    GPUTILS_ARG(VT == nullptr || ldvt >= n, 12, "ldvt < n");
    GPUTILS_ARG(strideVT >= 0, 13, "negative strideVT");
    GPUTILS_ARG(info != nullptr, 15, "info is NULL");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 17, "batch out of range");
    if (batch == 0) return 0;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(S != nullptr, 6, "S is NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int mode = gesvd_mode(m, n);
    if (mode == kResident)
        return launch_gesvd<T, kResident>(m, n, A, lda, strideA, S, strideS, U, ldu, strideU, VT, ldvt, strideVT,
                                          nullptr, info, sweeps, batch, st);
    GPUTILS_ARG(layout(m, n, mode).bytes <= kMaxLds, 1, "m too large for the batched SVD");
    GPUTILS_ARG(work != nullptr, 14, "work is NULL (query gputils_gesvd_workspace_bytes)");
    if (mode == kParked)
        return launch_gesvd<T, kParked>(m, n, A, lda, strideA, S, strideS, U, ldu, strideU, VT, ldvt, strideVT,
                                        static_cast<double *>(work), info, sweeps, batch, st);
    if (mode == kGlobalG)
        return launch_gesvd<T, kGlobalG>(m, n, A, lda, strideA, S, strideS, U, ldu, strideU, VT, ldvt, strideVT,
                                         static_cast<double *>(work), info, sweeps, batch, st);
    return launch_gesvd<T, kWide>(m, n, A, lda, strideA, S, strideS, U, ldu, strideU, VT, ldvt, strideVT,
                                  static_cast<double *>(work), info, sweeps, batch, st);
}

template <typename T>
int rank_batched(int n, const T *S, int64_t strideS, T eps, unsigned *rank, int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "negative n");
    GPUTILS_ARG(strideS >= 0, 3, "negative strideS");
    GPUTILS_ARG(batch >= 0, 6, "negative batch");
    if (batch == 0) return 0;
    GPUTILS_ARG(S != nullptr, 2, "S is NULL");
    GPUTILS_ARG(rank != nullptr, 5, "rank is NULL");
    const unsigned blocks = (unsigned)((batch + 255) / 256);
    hipLaunchKernelGGL(rank_kernel<T>, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), n, S, strideS,
                       eps, rank, batch);
    return launch_check("rank_kernel");
}

// b_i <- N_i (N_i^T b_i), N_i's first q_i = n - rank_i columns: the
// projection P_i b_i of Nullspace::project (P = N N^T) formed from the
// n x q_i block of N instead of the n x n P, so a call reads n q_i matrix
// elements instead of n^2.  One workgroup per matrix; rank_i >= n gives
// y = 0, as P = 0 does.
//
// Three paths, chosen by the host: the lane-group path below (n = 8 V RV,
// RV <= 4), else the vector path, else the generic one.
//
// Vector path (vec set by the host: n, ldn, strideN, strideB multiples of
// V = 16 / sizeof(T) and 16-byte aligned N and b; taken when q_i <= KC G):
// thread t owns the V-row slice r = t % (n/V) of the columns jg, jg + G, ...
// (jg = t / (n/V), G = kThreads / (n/V)) and loads all of them, and its
// slice of b, as 16-byte loads in flight together, so N is read from memory
// once and held in registers for both products.  KC, the columns a thread
// can hold, is the least of 4, 8, 12, 16 that covers q = n (16 when none
// does; a larger q_i then takes the generic path).  z = N^T b: per-slice dot
// products summed over the n/V slices through LDS; y = N z: per-thread
// V-row partial sums added over the G column groups through LDS.
//
// Generic path (any n, ld, alignment): each 16-lane group takes one column
// at a time for z (rows l + 16 k, a four-step DPP sum, b from LDS), then
// S = 256 / n threads per row (a power of two) split the q columns of y and
// add their partial sums through LDS, reading N a second time (from L2).
// Word 3 of a raw (stride 0) buffer resource on gfx9: 32-bit data format,
// default number format, no swizzle.
constexpr int kBufferWord3 = 0x00020000;

template <typename T, int KC>
__device__ __forceinline__ void project_vec(int n, int q, const T *__restrict__ nm, int64_t ldn, T *bm,
                                            unsigned char *lds) {
    constexpr int V = 16 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(V)));
    const int tid = threadIdx.x, rq = n / V, G = kThreads / rq;
    const int r = tid % rq, jg = tid / rq;
    const bool act = jg < G;
    T *part = reinterpret_cast<T *>(lds);  // rq x q dot products, later G x rq vec partial sums
    vec *part2 = reinterpret_cast<vec *>(lds);
    const size_t part_bytes = (size_t)rq * n * sizeof(T) > (size_t)kThreads * 16 ? (size_t)rq * n * sizeof(T)
                                                                                    : (size_t)kThreads * 16;
    T *zs = reinterpret_cast<T *>(lds + part_bytes);
    // One buffer resource per held column slot u, based at column u G and
    // sized to the q - u G columns left from there: a thread's load of
    // column jg + u G is then the shared offset jg ldn + V r of resource u,
    // and a load at or past column q (or any load of an idle thread, sent
    // out of range) returns zeros without touching memory, so all KC loads
    // issue unconditionally.  Base and size are workgroup-uniform;
    // readfirstlane states it, so the resources live in scalar registers.
    const uint64_t base = reinterpret_cast<uintptr_t>(nm);
    const uint64_t ubase = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
    const int uq = __builtin_amdgcn_readfirstlane(q);
    const int off = act ? (int)(((int64_t)jg * ldn + V * r) * (int64_t)sizeof(T)) : 0x7ffffff0;
    vec nv[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u) {
        const int left = uq - u * G;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            reinterpret_cast<void *>(ubase + (uint64_t)((int64_t)u * G * ldn * (int64_t)sizeof(T))), 0,
            left > 0 ? (int)((int64_t)left * ldn * (int64_t)sizeof(T)) : 0, kBufferWord3);
        nv[u] = __builtin_bit_cast(vec, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
    }
    vec bv = {};
    if (act) {
        bv = *reinterpret_cast<const vec *>(bm + V * r);
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int j = jg + u * G;
            if (j < q) {
                const vec p = nv[u] * bv;
                T d = p[0];
#pragma unroll
                for (int v = 1; v < V; ++v) d += p[v];
                part[r * q + j] = d;
            }
        }
    }
    __syncthreads();
    for (int j = tid; j < q; j += kThreads) {
        T z0 = T(0), z1 = T(0);
        int s = 0;
        for (; s + 1 < rq; s += 2) {
            z0 += part[s * q + j];
            z1 += part[(s + 1) * q + j];
        }
        if (s < rq) z0 += part[s * q + j];
        zs[j] = z0 + z1;
    }
    __syncthreads();
    vec acc = {};
    if (act) {
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const int j = jg + u * G;
            if (j < q) acc += nv[u] * zs[j];
        }
        part2[jg * rq + r] = acc;
    }
    __syncthreads();
    if (tid < rq) {
        vec y = part2[tid];
        for (int g = 1; g < G; ++g) y += part2[g * rq + tid];
        *reinterpret_cast<vec *>(bm + V * tid) = y;
    }
}

template <typename T>
__device__ __forceinline__ void project_generic(int n, int q, const T *__restrict__ nm, int64_t ldn, T *bm, int S,
                                                unsigned char *lds) {
    T *bs = reinterpret_cast<T *>(lds);  // b (n)
    T *zs = bs + n;                      // z (n >= q)
    T *ps = zs + n;                      // y partial sums (S x n)
    const int tid = threadIdx.x, g16 = tid >> 4, l16 = tid & 15;
    for (int i = tid; i < n; i += kThreads) bs[i] = bm[i];
    __syncthreads();
    for (int c = g16; c < q; c += kThreads / 16) {
        const T *col = nm + (int64_t)c * ldn;
        T acc = T(0);
#pragma unroll 4
        for (int i = l16; i < n; i += 16) acc = fma(col[i], bs[i], acc);
        acc = hh::row16_sum(acc);
        if (l16 == 0) zs[c] = acc;
    }
    __syncthreads();
    // row i, column slice h of S: columns c = h, h + S, ...
    for (int e = tid; e < n * S; e += kThreads) {
        const int i = e % n, h = e / n;
        T a0 = T(0), a1 = T(0);
        int c = h;
        for (; c + S < q; c += 2 * S) {
            a0 = fma(nm[i + (int64_t)c * ldn], zs[c], a0);
            a1 = fma(nm[i + (int64_t)(c + S) * ldn], zs[c + S], a1);
        }
        if (c < q) a0 = fma(nm[i + (int64_t)c * ldn], zs[c], a0);
        ps[h * n + i] = a0 + a1;
    }
    __syncthreads();
    for (int i = tid; i < n; i += kThreads) {
        T y = T(0);
        for (int h = 0; h < S; ++h) y += ps[h * n + i];
        bm[i] = y;
    }
}

// Lane-group path, n = 8 V RV (V = 16 / sizeof(T), RV = 1..4): each
// column is read by one aligned group of 8 lanes, lane e taking rows
// e RL ... e RL + RL - 1 (RL = V RV) as RV 16-byte loads; the 32 groups of
// the workgroup take columns g, g + 32, ... (KS slots, all n covered), all
// loaded at once through one range-checked buffer resource per slot, as in
// the vector path.  z_j = (N^T b)_j is a three-step DPP sum within the
// group, so z needs no LDS and no barrier.  y = N z: each lane sums its
// rows over its KS columns; the two groups of a 16-lane row are added by a
// row rotation, the 16 rows of the workgroup through LDS after the one
// barrier (4 lanes per output vector, 4 rows each, then a DPP quad sum).
template <typename T, int RV>
__device__ __forceinline__ void project_lanes(int q, const T *__restrict__ nm, int64_t ldn, T *bm,
                                              unsigned char *lds) {
    constexpr int V = 16 / sizeof(T);
    constexpr int n = 8 * V * RV;
    constexpr int RL = V * RV;            // rows per lane
    constexpr int CG = kThreads / 8;      // column groups
    constexpr int KS = (n + CG - 1) / CG;  // column slots
    constexpr int NV = n / V;             // vectors per column
    typedef T vec __attribute__((ext_vector_type(V)));
    const int tid = threadIdx.x, g = tid >> 3, e = tid & 7;
    const uint64_t base = reinterpret_cast<uintptr_t>(nm);
    const uint64_t ubase = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
    const int uq = __builtin_amdgcn_readfirstlane(q);
    const int off = (int)(((int64_t)g * ldn + e * RL) * (int64_t)sizeof(T));
    vec nv[KS][RV];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int left = uq - s * CG;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            reinterpret_cast<void *>(ubase + (uint64_t)((int64_t)s * CG * ldn * (int64_t)sizeof(T))), 0,
            left > 0 ? (int)((int64_t)left * ldn * (int64_t)sizeof(T)) : 0, kBufferWord3);
#pragma unroll
        for (int v = 0; v < RV; ++v)
            nv[s][v] = __builtin_bit_cast(vec, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16 * v, 0, 0));
    }
    vec bv[RV];
#pragma unroll
    for (int v = 0; v < RV; ++v) bv[v] = *reinterpret_cast<const vec *>(bm + e * RL + v * V);
    vec acc[RV];
#pragma unroll
    for (int v = 0; v < RV; ++v) acc[v] = vec{};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        vec p = nv[s][0] * bv[0];
#pragma unroll
        for (int v = 1; v < RV; ++v) p += nv[s][v] * bv[v];
        T d = p[0];
#pragma unroll
        for (int c = 1; c < V; ++c) d += p[c];
        d = hh::row16_plus<0xB1>(d);  // quad_perm [1,0,3,2]
        d = hh::row16_plus<0x4E>(d);  // quad_perm [2,3,0,1]
        d = hh::row16_plus<0x141>(d);  // row_half_mirror: the 8-lane sum
#pragma unroll
        for (int v = 0; v < RV; ++v) acc[v] += nv[s][v] * d;
    }
    vec *part = reinterpret_cast<vec *>(lds);  // 16 rows x NV vectors
#pragma unroll
    for (int v = 0; v < RV; ++v)
#pragma unroll
        for (int c = 0; c < V; ++c) acc[v][c] = hh::row16_plus<0x128>(acc[v][c]);  // row_ror:8
    if ((tid & 8) == 0) {
#pragma unroll
        for (int v = 0; v < RV; ++v) part[(tid >> 4) * NV + e * RV + v] = acc[v];
    }
    __syncthreads();
    if (tid < 4 * NV) {  // whole quads: NV vectors x 4 lanes
        const int o = tid >> 2, h = tid & 3;
        vec y = part[(4 * h) * NV + o] + part[(4 * h + 1) * NV + o] + part[(4 * h + 2) * NV + o] +
                part[(4 * h + 3) * NV + o];
#pragma unroll
        for (int c = 0; c < V; ++c) {
            y[c] = hh::row16_plus<0xB1>(y[c]);
            y[c] = hh::row16_plus<0x4E>(y[c]);
        }
        if (h == 0) *reinterpret_cast<vec *>(bm + o * V) = y;
    }
}

template <typename T, int KC, int RV>
__global__ __launch_bounds__(kThreads) void nullspace_project_kernel(int n, const T *__restrict__ N, int64_t ldn,
                                                                     int64_t strideN, const unsigned *__restrict__ rank,
                                                                     T *b, int64_t strideB, int64_t batch, int S,
                                                                     int vec) {
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn_raw[];
    const int64_t m = blockIdx.x + (int64_t)blockIdx.y * 65535;
    if (m >= batch) return;  // uniform per workgroup
    const unsigned rk = rank[m];
    const int q = n - (int)(rk < (unsigned)n ? rk : (unsigned)n);
    const T *nm = N + m * strideN;
    T *bm = b + m * strideB;
    if constexpr (RV > 0) {
        project_lanes<T, RV>(q, nm, ldn, bm, dyn_raw);
        return;
    }
    if constexpr (KC > 0) {
        if (vec && q <= KC * (kThreads / (n / (16 / (int)sizeof(T))))) {
            project_vec<T, KC>(n, q, nm, ldn, bm, dyn_raw);
            return;
        }
    }
    project_generic<T>(n, q, nm, ldn, bm, S, dyn_raw);
}

template <typename T>
int nullspace_project(int n, const T *N, int64_t ldn, int64_t strideN, const unsigned *rank, T *b, int64_t strideB,
                      int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 1, 1, "n must be >= 1");
    GPUTILS_ARG(ldn >= n, 3, "ldn < n");
    GPUTILS_ARG(strideN >= 0, 4, "negative strideN");
    GPUTILS_ARG(strideB >= 0, 7, "negative strideB");
    GPUTILS_ARG(batch >= 0, 8, "negative batch");
    if (batch == 0) return 0;
    GPUTILS_ARG(N != nullptr, 2, "N is NULL");
    GPUTILS_ARG(rank != nullptr, 5, "rank is NULL");
    GPUTILS_ARG(b != nullptr, 6, "b is NULL");
    int S = 1;
    while (S * 2 * n <= kThreads) S *= 2;
    const size_t generic_lds = sizeof(T) * (size_t)n * (2 + S);
    GPUTILS_ARG(generic_lds <= (size_t)kMaxLds, 1, "n too large");
    constexpr int V = 16 / sizeof(T);
    const bool vec = n % V == 0 && n / V <= kThreads && ldn % V == 0 && strideN % V == 0 && strideB % V == 0 &&
                     reinterpret_cast<uintptr_t>(N) % 16 == 0 && reinterpret_cast<uintptr_t>(b) % 16 == 0 &&
                     (int64_t)(n + kThreads) * ldn * (int64_t)sizeof(T) < ((int64_t)1 << 31);  // 32-bit offsets
    size_t lds = generic_lds;
    if (vec) {
        const size_t part = std::max((size_t)(n / V) * n * sizeof(T), (size_t)kThreads * 16);
        lds = std::max(lds, part + sizeof(T) * (size_t)n);
        lds = std::max(lds, (size_t)16 * 16 * (n / V));  // lane-group path: 16 rows of n / V vectors
    }
    GPUTILS_ARG(lds <= (size_t)kMaxLds, 1, "n too large");
    int kc = 0, rv = 0;
    if (vec) {
        const int G = kThreads / (n / V), need = (n + G - 1) / G;
        kc = need <= 4 ? 4 : need <= 8 ? 8 : need <= 12 ? 12 : 16;
        if (n % (8 * V) == 0 && n / (8 * V) <= 4) rv = n / (8 * V);
    }
    auto kern = rv == 1    ? nullspace_project_kernel<T, 0, 1>
                : rv == 2  ? nullspace_project_kernel<T, 0, 2>
                : rv == 3  ? nullspace_project_kernel<T, 0, 3>
                : rv == 4  ? nullspace_project_kernel<T, 0, 4>
                : kc == 4  ? nullspace_project_kernel<T, 4, 0>
                : kc == 8  ? nullspace_project_kernel<T, 8, 0>
                : kc == 12 ? nullspace_project_kernel<T, 12, 0>
                : kc == 16 ? nullspace_project_kernel<T, 16, 0>
                           : nullspace_project_kernel<T, 0, 0>;
    if (lds > 64 * 1024)
        GPUTILS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const unsigned gx = (unsigned)(batch < 65535 ? batch : 65535);
    const unsigned gy = (unsigned)((batch + 65534) / 65535);
    hipLaunchKernelGGL(kern, dim3(gx, gy), dim3(kThreads), lds, static_cast<hipStream_t>(stream), n, N, ldn, strideN,
                       rank, b, strideB, batch, S, (int)vec);
    return launch_check("nullspace_project_kernel");
}

template <typename T>
int nullspace_build(int n, const T *U, int64_t ldu, int64_t strideU, const unsigned *rank, T *N, int64_t ldn,
                    int64_t strideN, T *P, int64_t ldp, int64_t strideP, int64_t batch, void *stream) {
    GPUTILS_ARG(n >= 1, 1, "n must be >= 1");
    GPUTILS_ARG(ldu >= n, 3, "ldu < n");
    GPUTILS_ARG(strideU >= 0, 4, "negative strideU");
    GPUTILS_ARG(ldn >= n, 7, "ldn < n");
    GPUTILS_ARG(strideN >= 0, 8, "negative strideN");
    GPUTILS_ARG(ldp >= n, 10, "ldp < n");
    GPUTILS_ARG(strideP >= 0, 11, "negative strideP");
    GPUTILS_ARG(batch >= 0 && batch <= 0x7fffffff, 12, "batch out of range");
    if (batch == 0) return 0;
    GPUTILS_ARG(U != nullptr, 2, "U is NULL");
    GPUTILS_ARG(rank != nullptr, 5, "rank is NULL");
    GPUTILS_ARG(N != nullptr, 6, "N is NULL");
    GPUTILS_ARG(P != nullptr, 9, "P is NULL");
    // the matrix-core form when its staged block fits the LDS
    const int np = 16 * ((n + 15) >> 4), kp = (n + 3) & ~3;
    const size_t lds_mfma = (size_t)np * (size_t)(n > kp ? n : kp) * sizeof(T);
    if (lds_mfma <= (size_t)kMaxLds) {
        auto kern = nullspace_build_mfma_kernel<T>;
        if (lds_mfma > 64 * 1024)
            GPUTILS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_mfma));
        const int vec4 = (sizeof(T) == 4 && ldp % 4 == 0 && strideP % 4 == 0 &&
                          reinterpret_cast<uintptr_t>(P) % 16 == 0)
                             ? 1
                             : 0;
        hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kThreads), lds_mfma, static_cast<hipStream_t>(stream), n,
                           U, ldu, strideU, rank, N, ldn, strideN, P, ldp, strideP, batch, vec4);
        return launch_check("nullspace_build_mfma_kernel");
    }
    const size_t lds = (size_t)kNbKc * n * sizeof(T);
    GPUTILS_ARG(lds <= (size_t)kMaxLds, 1, "n too large");
    auto kern = nullspace_build_kernel<T>;
    if (lds > 64 * 1024)
        GPUTILS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kThreads), lds, static_cast<hipStream_t>(stream), n, U, ldu,
                       strideU, rank, N, ldn, strideN, P, ldp, strideP, batch);
    return launch_check("nullspace_build_kernel");
}

}  // namespace svd
}  // namespace gputils

using namespace gputils;

extern "C" {

int gputils_gesvd_set_fp32_sweeps(int enable) {
    return svd::g_fp32_sweeps.exchange(enable != 0 ? 1 : 0);
}

int64_t gputils_gesvd_workspace_bytes(int m, int n, int64_t batch) {
    if (m < n || n < 1 || batch < 0) return -1;
    const int mode = svd::gesvd_mode(m, n);
    if (svd::layout(m, n, mode).bytes > svd::kMaxLds) return -1;
    return svd::layout(m, n, mode).wstride * (int64_t)sizeof(double) * batch;
}

int gputils_sgesvd_batched(int m, int n, const float *A, int64_t lda, int64_t strideA, float *S, int64_t strideS,
                           float *U, int64_t ldu, int64_t strideU, float *VT, int64_t ldvt, int64_t strideVT,
                           void *work, int *info, int *sweeps, int64_t batch, void *stream) {
    return svd::gesvd_batched<float>(m, n, A, lda, strideA, S, strideS, U, ldu, strideU, VT, ldvt, strideVT, work,
                                     info, sweeps, batch, stream);
}
int gputils_dgesvd_batched(int m, int n, const double *A, int64_t lda, int64_t strideA, double *S, int64_t strideS,
                           double *U, int64_t ldu, int64_t strideU, double *VT, int64_t ldvt, int64_t strideVT,
                           void *work, int *info, int *sweeps, int64_t batch, void *stream) {
    return svd::gesvd_batched<double>(m, n, A, lda, strideA, S, strideS, U, ldu, strideU, VT, ldvt, strideVT, work,
                                      info, sweeps, batch, stream);
}

int gputils_ssvd_rank_batched(int n, const float *S, int64_t strideS, float eps, unsigned *rank, int64_t batch,
                              void *stream) {
    return svd::rank_batched<float>(n, S, strideS, eps, rank, batch, stream);
}
int gputils_dsvd_rank_batched(int n, const double *S, int64_t strideS, double eps, unsigned *rank, int64_t batch,
                              void *stream) {
    return svd::rank_batched<double>(n, S, strideS, eps, rank, batch, stream);
}

int gputils_snullspace_build_batched(int n, const float *U, int64_t ldu, int64_t strideU, const unsigned *rank,
                                     float *N, int64_t ldn, int64_t strideN, float *P, int64_t ldp, int64_t strideP,
                                     int64_t batch, void *stream) {
    return svd::nullspace_build<float>(n, U, ldu, strideU, rank, N, ldn, strideN, P, ldp, strideP, batch, stream);
}
int gputils_dnullspace_build_batched(int n, const double *U, int64_t ldu, int64_t strideU, const unsigned *rank,
                                     double *N, int64_t ldn, int64_t strideN, double *P, int64_t ldp,
                                     int64_t strideP, int64_t batch, void *stream) {
    return svd::nullspace_build<double>(n, U, ldu, strideU, rank, N, ldn, strideN, P, ldp, strideP, batch, stream);
}

int gputils_snullspace_project_batched(int n, const float *N, int64_t ldn, int64_t strideN, const unsigned *rank,
                                       float *b, int64_t strideB, int64_t batch, void *stream) {
    return svd::nullspace_project<float>(n, N, ldn, strideN, rank, b, strideB, batch, stream);
}
int gputils_dnullspace_project_batched(int n, const double *N, int64_t ldn, int64_t strideN, const unsigned *rank,
                                       double *b, int64_t strideB, int64_t batch, void *stream) {
    return svd::nullspace_project<double>(n, N, ldn, strideN, rank, b, strideB, batch, stream);
}

}  // extern "C"
