// gemm.hip -- batched C_i <- alpha A_i B_i + beta C_i for gfx950.
//
// Replaces cublas?gemmBatched / ?gemm behind DTensor::addAB and operator*
// (include/tensor.cuh:1286-1338, 645-650).
//
// Tiled path: one 256-thread workgroup computes a 64x64 tile of one C_i.
// K is staged through LDS in chunks (32 for float, 16 for double) with the
// k index contiguous for both operands (A tile stored transposed).  Each
// wavefront owns a 32x32 quadrant = 2x2 MFMA tiles of 16x16:
//   float, FP32 mode   v_mfma_f32_16x16x4_f32      (exact fp32 products)
//   float, BF16X3 mode v_mfma_f32_16x16x32_bf16 x3: a = ah + al, b = bh + bl
//                      (bf16 terms), C += ah bh + ah bl + al bh, fp32 acc.
//   double             v_mfma_f64_16x16x4_f64
// The C tile is staged through LDS and written column by column (coalesced).
//
// Per-matrix path (C aliasing B, or very narrow C): one workgroup per
// matrix stages all of B_i in LDS before any C_i element is written, which
// makes the in-place product of Nullspace::project (tensor.cuh:2084) safe.
#include "common.hpp"

#include <cstdint>
#include <type_traits>

namespace gputils {
namespace gemm {

constexpr int kThreads = 256;
constexpr int kTM = 64, kTN = 64;

template <typename T>
struct Cfg;
template <>
struct Cfg<float> {
    static constexpr int TK = 32, S = 36;  // k-chunk, LDS row stride (16-byte aligned rows)
};
template <>
struct Cfg<double> {
    static constexpr int TK = 16, S = 17;
};

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline void split_bf16(const float *v, bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)v[e];
        hi[e] = h;
        lo[e] = (__bf16)(v[e] - (float)h);
    }
}

// Stage A[i0:i0+64, k0:k0+TK] as As[i][k] and B[k0:k0+TK, j0:j0+64] as
// Bs[j][k] (both with row stride S), zero outside the matrix.
template <typename T>
__device__ void stage(const T *a, int64_t lda, const T *b, int64_t ldb, int m, int n, int k, int i0,
                      int j0, int k0, T *As, T *Bs) {
    constexpr int TK = Cfg<T>::TK, S = Cfg<T>::S;
    // A: consecutive threads walk down a column (contiguous in memory)
    for (int e = threadIdx.x; e < kTM * TK; e += kThreads) {
        const int i = e % kTM, kk = e / kTM;
        const int gi = i0 + i, gk = k0 + kk;
        As[i * S + kk] = (gi < m && gk < k) ? a[gi + (int64_t)gk * lda] : T(0);
    }
    // B: consecutive threads walk down a column of B (k contiguous)
    for (int e = threadIdx.x; e < kTN * TK; e += kThreads) {
        const int kk = e % TK, j = e / TK;
        const int gk = k0 + kk, gj = j0 + j;
        Bs[j * S + kk] = (gk < k && gj < n) ? b[gk + (int64_t)gj * ldb] : T(0);
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(kThreads) void gemm_tiled_kernel(int m, int n, int k, T alpha,
                                                              const T *__restrict__ A, int64_t lda,
                                                              int64_t strideA,
                                                              const T *__restrict__ B, int64_t ldb,
                                                              int64_t strideB, T beta, T *C,
                                                              int64_t ldc, int64_t strideC,
                                                              int tiles_m, int64_t batch) {
    constexpr int TK = Cfg<T>::TK, S = Cfg<T>::S;
    __shared__ __attribute__((aligned(16))) T smem[2 * kTM * S > kTM * (kTN + 1) ? 2 * kTM * S : kTM * (kTN + 1)];
    T *As = smem, *Bs = smem + kTM * S;
    using M = Mfma16<T>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = 32 * (wave & 1), wn = 32 * (wave >> 1);
    const int i0 = (blockIdx.x % tiles_m) * kTM, j0 = (blockIdx.x / tiles_m) * kTN;
    const int64_t bidx = blockIdx.y + (int64_t)blockIdx.z * 65535;
    if (bidx >= batch) return;  // uniform per workgroup
    const T *a = A + bidx * strideA;
    const T *b = B + bidx * strideB;
    T *c = C + bidx * strideC;

    typename M::acc_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = M::zero();

    for (int k0 = 0; k0 < k; k0 += TK) {
        __syncthreads();
        stage<T>(a, lda, b, ldb, m, n, k, i0, j0, k0, As, Bs);
        __syncthreads();
        if constexpr (MODE == GPUTILS_GEMM_BF16X3) {
            // one 16x16x32 step per tile: lane holds 8 consecutive k
            const int r = lane & 15, g = lane >> 4;
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const float4 *pa = reinterpret_cast<const float4 *>(As + (wm + 16 * x + r) * S + 8 * g);
                const float4 *pb = reinterpret_cast<const float4 *>(Bs + (wn + 16 * x + r) * S + 8 * g);
                float va[8], vb[8];
                const float4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
                va[0] = a0.x; va[1] = a0.y; va[2] = a0.z; va[3] = a0.w;
                va[4] = a1.x; va[5] = a1.y; va[6] = a1.z; va[7] = a1.w;
                vb[0] = b0.x; vb[1] = b0.y; vb[2] = b0.z; vb[3] = b0.w;
                vb[4] = b1.x; vb[5] = b1.y; vb[6] = b1.z; vb[7] = b1.w;
                split_bf16(va, ah[x], al[x]);
                split_bf16(vb, bh[x], bl[x]);
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                }
        } else {
            const int r = lane & 15, g = lane >> 4;
#pragma unroll
            for (int kk = 0; kk < TK; kk += 4) {
                T av[2], bv[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    av[x] = As[(wm + 16 * x + r) * S + kk + g];
                    bv[x] = Bs[(wn + 16 * x + r) * S + kk + g];
                }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) acc[x][y] = M::fma(av[x], bv[y], acc[x][y]);
            }
        }
    }
    // stage the 64x64 tile column-major (ld 65) and write coalesced
    __syncthreads();
    T *Cs = smem;
    {
        const int col = lane & 15;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Cs[(wm + 16 * x + M::drow(lane, r)) + (wn + 16 * y + col) * (kTM + 1)] = acc[x][y][r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTM * kTN; e += kThreads) {
        const int i = e % kTM, j = e / kTM;
        const int gi = i0 + i, gj = j0 + j;
        if (gi < m && gj < n) {
            T v = alpha * Cs[i + j * (kTM + 1)];
            T *dst = c + gi + (int64_t)gj * ldc;
            if (beta != T(0)) v += beta * *dst;
            *dst = v;
        }
    }
}

// Register-direct kernel for m, n <= 64 and k in {16, 32, 48, 64} (FP32):
// one workgroup per matrix, wave w computes the 32x32 quadrant (wm, wn) of
// C.  The reduction index is permuted so that lane (g, c) owns
// k = 16g + t, t = 0..15: its B elements are 16 consecutive floats of one
// column (four 16-byte loads) and its A elements are row wm + 16x + c of 16
// columns (loads coalesced across the 16 lanes of a group).  FP32 mode
// issues one 16x16x4 MFMA per t (the four lane groups supply
// k = 16g + t); BF16X3 mode one 16x16x32 step per 8 consecutive t.  No LDS
// and no staging: every load is in flight at once, and C leaves the
// accumulators as 16-byte stores of four consecutive rows when vec_c.
template <int MODE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MODE == GPUTILS_GEMM_BF16X3 ? 6 : 7))) void gemm_reg_kernel(int m, int n, int k, float alpha,
                                                            const float *__restrict__ A, int64_t lda,
                                                            int64_t strideA, const float *__restrict__ B,
                                                            int64_t ldb, int64_t strideB, float beta, float *C,
                                                            int64_t ldc, int64_t strideC, int vec_c,
                                                            int64_t batch) {
    const int64_t bidx = blockIdx.x + (int64_t)blockIdx.y * 65535;
    if (bidx >= batch) return;  // uniform per workgroup
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    const int wm = 32 * (wave & 1), wn = 32 * (wave >> 1);
    const float *a = A + bidx * strideA;
    const float *b = B + bidx * strideB;
    float *cm = C + bidx * strideC;
    const bool kin = 16 * g < k;  // k is a multiple of 16: a lane group is all in or all out
    // loads: addresses clamped into the matrix, values zeroed outside it
    float av[2][16], bv[2][16];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const int row = wm + 16 * x + c;
        const int64_t rc = row < m ? row : m - 1;
        const int64_t k0 = kin ? 16 * g : 0;
#pragma unroll
        for (int t = 0; t < 16; ++t) av[x][t] = a[rc + (k0 + t) * lda];
        const bool in = kin && row < m;
#pragma unroll
        for (int t = 0; t < 16; ++t) av[x][t] = in ? av[x][t] : 0.f;
    }
#pragma unroll
    for (int y = 0; y < 2; ++y) {
        const int col = wn + 16 * y + c;
        const int64_t cc = col < n ? col : n - 1;
        const float4 *p = reinterpret_cast<const float4 *>(b + cc * ldb + (kin ? 16 * g : 0));
        const bool in = kin && col < n;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 v = p[q];
            bv[y][4 * q + 0] = in ? v.x : 0.f;
            bv[y][4 * q + 1] = in ? v.y : 0.f;
            bv[y][4 * q + 2] = in ? v.z : 0.f;
            bv[y][4 * q + 3] = in ? v.w : 0.f;
        }
    }
    f32x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (MODE == GPUTILS_GEMM_BF16X3) {
#pragma unroll
        for (int s8 = 0; s8 < 2; ++s8) {
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                split_bf16(&av[x][8 * s8], ah[x], al[x]);
                split_bf16(&bv[x][8 * s8], bh[x], bl[x]);
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                }
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x][t], bv[y][t], acc[x][y], 0, 0, 0);
    }
    // every wave has consumed its A and B before any C element is written
    // (C may alias A)
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int row0 = wm + 16 * x + 4 * g, col = wn + 16 * y + c;
            if (col >= n) continue;
            float *dst = cm + row0 + (int64_t)col * ldc;
            if (vec_c && row0 + 3 < m) {
                float4 v = {alpha * acc[x][y][0], alpha * acc[x][y][1], alpha * acc[x][y][2], alpha * acc[x][y][3]};
                if (beta != 0.f) {
                    const float4 o = *reinterpret_cast<const float4 *>(dst);
                    v.x += beta * o.x;
                    v.y += beta * o.y;
                    v.z += beta * o.z;
                    v.w += beta * o.w;
                }
                *reinterpret_cast<float4 *>(dst) = v;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (row0 + r >= m) continue;
                    float v = alpha * acc[x][y][r];
                    if (beta != 0.f) v += beta * dst[r];
                    dst[r] = v;
                }
            }
        }
}

// Staged kernel for m, n <= 64 and k in {16, 32, 48, 64} (FP32) with
// 16-byte aligned columns (m % 4 == 0 and ld % 4 == 0): one workgroup per
// matrix reads A_i and B_i with 16-byte loads, consecutive threads on
// consecutive addresses (each column one contiguous read, all eight loads
// of a thread in flight at once), into LDS tiles with row stride 65 (bank
// conflict free for the operand reads below); after one barrier every
// wave computes its 32x32 quadrant of C as in gemm_reg_kernel (k = 16g + t
// for lane group g at step t) and stores it from the accumulators.
constexpr int kStS = 65;  // LDS row stride of the staged tiles (floats)
template <int MODE>
__global__ __launch_bounds__(kThreads) void gemm_staged_kernel(int m, int n, int k, float alpha,
                                                               const float *__restrict__ A, int64_t lda,
                                                               int64_t strideA, const float *__restrict__ B,
                                                               int64_t ldb, int64_t strideB, float beta, float *C,
                                                               int64_t ldc, int64_t strideC, int vec_c,
                                                               int64_t batch) {
    __shared__ float As[64 * kStS];  // As[j * kStS + i] = A(i, j), zero outside m x k
    __shared__ float Bs[64 * kStS];  // Bs[j * kStS + p] = B(p, j), zero outside k x n
    const int64_t bidx = blockIdx.x + (int64_t)blockIdx.y * 65535;
    if (bidx >= batch) return;  // uniform per workgroup
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const float *a = A + bidx * strideA;
    const float *b = B + bidx * strideB;
    float *cm = C + bidx * strideC;
    {
        f32x4 va[4], vb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = t + kThreads * u, j = e >> 4, q4 = 4 * (e & 15);
            const bool ina = j < k && q4 < m, inb = j < n && q4 < k;
            va[u] = *reinterpret_cast<const f32x4 *>(a + (int64_t)(ina ? j : 0) * lda + (ina ? q4 : 0));
            vb[u] = *reinterpret_cast<const f32x4 *>(b + (int64_t)(inb ? j : 0) * ldb + (inb ? q4 : 0));
            if (!ina) va[u] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (!inb) vb[u] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = t + kThreads * u, j = e >> 4, q4 = 4 * (e & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                As[j * kStS + q4 + r] = va[u][r];
                Bs[j * kStS + q4 + r] = vb[u][r];
            }
        }
    }
    __syncthreads();
    const int g = lane >> 4, c = lane & 15;
    const int wm = 32 * (wave & 1), wn = 32 * (wave >> 1);
    f32x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (MODE == GPUTILS_GEMM_BF16X3) {
#pragma unroll
        for (int s8 = 0; s8 < 2; ++s8) {
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                float av[8], bv[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int kk = 16 * g + 8 * s8 + e;
                    av[e] = As[kk * kStS + wm + 16 * x + c];
                    bv[e] = Bs[(wn + 16 * x + c) * kStS + kk];
                }
                split_bf16(av, ah[x], al[x]);
                split_bf16(bv, bh[x], bl[x]);
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                }
        }
    } else {
#pragma unroll
        for (int tt = 0; tt < 16; ++tt) {
            const int kk = 16 * g + tt;
            float av[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                av[x] = As[kk * kStS + wm + 16 * x + c];
                bv[x] = Bs[(wn + 16 * x + c) * kStS + kk];
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x], bv[y], acc[x][y], 0, 0, 0);
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int row0 = wm + 16 * x + 4 * g, col = wn + 16 * y + c;
            if (col >= n) continue;
            float *dst = cm + row0 + (int64_t)col * ldc;
            if (vec_c && row0 + 3 < m) {
                f32x4 v = {alpha * acc[x][y][0], alpha * acc[x][y][1], alpha * acc[x][y][2], alpha * acc[x][y][3]};
                if (beta != 0.f) {
                    const f32x4 o = *reinterpret_cast<const f32x4 *>(dst);
                    v[0] += beta * o[0];
                    v[1] += beta * o[1];
                    v[2] += beta * o[2];
                    v[3] += beta * o[3];
                }
                *reinterpret_cast<f32x4 *>(dst) = v;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (row0 + r >= m) continue;
                    float v = alpha * acc[x][y][r];
                    if (beta != 0.f) v += beta * dst[r];
                    dst[r] = v;
                }
            }
        }
}

// Batched GEMV y_i <- alpha A_i x_i + beta y_i (FP32, m % 4 == 0,
// m / 4 <= kThreads, 16-byte aligned columns), y may alias x (the in-place
// Nullspace::project, tensor.cuh:2084).  One workgroup per matrix: x is
// staged in LDS first; thread t owns row quad q = t % (m/4) of column
// group jg = t / (m/4) and reads the columns jg, jg + G, ... of its quad
// as 16-byte loads, all in flight together (consecutive threads read
// consecutive quads of a column, so each column is one contiguous read);
// the G partial sums of a quad are reduced through LDS and written after
// the last barrier.
__global__ __launch_bounds__(kThreads) void gemv_quad_kernel(int m, int k, float alpha, const float *__restrict__ A,
                                                             int64_t lda, int64_t strideA, const float *x,
                                                             int64_t strideX, float beta, float *y,
                                                             int64_t strideY, int64_t batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
    float *xs = reinterpret_cast<float *>(dyn);                   // k
    f32x4 *part = reinterpret_cast<f32x4 *>(xs + ((k + 3) & ~3));  // kThreads quads
    const int64_t bidx = blockIdx.x + (int64_t)blockIdx.y * 65535;
    if (bidx >= batch) return;  // uniform per workgroup
    const float *a = A + bidx * strideA;
    const float *xb = x + bidx * strideX;
    float *yb = y + bidx * strideY;
    const int t = threadIdx.x, rq = m >> 2, G = kThreads / rq;
    const int q = t % rq, jg = t / rq;
    for (int j = t; j < k; j += kThreads) xs[j] = xb[j];
    __syncthreads();
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (jg < G) {
        constexpr int U = 8;  // loads in flight per thread
        int j = jg;
        for (; j + (U - 1) * G < k; j += U * G) {
            f32x4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                v[u] = *reinterpret_cast<const f32x4 *>(a + 4 * q + (int64_t)(j + u * G) * lda);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float xj = xs[j + u * G];
                acc[0] = fmaf(v[u][0], xj, acc[0]);
                acc[1] = fmaf(v[u][1], xj, acc[1]);
                acc[2] = fmaf(v[u][2], xj, acc[2]);
                acc[3] = fmaf(v[u][3], xj, acc[3]);
            }
        }
        for (; j < k; j += G) {
            const f32x4 v = *reinterpret_cast<const f32x4 *>(a + 4 * q + (int64_t)j * lda);
            const float xj = xs[j];
            acc[0] = fmaf(v[0], xj, acc[0]);
            acc[1] = fmaf(v[1], xj, acc[1]);
            acc[2] = fmaf(v[2], xj, acc[2]);
            acc[3] = fmaf(v[3], xj, acc[3]);
        }
    }
    part[t] = acc;
    __syncthreads();
    if (t < rq) {
        f32x4 s = part[t];
        for (int g = 1; g < G; ++g) {
            const f32x4 p = part[t + g * rq];
            s[0] += p[0];
            s[1] += p[1];
            s[2] += p[2];
            s[3] += p[3];
        }
        float *dst = yb + 4 * t;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = alpha * s[r];
            if (beta != 0.f) v += beta * dst[r];
            dst[r] = v;
        }
    }
}

// Per-matrix kernel: stage B_i (k x n) in LDS, then each thread group
// computes C elements with a split-K dot product reduced through LDS.
template <typename T>
__global__ __launch_bounds__(kThreads) void gemm_permatrix_kernel(int m, int n, int k, T alpha,
                                                                  const T *__restrict__ A,
                                                                  int64_t lda, int64_t strideA,
                                                                  const T *B, int64_t ldb,
                                                                  int64_t strideB, T beta, T *C,
                                                                  int64_t ldc, int64_t strideC,
                                                                  int split, int64_t batch) {
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
    T *Bs = reinterpret_cast<T *>(dyn);       // k * n
    T *red = Bs + (int64_t)k * n;              // kThreads partials
    const int64_t bidx = blockIdx.x + (int64_t)blockIdx.y * 65535;
    if (bidx >= batch) return;  // uniform per workgroup
    const T *a = A + bidx * strideA;
    const T *b = B + bidx * strideB;
    T *c = C + bidx * strideC;
    for (int e = threadIdx.x; e < k * n; e += kThreads) {
        const int p = e % k, j = e / k;
        Bs[e] = b[p + (int64_t)j * ldb];
    }
    __syncthreads();
    const int groups = kThreads / split;  // output elements per pass
    const int s = threadIdx.x / groups, oi = threadIdx.x % groups;
    const int total = m * n;
    for (int base = 0; base < total; base += groups) {
        const int o = base + oi;
        T acc = T(0);
        if (o < total) {
            const int i = o % m, j = o / m;
            for (int p = s; p < k; p += split) acc += a[i + (int64_t)p * lda] * Bs[p + j * k];
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        if (s == 0 && o < total) {
            T sum = acc;
            for (int q = 1; q < split; ++q) sum += red[q * groups + oi];
            const int i = o % m, j = o / m;
            T *dst = c + i + (int64_t)j * ldc;
            T v = alpha * sum;
            if (beta != T(0)) v += beta * *dst;
            *dst = v;
        }
        __syncthreads();
    }
}

template <typename T, int MODE>
int run(int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, int64_t strideA,
        const T *B, int64_t ldb, int64_t strideB, T beta, T *C, int64_t ldc, int64_t strideC,
        int64_t batch, hipStream_t st) {
    const bool alias = (static_cast<const void *>(B) == static_cast<const void *>(C));
    if (alias && (ldb != ldc || strideB != strideC)) {
        set_error("gemm: C may alias B only with identical ld and stride");
        return -10;
    }
    if constexpr (std::is_same<T, float>::value) {
        // matrix-vector products (the in-place project among them)
        const bool quad_ok = n == 1 && k > 0 && m % 4 == 0 && m / 4 <= kThreads && lda % 4 == 0 &&
                             strideA % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
                             (size_t)k * sizeof(T) + 16 * kThreads + 16 <= 64 * 1024;
        if (quad_ok) {
            const size_t lds = sizeof(T) * (size_t)((k + 3) & ~3) + 16 * kThreads;
            const unsigned gx = (unsigned)(batch < 65535 ? batch : 65535);
            const unsigned gy = (unsigned)ceil_div(batch, 65535);
            hipLaunchKernelGGL(gemv_quad_kernel, dim3(gx, gy), dim3(kThreads), lds, st, (int)m, (int)k, alpha, A, lda,
                               strideA, B, strideB, beta, C, strideC, batch);
            return launch_check("gemv_quad_kernel");
        }
    }
    const bool narrow = (n <= 4 && m * n <= 4096);
    if (alias || narrow) {
        const size_t lds = sizeof(T) * ((size_t)k * n + kThreads);
        if (lds > 64 * 1024) {
            set_error("gemm: in-place product needs k*n*sizeof(T) <= 62 KiB");
            return -10;
        }
        // split-K factor: power of two, at most 32, so that m*n*split <= kThreads
        const int64_t mn = m * n;
        int split = 1;
        while (split < 32 && split * 2 <= k && mn * split * 2 <= kThreads) split *= 2;
        const unsigned gx = (unsigned)(batch < 65535 ? batch : 65535);
        const unsigned gy = (unsigned)ceil_div(batch, 65535);
        hipLaunchKernelGGL((gemm_permatrix_kernel<T>), dim3(gx, gy), dim3(kThreads), lds, st, (int)m,
                           (int)n, (int)k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                           strideC, split, batch);
        return launch_check("gemm_permatrix_kernel");
    }
    if constexpr (std::is_same<T, float>::value) {
        const bool b_vec = (reinterpret_cast<uintptr_t>(B) % 16 == 0) && ldb % 4 == 0 && strideB % 4 == 0;
        if (m <= kTM && n <= kTN && k > 0 && k <= 64 && k % 16 == 0 && b_vec) {
            const int vec_c = (reinterpret_cast<uintptr_t>(C) % 16 == 0) && ldc % 4 == 0 && strideC % 4 == 0;
            const unsigned gx = (unsigned)(batch < 65535 ? batch : 65535);
            const unsigned gy = (unsigned)ceil_div(batch, 65535);
            const bool a_vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && lda % 4 == 0 && strideA % 4 == 0 &&
                               m % 4 == 0;
            if (a_vec) {
                hipLaunchKernelGGL((gemm_staged_kernel<MODE>), dim3(gx, gy), dim3(kThreads), 0, st, (int)m, (int)n,
                                   (int)k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC, vec_c,
                                   batch);
                return launch_check("gemm_staged_kernel");
            }
            hipLaunchKernelGGL((gemm_reg_kernel<MODE>), dim3(gx, gy), dim3(kThreads), 0, st, (int)m, (int)n,
                               (int)k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC, vec_c,
                               batch);
            return launch_check("gemm_reg_kernel");
        }
    }
    const int tiles_m = (int)ceil_div(m, kTM), tiles_n = (int)ceil_div(n, kTN);
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    if (tiles > 0x7fffffff) {
        set_error("gemm: too many tiles");
        return -1;
    }
    const unsigned gy = (unsigned)(batch < 65535 ? batch : 65535);
    const unsigned gz = (unsigned)ceil_div(batch, 65535);
    hipLaunchKernelGGL((gemm_tiled_kernel<T, MODE>), dim3((unsigned)tiles, gy, gz), dim3(kThreads), 0,
                       st, (int)m, (int)n, (int)k, alpha, A, lda, strideA, B, ldb, strideB, beta, C,
                       ldc, strideC, tiles_m, batch);
    return launch_check("gemm_tiled_kernel");
}

}  // namespace gemm
}  // namespace gputils

using namespace gputils;

extern "C" {

int gputils_sgemm_strided_batched(int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                                  int64_t lda, int64_t strideA, const float *B, int64_t ldb,
                                  int64_t strideB, float beta, float *C, int64_t ldc,
                                  int64_t strideC, int64_t batch, int mode, void *stream) {
    GPUTILS_ARG(m >= 0, 1, "m < 0");
    GPUTILS_ARG(n >= 0, 2, "n < 0");
    GPUTILS_ARG(k >= 0, 3, "k < 0");
    GPUTILS_ARG(lda >= (m > 1 ? m : 1), 6, "lda < max(1,m)");
    GPUTILS_ARG(ldb >= (k > 1 ? k : 1), 9, "ldb < max(1,k)");
    GPUTILS_ARG(ldc >= (m > 1 ? m : 1), 13, "ldc < max(1,m)");
    GPUTILS_ARG(batch >= 0 && batch <= 65535ll * 65535ll, 15, "batch out of range");
    GPUTILS_ARG(mode == GPUTILS_GEMM_FP32 || mode == GPUTILS_GEMM_BF16X3, 16, "unknown mode");
    GPUTILS_ARG(m <= 0x7fffffff && n <= 0x7fffffff && k <= 0x7fffffff, 1, "dimension too large");
    if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(C != nullptr, 12, "C is NULL");
    GPUTILS_ARG(k == 0 || A != nullptr, 5, "A is NULL");
    GPUTILS_ARG(k == 0 || B != nullptr, 8, "B is NULL");
    hipStream_t st = as_stream(stream);
    if (mode == GPUTILS_GEMM_BF16X3 && !(B == C))
        return gemm::run<float, GPUTILS_GEMM_BF16X3>(m, n, k, alpha, A, lda, strideA, B, ldb, strideB,
                                                     beta, C, ldc, strideC, batch, st);
    return gemm::run<float, GPUTILS_GEMM_FP32>(m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta,
                                               C, ldc, strideC, batch, st);
}

int gputils_dgemm_strided_batched(int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                                  int64_t lda, int64_t strideA, const double *B, int64_t ldb,
                                  int64_t strideB, double beta, double *C, int64_t ldc,
                                  int64_t strideC, int64_t batch, void *stream) {
    GPUTILS_ARG(m >= 0, 1, "m < 0");
    GPUTILS_ARG(n >= 0, 2, "n < 0");
    GPUTILS_ARG(k >= 0, 3, "k < 0");
    GPUTILS_ARG(lda >= (m > 1 ? m : 1), 6, "lda < max(1,m)");
    GPUTILS_ARG(ldb >= (k > 1 ? k : 1), 9, "ldb < max(1,k)");
    GPUTILS_ARG(ldc >= (m > 1 ? m : 1), 13, "ldc < max(1,m)");
    GPUTILS_ARG(batch >= 0 && batch <= 65535ll * 65535ll, 15, "batch out of range");
    GPUTILS_ARG(m <= 0x7fffffff && n <= 0x7fffffff && k <= 0x7fffffff, 1, "dimension too large");
    if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(C != nullptr, 12, "C is NULL");
    GPUTILS_ARG(k == 0 || A != nullptr, 5, "A is NULL");
    GPUTILS_ARG(k == 0 || B != nullptr, 8, "B is NULL");
    return gemm::run<double, 0>(m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                strideC, batch, as_stream(stream));
}

}  // extern "C"
