// blas1.hip -- whole-tensor reductions, scal/axpy, batched transpose and
// Givens rotations.
//
// Replace the cuBLAS level-1 calls behind DTensor::dotF/normF/sumAbs/
// maxAbs/minAbs, *=, +=, -=, tr() and applyLeft/RightGivensRotation, and
// GivensAnnihilator (include/tensor.cuh:968-1104, 1167-1276, 2204-2312).  Reductions are two launches: a grid-stride pass writing one
// partial per workgroup into the caller's workspace, then one workgroup
// combining the partials.  All accumulation happens in double.
#include "common.hpp"

namespace gputils {
namespace blas1 {

constexpr int kThreads = 256;
constexpr int kMaxPartials = 1024;

enum Op { kDot = 0, kSsq = 1, kAsum = 2, kAmax = 3, kAmin = 4 };

template <int OP>
__device__ inline double combine(double a, double b) {
    if (OP == kAmax) return a > b ? a : b;
    if (OP == kAmin) return a < b ? a : b;
    return a + b;
}

template <int OP>
__device__ inline double identity() {
    if (OP == kAmin) return __builtin_inf();
    return 0.0;
}

template <int OP>
__device__ double block_reduce(double v, double *s) {
    v = (OP == kAmax) ? wave_max(v) : (OP == kAmin) ? wave_min(v) : wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s[0];
        for (int w = 1; w < kThreads / kWave; ++w) r = combine<OP>(r, s[w]);
        s[0] = r;
    }
    __syncthreads();
    return s[0];
}

template <int OP, typename T>
__device__ inline double elem(const T *x, const T *y, int64_t i) {
    const double a = (double)x[i];
    if (OP == kDot) return a * (double)y[i];
    if (OP == kSsq) return a * a;
    return a < 0 ? -a : a;
}

// pass 1: partial[blockIdx.x] (+ for kSsq the max |x| in partial[kMaxPartials + blockIdx.x])
template <int OP, typename T>
__global__ __launch_bounds__(kThreads) void reduce_pass1(int64_t n, const T *x, const T *y,
                                                         double *partial) {
    __shared__ double s[kThreads / kWave];
    double acc = identity<OP>(), big = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        acc = combine<OP>(acc, elem<OP>(x, y, i));
        if (OP == kSsq) {
            const double a = fabs((double)x[i]);
            big = a > big ? a : big;
        }
    }
    acc = block_reduce<OP>(acc, s);
    if (OP == kSsq) {
        __syncthreads();
        big = block_reduce<kAmax>(big, s);
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = acc;
        if (OP == kSsq) partial[kMaxPartials + blockIdx.x] = big;
    }
}

// pass 2: combine partials, finish (sqrt for nrm2), cast to T.
// For kSsq with a maximum outside [1e-140, 1e140] the sum of squares may
// have lost range; the single workgroup then rescans x scaled by 1/max.
template <int OP, typename T>
__global__ __launch_bounds__(kThreads) void reduce_pass2(int nparts, const double *partial,
                                                         int64_t n, const T *x, T *result) {
    __shared__ double s[kThreads / kWave];
    double acc = identity<OP>(), big = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
        acc = combine<OP>(acc, partial[i]);
        if (OP == kSsq) big = fmax(big, partial[kMaxPartials + i]);
    }
    acc = block_reduce<OP>(acc, s);
    if (OP == kSsq) {
        __syncthreads();
        big = block_reduce<kAmax>(big, s);
        if (big > 0.0 && (big > 1e140 || big < 1e-140)) {
            const double inv = 1.0 / big;
            double q = 0.0;
            for (int64_t i = threadIdx.x; i < n; i += kThreads) {
                const double v = (double)x[i] * inv;
                q += v * v;
            }
            __syncthreads();
            q = block_reduce<kDot>(q, s);
            acc = q;
            if (threadIdx.x == 0) result[0] = (T)(big * sqrt(acc));
            return;
        }
        if (threadIdx.x == 0) result[0] = (T)sqrt(acc);
        return;
    }
    if (threadIdx.x == 0) result[0] = (T)acc;
}

template <int OP, typename T>
int reduce(int64_t n, const T *x, const T *y, T *result, void *ws, hipStream_t st) {
    if (n <= 0) {
        // empty reductions: 0 (cuBLAS returns 0 for n <= 0)
        return gputils_fill_bytes_async(result, 0, sizeof(T), st);
    }
    const int parts = (int)grid_1d(n, kThreads * 8, kMaxPartials);
    double *partial = static_cast<double *>(ws);
    hipLaunchKernelGGL((reduce_pass1<OP, T>), dim3(parts), dim3(kThreads), 0, st, n, x, y, partial);
    int rc = launch_check("reduce_pass1");
    if (rc) return rc;
    hipLaunchKernelGGL((reduce_pass2<OP, T>), dim3(1), dim3(kThreads), 0, st, parts, partial, n, x, result);
    return launch_check("reduce_pass2");
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scal_kernel(int64_t n, T alpha, T *x) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        x[i] *= alpha;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void axpy_kernel(int64_t n, T alpha, const T *x, T *y) {
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        y[i] += alpha * x[i];
}

// Plane rotation of two strided vectors (BLAS rot, behind
// DTensor::applyRight/LeftGivensRotation, tensor.cuh:1074-1104):
// x_k <- c x_k + s y_k, y_k <- c y_k - s x_k.
template <typename T>
__global__ __launch_bounds__(kThreads) void rot_kernel(int64_t n, T *x, int64_t incx, T *y, int64_t incy, T c,
                                                       T s) {
    for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const T xv = x[k * incx], yv = y[k * incy];
        x[k * incx] = c * xv + s * yv;
        y[k * incy] = c * yv - s * xv;
    }
}

// 1 / hypot(a, b) with the squares formed after scaling by the larger
// magnitude (no overflow or underflow): m = max(|a|, |b|), t = min / m,
// 1 / (m sqrt(1 + t^2)); +inf for a = b = 0, like 1 / hypot(0, 0).
template <typename T>
__device__ inline T rhypot_scaled(T a, T b) {
    const T aa = fabs(a), bb = fabs(b);
    const T m = aa > bb ? aa : bb, lo = aa > bb ? bb : aa;
    if (!(m > T(0))) return T(1) / m;
    const T t = lo / m;
    return T(1) / (m * sqrt(T(1) + t * t));
}

// GivensAnnihilator::annihilate (tensor.cuh:2250-2311) in one launch: every
// thread reads A(i,j) and A(k,j), the workgroup synchronises, and rows i
// and k are rotated by c = A(i,j) r, s = A(k,j) r (r = 1 / hypot), which
// zeroes A(k,j).  The sines and cosines never leave the device.  One
// workgroup: its barrier orders the two reads before the column-j update.
template <typename T>
__global__ __launch_bounds__(kThreads) void givens_annihilate_kernel(int64_t ncols, T *A, int64_t lda, int64_t i,
                                                                     int64_t k, int64_t j) {
    const T xij = A[i + j * lda], xkj = A[k + j * lda];
    const T r = rhypot_scaled(xij, xkj);
    const T c = xij * r, s = xkj * r;
    __syncthreads();
    for (int64_t col = threadIdx.x; col < ncols; col += kThreads) {
        T *p = A + col * lda;
        const T xv = p[i], yv = p[k];
        p[i] = c * xv + s * yv;
        p[k] = c * yv - s * xv;
    }
}

// B_b (n x m) = A_b^T (A_b m x n); 32x32 tiles through LDS, grid (tiles, batch).
template <typename T>
__global__ __launch_bounds__(kThreads) void transpose_kernel(int m, int n, const T *A, int64_t lda,
                                                             int64_t strideA, T *B, int64_t ldb,
                                                             int64_t strideB, int tiles_m) {
    __shared__ T tile[32][33];
    const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
    const T *a = A + blockIdx.y * strideA;
    T *bp = B + blockIdx.y * strideB;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int r = tm * 32 + tx, c = tn * 32 + j;
        if (r < m && c < n) tile[j][tx] = a[r + (int64_t)c * lda];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = tn * 32 + tx, c = tm * 32 + j;  // B row r (= A col), B col c (= A row)
        if (r < n && c < m) bp[r + (int64_t)c * ldb] = tile[tx][j];
    }
}

// Small matrices with no padding (lda = m, ldb = n): one workgroup per
// matrix, the whole matrix through LDS.  A is read and B written as the
// contiguous blocks they are (consecutive lanes on consecutive elements);
// the LDS copy holds B with the odd column stride n + 1, so both the
// transposing writes (stride n + 1 across lanes) and the reads are
// conflict-free.  Eight loads per thread are issued together.  Element e
// of a column-major p x q block sits in column floor((e + 1/2) / p),
// computed in FP32: exact for e < 2^16 and p <= 512, since (e + 1/2) / p
// lies at least 1 / (2p) from an integer and the float quotient within
// 2^-17 of it.
constexpr int kSmallTransposeElems = 8192;
constexpr int kSmallTransposeMaxDim = 512;

template <typename T>
__global__ __launch_bounds__(kThreads) void transpose_small_kernel(int m, int n, const T *__restrict__ A,
                                                                   int64_t strideA, T *__restrict__ B,
                                                                   int64_t strideB) {
    extern __shared__ unsigned char dyn_raw[];
    T *S = reinterpret_cast<T *>(dyn_raw);  // B (n x m) with column stride n + 1
    const T *a = A + (int64_t)blockIdx.x * strideA;
    T *bp = B + (int64_t)blockIdx.x * strideB;
    const int total = m * n, lds = n + 1;
    const float inv_m = 1.0f / (float)m, inv_n = 1.0f / (float)n;
    constexpr int kU = 8;
    for (int e0 = threadIdx.x; e0 < total; e0 += kU * kThreads) {
        T v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = e0 + u * kThreads;
            v[u] = a[e < total ? e : 0];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = e0 + u * kThreads;
            const int j = (int)(((float)e + 0.5f) * inv_m), i = e - j * m;  // A(i, j)
            if (e < total) S[j + i * lds] = v[u];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < total; e += kThreads) {
        const int c = (int)(((float)e + 0.5f) * inv_n), r = e - c * n;  // B(r, c)
        bp[e] = S[r + c * lds];
    }
}

template <typename T>
int transpose(int64_t m, int64_t n, const T *A, int64_t lda, int64_t strideA, T *B, int64_t ldb,
              int64_t strideB, int64_t batch, hipStream_t st) {
    const size_t lds = (size_t)(n + 1) * (size_t)m * sizeof(T);
    if (lda == m && ldb == n && m * n <= kSmallTransposeElems && m <= kSmallTransposeMaxDim &&
        n <= kSmallTransposeMaxDim && lds <= 64 * 1024) {
        for (int64_t b0 = 0; b0 < batch; b0 += 0x7fffffff) {
            const int64_t nb = (batch - b0 < 0x7fffffff) ? (batch - b0) : 0x7fffffff;
            hipLaunchKernelGGL((transpose_small_kernel<T>), dim3((unsigned)nb), dim3(kThreads), lds, st, (int)m,
                               (int)n, A + b0 * strideA, strideA, B + b0 * strideB, strideB);
            int rc = launch_check("transpose_small_kernel");
            if (rc) return rc;
        }
        return GPUTILS_OK;
    }
    const int tiles_m = (int)ceil_div(m, 32), tiles_n = (int)ceil_div(n, 32);
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
        const int64_t nb = (batch - b0 < 65535) ? (batch - b0) : 65535;
        hipLaunchKernelGGL((transpose_kernel<T>), dim3(tiles_m * tiles_n, (unsigned)nb), dim3(kThreads), 0, st,
                           (int)m, (int)n, A + b0 * strideA, lda, strideA, B + b0 * strideB, ldb, strideB,
                           tiles_m);
        int rc = launch_check("transpose_kernel");
        if (rc) return rc;
    }
    return GPUTILS_OK;
}

}  // namespace blas1
}  // namespace gputils

using namespace gputils;
using namespace gputils::blas1;

#define GPUTILS_REDUCE_ARGS(res_idx, ws_idx)                          \
    GPUTILS_ARG(n >= 0, 1, "n < 0");                                  \
    GPUTILS_ARG(result != nullptr, res_idx, "result is NULL");        \
    GPUTILS_ARG(n == 0 || workspace != nullptr, ws_idx, "workspace is NULL");

extern "C" {

size_t gputils_reduce_workspace_bytes(void) { return sizeof(double) * 2 * kMaxPartials; }

int gputils_sdot(int64_t n, const float *x, const float *y, float *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(4, 5);
    return reduce<kDot, float>(n, x, y, result, workspace, as_stream(stream));
}
int gputils_ddot(int64_t n, const double *x, const double *y, double *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(4, 5);
    return reduce<kDot, double>(n, x, y, result, workspace, as_stream(stream));
}
int gputils_snrm2(int64_t n, const float *x, float *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kSsq, float>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_dnrm2(int64_t n, const double *x, double *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kSsq, double>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_sasum(int64_t n, const float *x, float *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kAsum, float>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_dasum(int64_t n, const double *x, double *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kAsum, double>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_samax(int64_t n, const float *x, float *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kAmax, float>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_damax(int64_t n, const double *x, double *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kAmax, double>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_samin(int64_t n, const float *x, float *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kAmin, float>(n, x, x, result, workspace, as_stream(stream));
}
int gputils_damin(int64_t n, const double *x, double *result, void *workspace, void *stream) {
    GPUTILS_REDUCE_ARGS(3, 4);
    return reduce<kAmin, double>(n, x, x, result, workspace, as_stream(stream));
}

int gputils_sscal(int64_t n, float alpha, float *x, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    if (n == 0) return GPUTILS_OK;
    GPUTILS_ARG(x != nullptr, 3, "x is NULL");
    hipLaunchKernelGGL((scal_kernel<float>), dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0,
                       as_stream(stream), n, alpha, x);
    return launch_check("scal_kernel");
}
int gputils_dscal(int64_t n, double alpha, double *x, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    if (n == 0) return GPUTILS_OK;
    GPUTILS_ARG(x != nullptr, 3, "x is NULL");
    hipLaunchKernelGGL((scal_kernel<double>), dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0,
                       as_stream(stream), n, alpha, x);
    return launch_check("scal_kernel");
}
int gputils_saxpy(int64_t n, float alpha, const float *x, float *y, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    if (n == 0) return GPUTILS_OK;
    GPUTILS_ARG(x != nullptr, 3, "x is NULL");
    GPUTILS_ARG(y != nullptr, 4, "y is NULL");
    hipLaunchKernelGGL((axpy_kernel<float>), dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0,
                       as_stream(stream), n, alpha, x, y);
    return launch_check("axpy_kernel");
}
int gputils_daxpy(int64_t n, double alpha, const double *x, double *y, void *stream) {
    GPUTILS_ARG(n >= 0, 1, "n < 0");
    if (n == 0) return GPUTILS_OK;
    GPUTILS_ARG(x != nullptr, 3, "x is NULL");
    GPUTILS_ARG(y != nullptr, 4, "y is NULL");
    hipLaunchKernelGGL((axpy_kernel<double>), dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0,
                       as_stream(stream), n, alpha, x, y);
    return launch_check("axpy_kernel");
}

int gputils_stranspose_batched(int64_t m, int64_t n, const float *A, int64_t lda, int64_t strideA,
                               float *B, int64_t ldb, int64_t strideB, int64_t batch, void *stream) {
    GPUTILS_ARG(m >= 0, 1, "m < 0");
    GPUTILS_ARG(n >= 0, 2, "n < 0");
    GPUTILS_ARG(lda >= (m > 1 ? m : 1), 4, "lda < max(1,m)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(batch >= 0, 9, "batch < 0");
    if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(B != nullptr, 6, "B is NULL");
    return transpose<float>(m, n, A, lda, strideA, B, ldb, strideB, batch, as_stream(stream));
}
int gputils_dtranspose_batched(int64_t m, int64_t n, const double *A, int64_t lda, int64_t strideA,
                               double *B, int64_t ldb, int64_t strideB, int64_t batch, void *stream) {
    GPUTILS_ARG(m >= 0, 1, "m < 0");
    GPUTILS_ARG(n >= 0, 2, "n < 0");
    GPUTILS_ARG(lda >= (m > 1 ? m : 1), 4, "lda < max(1,m)");
    GPUTILS_ARG(ldb >= (n > 1 ? n : 1), 7, "ldb < max(1,n)");
    GPUTILS_ARG(batch >= 0, 9, "batch < 0");
    if (m == 0 || n == 0 || batch == 0) return GPUTILS_OK;
    GPUTILS_ARG(A != nullptr, 3, "A is NULL");
    GPUTILS_ARG(B != nullptr, 6, "B is NULL");
    return transpose<double>(m, n, A, lda, strideA, B, ldb, strideB, batch, as_stream(stream));
}

#define GPUTILS_ROT_ENTRY(NAME, T)                                                                       \
    int NAME(int64_t n, T *x, int64_t incx, T *y, int64_t incy, T c, T s, void *stream) {                 \
        GPUTILS_ARG(n >= 0, 1, "n < 0");                                                                  \
        GPUTILS_ARG(incx >= 1, 3, "incx < 1");                                                            \
        GPUTILS_ARG(incy >= 1, 5, "incy < 1");                                                            \
        if (n == 0) return GPUTILS_OK;                                                                    \
        GPUTILS_ARG(x != nullptr, 2, "x is NULL");                                                        \
        GPUTILS_ARG(y != nullptr, 4, "y is NULL");                                                        \
        hipLaunchKernelGGL((rot_kernel<T>), dim3(grid_1d(n, kThreads * 4)), dim3(kThreads), 0,            \
                           as_stream(stream), n, x, incx, y, incy, c, s);                                 \
        return launch_check("rot_kernel");                                                                \
    }
GPUTILS_ROT_ENTRY(gputils_srot, float)
GPUTILS_ROT_ENTRY(gputils_drot, double)

#define GPUTILS_GIVENS_ENTRY(NAME, T)                                                                     \
    int NAME(int64_t m, int64_t ncols, T *A, int64_t lda, int64_t i, int64_t k, int64_t j, void *stream) { \
        GPUTILS_ARG(m >= 0, 1, "m < 0");                                                                  \
        GPUTILS_ARG(ncols >= 0, 2, "ncols < 0");                                                          \
        GPUTILS_ARG(lda >= (m > 1 ? m : 1), 4, "lda < max(1,m)");                                         \
        GPUTILS_ARG(i >= 0 && i < m, 5, "row i out of range");                                            \
        GPUTILS_ARG(k >= 0 && k < m, 6, "row k out of range");                                            \
        GPUTILS_ARG(j >= 0 && j < ncols, 7, "column j out of range");                                     \
        GPUTILS_ARG(A != nullptr, 3, "A is NULL");                                                        \
        hipLaunchKernelGGL((givens_annihilate_kernel<T>), dim3(1), dim3(kThreads), 0, as_stream(stream),  \
                           ncols, A, lda, i, k, j);                                                       \
        return launch_check("givens_annihilate_kernel");                                                  \
    }
GPUTILS_GIVENS_ENTRY(gputils_sgivens_annihilate, float)
GPUTILS_GIVENS_ENTRY(gputils_dgivens_annihilate, double)

}  // extern "C"
