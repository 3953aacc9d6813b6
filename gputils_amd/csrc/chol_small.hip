// chol_small.hip -- FP64 batched Cholesky factor + solve for n <= 32 on
// gfx950: two matrices per wavefront (one per 32-lane half), lane i of a
// half owns row i of its matrix in registers.
//
// Used for the small-n cases of cusolverDn?potrfBatched / ?potrsBatched
// behind CholeskyBatchFactoriser (include/tensor.cuh:2135-2197) when the
// factor and the solve are fused.  Right-looking unblocked algorithm: at
// step k every lane writes its current A[i][k] to a per-half LDS column
// buffer and reads the pivot and the column back (broadcast reads), then
// updates its row with A[i][j] -= A[i][k] A[j][k] / A[k][k].  The matrix is
// padded with the identity to N = 16 or 32 rows so the unrolled code needs
// no bounds checks.  The forward solve uses the row layout; the backward
// solve (L^T x = y) reads L by columns from an LDS copy.  Both multiply by
// the reciprocal pivots kept from the factorisation instead of dividing.
#include "chol_small.hpp"
#include "common.hpp"

#include <type_traits>

namespace gputils {
namespace chols {

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int kLdT = 33;  // LDS copy of L: element (row i, column j) at j * kLdT + i

template <int N, bool WRITE_L>
__global__ __launch_bounds__(64) void posv_small_kernel(int n, int nrhs, double *A, int64_t lda, int64_t strideA,
                                                        const double *B, int64_t ldb, int64_t strideB, double *X,
                                                        int64_t ldx, int64_t strideX, int *info, int64_t batch,
                                                        const int *mask) {
    __shared__ __attribute__((aligned(16))) double LT[2][N * kLdT];
    __shared__ __attribute__((aligned(16))) double CB[2][N];  // per-half broadcast buffer
    __shared__ double RI[2][N];                               // per-half 1 / L(k, k)
    const int lane = threadIdx.x, i = lane & 31;
    const bool hi = lane >= 32;
    const int64_t b = 2 * (int64_t)blockIdx.x + (hi ? 1 : 0);
    // a half that has no matrix (odd batch) or that the mask skips runs on
    // the identity and stores nothing
    bool active = b < batch;
    if (active && mask != nullptr) active = mask[b] < 0;
    const bool row_in = active && i < n;
    double *a_g = A + (active ? b : 0) * strideA;
    double *cb = CB[hi ? 1 : 0];
    double *lt = LT[hi ? 1 : 0];
    double *ri = RI[hi ? 1 : 0];

    // every lane loads a valid element (indices clamped into the matrix) so
    // that the loads are unconditional and all in flight together
    const int ic = (i < n) ? i : n - 1;
    double a[N];
    static_for<0, N>([&](auto J) {
        const int jc = (J < n) ? (int)J : n - 1;
        a[J] = a_g[ic + (int64_t)jc * lda];
    });
    static_for<0, N>([&](auto J) {
        const bool in = row_in && J < n && J <= i;
        a[J] = in ? a[J] : (i == J ? 1.0 : 0.0);
    });

    // factorisation; a failing half continues on a unit pivot (so no NaN
    // spreads) and reports its first failure
    int fail = 0;
    static_for<0, N>([&](auto K) {
        if (i < N) cb[i] = a[K];
        // the broadcast reads below must see the other lanes' writes: lanes
        // that wrote nothing would otherwise reuse an earlier read of cb
        asm volatile("" ::: "memory");
        const double d0 = cb[K];
        const bool bad = !(d0 > 0.0);
        if (bad && fail == 0) fail = K + 1;
        const double d = bad ? 1.0 : d0;
        // 1/sqrt(d): hardware estimate and one Newton step (error ~ ulp^2
        // of the estimate, below FP64 rounding)
        double inv = __builtin_amdgcn_rsq(d);
        inv = inv * fma(-0.5 * d, inv * inv, 1.5);
        const double s = d * inv;
        ri[K] = inv;  // every lane of the half holds the same value
        const double t = a[K] * (inv * inv);  // A[i][k] / A[k][k]
        static_for<K + 1, N>([&](auto J) { a[J] = fma(-t, cb[J], a[J]); });
        a[K] = (i > K) ? a[K] * inv : ((i == K) ? s : 0.0);
    });
    const int fail_h = hi ? __builtin_amdgcn_readlane(fail, 32) : __builtin_amdgcn_readlane(fail, 0);
    if (active && i == 0 && info != nullptr) info[b] = fail_h;
    const bool ok = active && fail_h == 0;

    if (WRITE_L && ok && i < n) {
        static_for<0, N>([&](auto J) {
            if (J <= i && J < n) a_g[i + (int64_t)J * lda] = a[J];
        });
    }
    // L by columns for the backward solve
    static_for<0, N>([&](auto J) { lt[J * kLdT + i] = a[J]; });

    const double *bb = B + (active ? b : 0) * strideB;
    double *xx = X + (active ? b : 0) * strideX;
    for (int r = 0; r < nrhs; ++r) {
        // keeps the LDS reads below inside the loop: hoisted out of it, the
        // loop-invariant column and diagonal reads would occupy ~128 VGPRs
        asm volatile("" ::: "memory");
        double y = row_in ? bb[i + (int64_t)r * ldb] : 0.0;
        // forward: L y = b (row layout, y_k broadcast through the column buffer)
        static_for<0, N>([&](auto K) {
            if (i < N) cb[i] = y;  // lanes past N own no row of the buffer
            asm volatile("" ::: "memory");
            const double yk = cb[K] * ri[K];
            y = (i == K) ? yk : ((i > K) ? fma(-a[K], yk, y) : y);
        });
        // backward: L^T x = y; lane i reads column i of L from LDS
        static_for<0, N>([&](auto tt) {
            constexpr int K = N - 1 - tt;
            if (i < N) cb[i] = y;  // lanes past N own no row of the buffer
            asm volatile("" ::: "memory");
            const double xk = cb[K] * ri[K];
            y = (i == K) ? xk : ((i < K) ? fma(-lt[i * kLdT + K], xk, y) : y);  // L[K][i]
        });
        if (ok && i < n) xx[i + (int64_t)r * ldx] = y;
    }
}

int launch_posv_small(int n, int nrhs, double *A, int64_t lda, int64_t strideA, const double *B, int64_t ldb,
                      int64_t strideB, double *X, int64_t ldx, int64_t strideX, int *info, int64_t batch,
                      const int *mask, bool write_l, hipStream_t st) {
    const int64_t blocks = (batch + 1) / 2;
    const dim3 grid((unsigned)blocks), block(64);
#define GPUTILS_SMALL(NV, WL)                                                                                        \
    hipLaunchKernelGGL((posv_small_kernel<NV, WL>), grid, block, 0, st, n, nrhs, A, lda, strideA, B, ldb, strideB, X, \
                       ldx, strideX, info, batch, mask)
    if (n <= 16) {
        if (write_l) GPUTILS_SMALL(16, true);
        else GPUTILS_SMALL(16, false);
    } else {
        if (write_l) GPUTILS_SMALL(32, true);
        else GPUTILS_SMALL(32, false);
    }
#undef GPUTILS_SMALL
    return launch_check("posv_small_kernel");
}

}  // namespace chols
}  // namespace gputils
