// chol_small.hip -- FP32/FP64 batched Cholesky factor + solve for n <= 32 on
// gfx950: N = 32 (n > 16) puts two matrices in a wavefront (one per 32-lane
// group), N = 16 four (one per 16-lane group); lane i of a group owns row i
// of its matrix in registers.
//
// Used for the small-n cases of cusolverDn?potrfBatched / ?potrsBatched
// behind CholeskyBatchFactoriser (include/tensor.cuh:2135-2197) when the
// factor and the solve are fused.  Right-looking unblocked algorithm: at
// step k every lane writes its current A[i][k] to a per-group LDS column
// buffer and reads the pivot and the column back (broadcast reads), then
// updates its row with A[i][j] -= A[i][k] A[j][k] / A[k][k].  Only the lower
// triangle is used: the entries right of the diagonal, loaded with the row,
// take part in the row updates but never reach the lower part, and each is replaced by zero when
// its column is reached, so the registers end with the strictly lower L and
// the diagonal and the reciprocal pivots sit in LDS.  Rows past n are zero
// rows whose pivots are taken as 1 (identity padding to N = 16 or 32) so the
// unrolled code needs no bounds checks and no per-element lane masks.  The
// solves keep every lane's value unscaled and publish the scaled one
// (value * 1 / L(k, k)) at step k, which is the same product the scaled
// form computes; the backward solve reads L(k, i) from a copy of the
// strictly lower L packed by rows in LDS (4 KB per matrix for n = 32, so a
// CU's LDS holds sixteen workgroups).  The same kernel factors only
// (nrhs = 0, potrf) and solves only (GIVEN_L: L is read, potrs).
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

// a copy of v the compiler cannot relate to v: comparisons against it are
// recomputed where they are used instead of being kept live across phases
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// the value is materialised here: the compiler may not defer the arithmetic
// producing it past this point
template <typename T>
__device__ __forceinline__ void pin(T &x) { asm volatile("" : "+v"(x)); }

// hardware reciprocal square root estimate (refined by the caller)
__device__ __forceinline__ double rsq_estimate(double d) { return __builtin_amdgcn_rsq(d); }
__device__ __forceinline__ float rsq_estimate(float d) { return __builtin_amdgcn_rsqf(d); }

// Word 3 of a raw (stride 0) buffer resource: 32-bit data format, default
// number format, no swizzle.  A buffer store at an offset at or past the
// resource's size is dropped by the range check.
constexpr int kBufferWord3 = 0x00020000;
constexpr uint32_t kStoreDropped = 0x80000000u;

__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t rs, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, (int)off, 0, 0);
}
__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t rs, uint32_t off, double v) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const u32x2 d = {(unsigned)b, (unsigned)(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(d, rs, (int)off, 0, 0);
}

template <typename T, int N, bool WRITE_L, bool GIVEN_L>
__global__ __launch_bounds__(64) void posv_small_kernel(int n, int nrhs, T *A, int64_t lda, int64_t strideA,
                                                        const T *B, int64_t ldb, int64_t strideB, T *X,
                                                        int64_t ldx, int64_t strideX, int *info, int64_t batch,
                                                        const int *mask, uint32_t l_bytes) {
    constexpr int kTri = N * (N - 1) / 2;
    constexpr int G = 64 / N;  // matrices per wavefront, one per group of N lanes
    __shared__ __attribute__((aligned(16))) T LR[G][kTri];  // strict lower L, packed by rows
    // per-group broadcast buffer: columns 0..N-1 carry the factor's column
    // and the first right-hand side, N..2N-1 L(k, k) during the
    // factorisation and the second right-hand side during the solves
    __shared__ __attribute__((aligned(16))) T CB[G][2 * N];
    __shared__ T RI[G][N];                                   // per-group 1 / L(k, k)
    const int lane = threadIdx.x, i = lane % N, grp = lane / N;
    const int64_t b = G * (int64_t)blockIdx.x + grp;
    bool active = b < batch;
    if (active && mask != nullptr) active = mask[b] < 0;
    const bool row_in = active && i < n;
    T *a_g = A + (active ? b : 0) * strideA;
    T *cb = CB[grp];
    T *lr = LR[grp];
    T *ri = RI[grp];
    T *sd = CB[grp] + N;  // read for the last time before the solves

    // every lane loads a valid element (indices clamped into the matrix) so
    // that the loads are unconditional and all in flight together.  Entries
    // above the diagonal are never used: those lanes load the diagonal
    // element instead, so a column touches only the cache lines of its lower
    // part.
    const int ic = (i < n) ? i : n - 1;
    T a[N];
    static_for<0, N>([&](auto J) {
        const int jc = (J < n) ? (int)J : n - 1;
        const int r = ic > jc ? ic : jc;
        a[J] = a_g[r + (int64_t)jc * lda];
    });
    {
        const bool pad = !row_in;
        static_for<0, N>([&](auto J) { a[J] = pad ? T(0) : a[J]; });
    }

    // factorisation; a failing group continues on a unit pivot (so no NaN
    // spreads) and reports its first failure
    int fail = 0;
    T myri_given = T(1);
    if constexpr (GIVEN_L) {
        // the factor is an input (solve only): keep its strictly lower part
        // and 1 / L(i, i); rows past n get a unit diagonal
        const int iw = opaque(i);
        T diag = T(1);
        static_for<0, N>([&](auto J) {
            diag = (iw == J) ? a[J] : diag;
            a[J] = (iw > J) ? a[J] : T(0);
        });
        myri_given = T(1) / (row_in ? diag : T(1));
    } else {
        const int iw = opaque(i);
        static_for<0, N>([&](auto K) {
            if (i < N) cb[i] = a[K];
            // the broadcast reads below must see the other lanes' writes: lanes
            // that wrote nothing would otherwise reuse an earlier read of cb
            asm volatile("" ::: "memory");
            // the rows past n are zero, so their pivots read as failed and
            // become 1; only failures of the real rows are reported
            const T d0 = cb[K];
            const bool bad = !(d0 > T(0));
            if (K < n && bad && fail == 0) fail = K + 1;
            const T d = bad ? T(1) : d0;
            // 1/sqrt(d): hardware estimate and one Newton step in T (error ~
            // ulp^2 of the estimate, below T's rounding)
            T inv = rsq_estimate(d);
            inv = inv * fma(T(-0.5) * d, inv * inv, T(1.5));
            ri[K] = inv;  // every lane of the group holds the same values
            sd[K] = d * inv;
            const T t = a[K] * (inv * inv);
            static_for<K + 1, N>([&](auto J) { a[J] = fma(-t, cb[J], a[J]); });
            a[K] = (iw > K) ? a[K] * inv : T(0);
            // finish this step's row update here: deferred to the step that
            // first needs each element, it keeps every earlier broadcast
            // column live (over 200 VGPRs instead of 127)
            static_for<K, N>([&](auto J) { pin(a[J]); });
        });
    }
    // every lane of a group read the same pivots, so fail is the group's;
    // taking it from the group's first lane makes that explicit
    int fail_h = __builtin_amdgcn_readlane(fail, 0);
    static_for<1, G>([&](auto Gc) {
        constexpr int g = Gc;
        const int f = __builtin_amdgcn_readlane(fail, g * N);
        fail_h = (grp == g) ? f : fail_h;
    });
    if (active && i == 0 && info != nullptr) info[b] = fail_h;
    const bool ok = active && fail_h == 0;

    // strictly lower L by rows for the backward solve
    {
        const int iw = opaque(i);
        const int base = iw * (iw - 1) / 2;
        static_for<0, N - 1>([&](auto J) {
            if (i < N && iw > J) lr[base + J] = a[J];
        });
    }
    asm volatile("" ::: "memory");
    T myri;
    if constexpr (GIVEN_L) myri = myri_given;
    else myri = ri[i < N ? i : 0];
    if constexpr (WRITE_L) {
        if (l_bytes != 0) {
            // one buffer resource over the wavefront's G matrices (l_bytes
            // from the first one's base, < 2^31, checked by the host): every
            // lane issues every store of its row, and a lane with no element
            // of L there (right of the diagonal, a row past n, a matrix that
            // failed or is not written) stores past the range, which the
            // range check drops, so no store needs an exec mask of its own
            const uint64_t base = reinterpret_cast<uintptr_t>(A + (int64_t)G * blockIdx.x * strideA);
            const uint64_t ubase = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                                   (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                reinterpret_cast<void *>(ubase), 0, __builtin_amdgcn_readfirstlane((int)l_bytes), kBufferWord3);
            // a storing lane's offsets row0 + J col (J <= i < n) lie inside
            // the range; the others are replaced by the dropped offset
            const bool lane_ok = ok & (i < n);
            const uint32_t row0 = (uint32_t)(((int64_t)grp * strideA + i) * (int64_t)sizeof(T));
            const uint32_t col = (uint32_t)(lda * (int64_t)sizeof(T));
            const int iw = opaque(i);
            static_for<0, N>([&](auto J) {
                const bool st = lane_ok & (J < n) & (iw > J);
                buffer_store(rs, st ? row0 + (uint32_t)J * col : kStoreDropped, a[J]);
            });
            buffer_store(rs, lane_ok ? row0 + (uint32_t)i * col : kStoreDropped, sd[i]);
        } else if (ok && i < n) {
            const int iw = opaque(i);
            static_for<0, N>([&](auto J) {
                if (J < n && iw > J) a_g[i + (int64_t)J * lda] = a[J];
            });
            a_g[i + (int64_t)i * lda] = sd[i];
        }
    }

    const T *bb = B + (active ? b : 0) * strideB;
    T *xx = X + (active ? b : 0) * strideX;
    // right-hand sides two at a time (y0, y1), each with the same
    // arithmetic as alone; an odd last one runs the single-column code
    auto solve = [&](int r, auto PAIR) {
        constexpr bool two = decltype(PAIR)::value;
        // keeps the LDS reads below inside the loop: hoisted out of it, the
        // loop-invariant reads of L would occupy ~128 VGPRs
        asm volatile("" ::: "memory");
        T y0 = row_in ? bb[i + (int64_t)r * ldb] : T(0);
        T y1 = T(0);
        if constexpr (two) y1 = row_in ? bb[i + (int64_t)(r + 1) * ldb] : T(0);
        // forward: L y = b.  Each lane keeps its unscaled value; lane k's
        // scaled value y_k = y * (1 / L(k, k)) is published at step k.
        static_for<0, N>([&](auto K) {
            cb[i] = y0 * myri;
            if constexpr (two) cb[N + i] = y1 * myri;
            asm volatile("" ::: "memory");
            y0 = fma(-a[K], cb[K], y0);
            if constexpr (two) y1 = fma(-a[K], cb[N + K], y1);
        });
        y0 *= myri;
        y1 *= myri;
        // backward: L^T x = y; lane i reads L(k, i) from row k of the packed copy
        const int iw = opaque(i);
        static_for<0, N>([&](auto tt) {
            constexpr int K = N - 1 - tt;
            if constexpr (K > 0) {
                cb[i] = y0 * myri;
                if constexpr (two) cb[N + i] = y1 * myri;
                asm volatile("" ::: "memory");
                constexpr int rowk = K * (K - 1) / 2;
                const T l = lr[rowk + (iw < K ? iw : 0)];
                const T lk = iw < K ? l : T(0);
                y0 = fma(-lk, cb[K], y0);
                if constexpr (two) y1 = fma(-lk, cb[N + K], y1);
            }
        });
        y0 *= myri;
        y1 *= myri;
        if (ok && i < n) {
            xx[i + (int64_t)r * ldx] = y0;
            if constexpr (two) xx[i + (int64_t)(r + 1) * ldx] = y1;
        }
    };
    int r = 0;
    for (; r + 1 < nrhs; r += 2) solve(r, std::true_type{});
    if (r < nrhs) solve(r, std::false_type{});
}

template <typename T>
int launch_posv_small(int n, int nrhs, T *A, int64_t lda, int64_t strideA, const T *B, int64_t ldb,
                      int64_t strideB, T *X, int64_t ldx, int64_t strideX, int *info, int64_t batch,
                      const int *mask, bool write_l, hipStream_t st) {
    const int64_t per_wave = (n <= 16) ? 4 : 2;
    const int64_t blocks = (batch + per_wave - 1) / per_wave;
    const dim3 grid((unsigned)blocks), block(64);
    // bytes a wavefront's buffer resource spans for the stores of L (from
    // its first matrix to the last element of its last one); 0 when that
    // is 2 GiB or more, or the strides are negative: then masked stores
    const int64_t span = (per_wave - 1) * strideA + (int64_t)(n - 1) * lda + n;
    const uint32_t l_bytes =
        (write_l && strideA >= 0 && lda >= n && span * (int64_t)sizeof(T) < (int64_t)kStoreDropped)
            ? (uint32_t)(span * (int64_t)sizeof(T))
            : 0u;
#define GPUTILS_SMALL(NV, WL)                                                                                        \
    hipLaunchKernelGGL((posv_small_kernel<T, NV, WL, false>), grid, block, 0, st, n, nrhs, A, lda, strideA, B, ldb, strideB, X, \
                       ldx, strideX, info, batch, mask, l_bytes)
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

template <typename T>
int launch_potrs_small(int n, int nrhs, const T *L, int64_t lda, int64_t strideL, T *B, int64_t ldb,
                       int64_t strideB, int *info, int64_t batch, hipStream_t st) {
    const int64_t per_wave = (n <= 16) ? 4 : 2;
    const int64_t blocks = (batch + per_wave - 1) / per_wave;
    const dim3 grid((unsigned)blocks), block(64);
    // L is only read (WRITE_L = false); X aliases B
    T *l = const_cast<T *>(L);
    if (n <= 16)
        hipLaunchKernelGGL((posv_small_kernel<T, 16, false, true>), grid, block, 0, st, n, nrhs, l, lda, strideL, B, ldb,
                           strideB, B, ldb, strideB, info, batch, nullptr, 0u);
    else
        hipLaunchKernelGGL((posv_small_kernel<T, 32, false, true>), grid, block, 0, st, n, nrhs, l, lda, strideL, B, ldb,
                           strideB, B, ldb, strideB, info, batch, nullptr, 0u);
    return launch_check("posv_small_kernel (solve only)");
}


#define GPUTILS_SMALL_INST(T)                                                                                       \
    template int launch_posv_small<T>(int, int, T *, int64_t, int64_t, const T *, int64_t, int64_t, T *, int64_t,  \
                                      int64_t, int *, int64_t, const int *, bool, hipStream_t);                   \
    template int launch_potrs_small<T>(int, int, const T *, int64_t, int64_t, T *, int64_t, int64_t, int *, int64_t, \
                                       hipStream_t);
GPUTILS_SMALL_INST(float)
GPUTILS_SMALL_INST(double)
#undef GPUTILS_SMALL_INST

}  // namespace chols
}  // namespace gputils
