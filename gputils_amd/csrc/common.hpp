// common.hpp -- shared device/host helpers for the gfx950 kernels of
// libgputils_hip.so.  Everything here is CDNA-specific: 64-wide wavefronts,
// 16x16x4 MFMA tiles, LDS-resident per-matrix working sets.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "gputils_hip.h"

namespace gputils {

constexpr int kWave = 64;

// ------------------------------------------------------------ error state --
// Thread-local message of the last failing call; read by gputils_last_error().
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

#define GPUTILS_HIP(call)                                      \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ::gputils::hip_fail(e_, #call); \
    } while (0)

#define GPUTILS_ARG(cond, idx, msg)                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            ::gputils::set_error("invalid argument %d: %s", (idx), (msg)); \
            return -(idx);                                                 \
        }                                                                  \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Kernel launch epilogue: report launch-configuration errors.
inline int launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return GPUTILS_OK;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grid-stride loops must never exceed the 2^31-1 blocks limit of x.
inline unsigned grid_1d(int64_t work, int per_block, int64_t cap = 65536) {
    int64_t g = ceil_div(work, per_block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// -------------------------------------------------------------- MFMA ops --
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA in float or double.  Operand layout (measured on gfx950):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, item r holds D[drow(l, r)][l % 16] with
//              f32: drow = 4 * (l / 16) + r      (rows blocked by lane group)
//              f64: drow = (l / 16) + 4 * r      (rows interleaved)
template <typename T>
struct Mfma16;

template <>
struct Mfma16<float> {
    using acc_t = f32x4;
    __device__ static inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
    __device__ static inline acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    __device__ static inline acc_t fma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma16<double> {
    using acc_t = f64x4;
    __device__ static inline int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
    __device__ static inline acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
    __device__ static inline acc_t fma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// ------------------------------------------------------- lane exchanges --
__device__ inline float readlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline double readlane(double v, int lane) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline int readlane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Wave-wide sum / max through xor shuffles (ds_swizzle/bpermute).
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// v + the value of the lane that DPP control CTRL selects (0 where that lane
// lies outside the row or the row is masked off).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, true);
    return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Floating-point wave sums in registers: row shifts 1, 2, 4, 8 leave each
// 16-lane row's sum in its lane 15, row broadcasts 15 and 31 gather the four
// row sums in lane 63, and readlane hands that to every lane.  No LDS
// round trips, unlike the xor shuffles.  All 64 lanes must be active.
template <typename T>
__device__ __forceinline__ T wave_sum_dpp(T v) {
    v = dpp_add<0x111, 0xF>(v);  // row_shr:1
    v = dpp_add<0x112, 0xF>(v);  // row_shr:2
    v = dpp_add<0x114, 0xF>(v);  // row_shr:4
    v = dpp_add<0x118, 0xF>(v);  // row_shr:8
    v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return readlane(v, 63);
}
template <>
__device__ inline float wave_sum<float>(float v) { return wave_sum_dpp(v); }
template <>
__device__ inline double wave_sum<double>(double v) { return wave_sum_dpp(v); }
template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = (o > v) ? o : v;
    }
    return v;
}
template <typename T>
__device__ inline T wave_min(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = (o < v) ? o : v;
    }
    return v;
}

template <typename T>
__device__ inline T dev_sqrt(T x);
template <>
__device__ inline float dev_sqrt<float>(float x) { return __builtin_sqrtf(x); }
template <>
__device__ inline double dev_sqrt<double>(double x) { return __builtin_sqrt(x); }

}  // namespace gputils
