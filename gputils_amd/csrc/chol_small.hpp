// chol_small.hpp -- launchers of the FP32/FP64 small-n (n <= 32) Cholesky
// factor + solve (chol_small.hip).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace gputils {
namespace chols {

// Two matrices per wavefront (four for n <= 16).  write_l: A_i <- L_i (lower triangle).
// mask != nullptr: only matrices with mask[b] < 0 are processed.
template <typename T>  // float or double
int launch_posv_small(int n, int nrhs, T *A, int64_t lda, int64_t strideA, const T *B, int64_t ldb,
                      int64_t strideB, T *X, int64_t ldx, int64_t strideX, int *info, int64_t batch,
                      const int *mask, bool write_l, hipStream_t st);

// Solve only, given the factor L (lower triangle read): B <- A^{-1} B in
// place; info[b] = 0.
template <typename T>  // float or double
int launch_potrs_small(int n, int nrhs, const T *L, int64_t lda, int64_t strideL, T *B, int64_t ldb,
                       int64_t strideB, int *info, int64_t batch, hipStream_t st);

}  // namespace chols
}  // namespace gputils
