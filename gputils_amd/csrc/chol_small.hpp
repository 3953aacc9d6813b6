// chol_small.hpp -- launcher of the FP64 small-n (n <= 32) fused Cholesky
// factor + solve (chol_small.hip).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace gputils {
namespace chols {

// Two matrices per wavefront (four for n <= 16).  write_l: A_i <- L_i (lower triangle).
// mask != nullptr: only matrices with mask[b] < 0 are processed.
int launch_posv_small(int n, int nrhs, double *A, int64_t lda, int64_t strideA, const double *B, int64_t ldb,
                      int64_t strideB, double *X, int64_t ldx, int64_t strideX, int *info, int64_t batch,
                      const int *mask, bool write_l, hipStream_t st);

// Solve only, given the factor L (lower triangle read): B <- A^{-1} B in
// place; info[b] = 0.
int launch_potrs_small(int n, int nrhs, const double *L, int64_t lda, int64_t strideL, double *B, int64_t ldb,
                       int64_t strideB, int *info, int64_t batch, hipStream_t st);

}  // namespace chols
}  // namespace gputils
