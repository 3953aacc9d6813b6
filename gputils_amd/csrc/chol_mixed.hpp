// chol_mixed.hpp -- launcher of the one-wavefront-per-matrix mixed-precision
// Cholesky solve (chol_mixed.hip).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace gputils {
namespace cholw {

// Largest n of the fused mixed-precision kernel.
constexpr int kMixedMaxN = 128;

// Launches one wavefront per matrix (n <= kMixedMaxN).  iters[b] receives the
// number of refinement steps, or -1 when the FP32 factorisation failed or
// the refinement did not reach tol within max_iter steps.
int launch_posv_mixed(int n, int nrhs, const double *A, int64_t lda, int64_t strideA, const double *B, int64_t ldb,
                      int64_t strideB, double *X, int64_t ldx, int64_t strideX, double tol, int max_iter, int *iters,
                      int *info, int64_t batch, hipStream_t st);

}  // namespace cholw
}  // namespace gputils
