/*
 * oracle.h -- FP64 host restatement of GPUtils' batched dense linear algebra.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It is used by tests/ (as the parity checker), by
 * __graft_entry__.smoke() (as the checker of one small GPU call) and by
 * bench.py's cpu_baseline leg (as the timed host baseline).
 *
 * Every routine restates the algorithm behind one reference call site
 * (the reference delegates the arithmetic to cuBLAS/cuSOLVER, which follow
 * the LAPACK definitions restated here).  Citations use
 * include/tensor.cuh:LINE of GPUtils.
 *
 * All matrices are column-major with an explicit leading dimension.
 */
#ifndef GPUTILS_ORACLE_H
#define GPUTILS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C <- beta*C + alpha*A*B  (tensor.cuh:1286-1338, cublas?gemm semantics). */
void orc_dgemm(int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb,
               double beta, double *C, int64_t ldc);

/* Lower Cholesky A = L L^T in place (tensor.cuh:2136-2146, LAPACK dpotrf 'L').
 * Returns info: 0 ok, j>0 if the leading minor of order j is not positive
 * definite (the factorisation stops there). The upper triangle is untouched. */
int orc_dpotrf_lower(int64_t n, double *A, int64_t lda);

/* Solve L L^T X = B in place (tensor.cuh:2162-2178, LAPACK dpotrs 'L'). */
void orc_dpotrs_lower(int64_t n, int64_t nrhs, const double *L, int64_t lda,
                      double *B, int64_t ldb);

/* Householder QR, LAPACK storage: R on/above the diagonal, v below, tau[j]
 * (tensor.cuh:1866-1876, LAPACK dgeqrf/dgeqr2 with the dlarfg reflector). */
void orc_dgeqrf(int64_t m, int64_t n, double *A, int64_t lda, double *tau);

/* C <- Q^T C (trans=1) or C <- Q C (trans=0), Q = H_0 ... H_{k-1}
 * (tensor.cuh:1896-1902 and 1946-1952, LAPACK dormqr side 'L'). */
void orc_dormqr(int trans, int64_t m, int64_t ncols, int64_t k,
                const double *A, int64_t lda, const double *tau,
                double *C, int64_t ldc);

/* Solve R X = B, R upper triangular non-unit (tensor.cuh:1903-1907,
 * cublas?trsm left/upper/N/non-unit). */
void orc_dtrsm_lun(int64_t n, int64_t nrhs, const double *R, int64_t ldr,
                   double *B, int64_t ldb);

/* Least squares min ||A x - b|| for m >= n, x written to B[0:n]; A is
 * overwritten with its QR factors (tensor.cuh:1340-1366, cublas?gelsBatched). */
void orc_dgels(int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda,
               double *B, int64_t ldb);

/* SVD of a tall m x n matrix (m >= n) by one-sided Jacobi in FP64
 * (tensor.cuh:1624-1649: gesvd jobu 'A'/'N', jobvt 'A').
 * S: n singular values, descending.  VT: n x n (ldvt) holding V^T.
 * U: m x m (ldu) or NULL.  Columns of U/V for sigma > 0 follow the
 * convention that the largest-magnitude entry of each right singular vector
 * is negative; the trailing columns of U complete an orthonormal basis.
 * A is destroyed.  Returns the number of sweeps. */
int orc_dgesvd(int64_t m, int64_t n, double *A, int64_t lda, double *S,
               double *U, int64_t ldu, double *VT, int64_t ldvt);

/* #{ i : S[i] > eps }  (tensor.cuh:1484-1491 and 1600-1609). */
int orc_rank(int64_t n, const double *S, double eps);

/* Nullspace of a fat m x n matrix A (m <= n), tensor.cuh:2046-2079:
 * N (n x n, zero padded, first `nullity` columns span ker A) and
 * P = N N^T (n x n).  Returns the nullity. */
int orc_nullspace(int64_t m, int64_t n, const double *A, int64_t lda,
                  double eps, double *N, double *P);

/* Plane rotation of two strided vectors, the BLAS drot definition behind
 * DTensor::applyRight/LeftGivensRotation (tensor.cuh:1074-1104):
 * x_k <- c x_k + s y_k, y_k <- c y_k - s x_k, k < n (incx, incy >= 1). */
void orc_drot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s);

/* GivensAnnihilator::annihilate (tensor.cuh:2261-2311, kernel 2250-2258):
 * r = 1 / hypot(A(i,j), A(k,j)), c = A(i,j) r, s = A(k,j) r, then rows i
 * and k of the m x ncols matrix A rotated by (c, s) as in orc_drot, which
 * zeroes A(k,j). */
void orc_dgivens_annihilate(int64_t m, int64_t ncols, double *A, int64_t lda, int64_t i, int64_t k,
                            int64_t j);

/* Batched host loops (OpenMP over matrices) used for the CPU baseline. */
int orc_dposv_batched(int64_t n, int64_t nrhs, double *A, double *B,
                      int64_t batch, int threads);
void orc_dgemm_batched(int64_t m, int64_t n, int64_t k, const double *A,
                       const double *B, double *C, int64_t batch, int threads);
void orc_sgemm_batched(int64_t m, int64_t n, int64_t k, const float *A,
                       const float *B, float *C, int64_t batch, int threads);
void orc_dgels_ragged(int64_t count, const int32_t *m, const int32_t *n,
                      const int64_t *offA, const int64_t *offB, double *A,
                      double *B, int threads);
int orc_nullspace_batched(int64_t m, int64_t n, const double *A, double eps, double *N, double *P,
                          int32_t *nullity, int64_t batch, int threads);

#ifdef __cplusplus
}
#endif
#endif
