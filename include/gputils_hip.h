/*
 * gputils_hip.h -- C ABI of libgputils_hip.so, the MI355X (gfx950) kernels
 * behind GPUtils' batched dense linear-algebra API.
 *
 * Conventions
 *  - Matrices are column-major with an explicit leading dimension.  A batch
 *    of matrices is addressed as base pointer + element stride between
 *    consecutive matrices (a GPUtils DTensor (m,n,k) has stride m*n).
 *  - Sizes are int64_t.  `stream` is a hipStream_t passed as void*
 *    (NULL = the null stream).  All compute calls are asynchronous on that
 *    stream, allocate nothing and are therefore capturable in HIP graphs.
 *  - Return value: 0 on success; -i if argument i (1-based) is invalid
 *    (LAPACK convention, nothing is launched); a positive value for a HIP
 *    runtime failure.  gputils_last_error() returns a message for the last
 *    failing call of the calling thread.
 *  - Numerical failures are reported per matrix through device `info`
 *    arrays (LAPACK convention: 0 = success, i > 0 = failure at step i).
 *
 * Each entry point names the place in GPUtils' include/tensor.cuh whose
 * vendor-library call it replaces.
 */
#ifndef GPUTILS_HIP_H
#define GPUTILS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPUTILS_OK 0

/* ------------------------------------------------------------ runtime --- */
/* Message describing the last failing call on this thread ("" if none). */
const char *gputils_last_error(void);
/* Library version string. */
const char *gputils_version(void);
/* Number of visible GPUs (0 when no GPU is present). */
int gputils_device_count(int *count);

/* Device allocation; replaces cudaMalloc/cudaFree of Session::cudaAllocate
 * and DTensor::destroy (tensor.cuh:209-215, 283-294). */
int gputils_malloc(void **ptr, size_t bytes);
int gputils_free(void *ptr);
/* Synchronous copies with null-stream semantics (tensor.cuh:687, 893, 1142,
 * 1150, 1204, 1282).  kind: 0 host->device, 1 device->host,
 * 2 device->device, 3 inferred from the pointers. */
int gputils_memcpy(void *dst, const void *src, size_t bytes, int kind);
int gputils_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);
/* Synchronous memset (tensor.cuh:1114) and its stream-ordered variant. */
int gputils_memset(void *ptr, int value, size_t bytes);
/* Byte fill by a kernel (no memset node, so it is safe inside HIP graphs). */
int gputils_fill_bytes_async(void *ptr, unsigned char value, size_t bytes, void *stream);

/* Streams (Session, tensor.cuh:154-166, 232-246).  Streams are created
 * blocking with default priority. */
int gputils_stream_create(void **stream);
int gputils_stream_destroy(void *stream);
int gputils_stream_synchronize(void *stream);
int gputils_device_synchronize(void);

/* Device array of per-matrix pointers base + i*stride_bytes, i < count,
 * built by a kernel (DTensor::initialisePointersToMatricesData,
 * tensor.cuh:672-688, does this with a host loop and a copy). */
int gputils_fill_pointer_array(void **dst, const void *base, int64_t stride_bytes,
                               int64_t count, void *stream);

/* --------------------------------------------------- BLAS-1 reductions --- */
/* Whole-array reductions over n contiguous elements.  The result is written
 * to *result (device memory).  `workspace` is device memory of at least
 * gputils_reduce_workspace_bytes() bytes.
 * Replace cublas?dot (tensor.cuh:968-992), ?nrm2 (994-1010), ?asum
 * (1012-1028), I?amax/I?amin + one element copy (1030-1072: maxAbs/minAbs
 * return |x| at that index, which is the max/min of |x|). */
size_t gputils_reduce_workspace_bytes(void);
int gputils_sdot(int64_t n, const float *x, const float *y, float *result, void *workspace, void *stream);
int gputils_ddot(int64_t n, const double *x, const double *y, double *result, void *workspace, void *stream);
int gputils_snrm2(int64_t n, const float *x, float *result, void *workspace, void *stream);
int gputils_dnrm2(int64_t n, const double *x, double *result, void *workspace, void *stream);
int gputils_sasum(int64_t n, const float *x, float *result, void *workspace, void *stream);
int gputils_dasum(int64_t n, const double *x, double *result, void *workspace, void *stream);
int gputils_samax(int64_t n, const float *x, float *result, void *workspace, void *stream);
int gputils_damax(int64_t n, const double *x, double *result, void *workspace, void *stream);
int gputils_samin(int64_t n, const float *x, float *result, void *workspace, void *stream);
int gputils_damin(int64_t n, const double *x, double *result, void *workspace, void *stream);

/* x <- alpha x (cublas?scal, tensor.cuh:1210-1237). */
int gputils_sscal(int64_t n, float alpha, float *x, void *stream);
int gputils_dscal(int64_t n, double alpha, double *x, void *stream);
/* y <- y + alpha x (cublas?axpy, tensor.cuh:1239-1276). */
int gputils_saxpy(int64_t n, float alpha, const float *x, float *y, void *stream);
int gputils_daxpy(int64_t n, double alpha, const double *x, double *y, void *stream);

/* B_i = A_i^T, A_i m x n, B_i n x m (cublas?geam per matrix,
 * tensor.cuh:1167-1197). */
int gputils_stranspose_batched(int64_t m, int64_t n, const float *A, int64_t lda, int64_t strideA,
                               float *B, int64_t ldb, int64_t strideB, int64_t batch, void *stream);
int gputils_dtranspose_batched(int64_t m, int64_t n, const double *A, int64_t lda, int64_t strideA,
                               double *B, int64_t ldb, int64_t strideB, int64_t batch, void *stream);

/* Plane rotation of two strided vectors: x_k <- c x_k + s y_k,
 * y_k <- c y_k - s x_k for k < n, incx, incy >= 1 (cublas?rot behind
 * DTensor::applyRightGivensRotation, columns i and j with increment 1, and
 * applyLeftGivensRotation, rows i and j with increment numRows;
 * tensor.cuh:1074-1104, which pass s = minus_s). */
int gputils_srot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s, void *stream);
int gputils_drot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s, void *stream);
/* GivensAnnihilator::annihilate(i, k, j) (tensor.cuh:2261-2311 and its
 * rhypot kernel 2250-2258) in one launch: with r = 1 / hypot(A(i,j),
 * A(k,j)), c = A(i,j) r and s = A(k,j) r, rows i and k of the m x ncols
 * matrix A are rotated as by gputils_?rot, which zeroes A(k,j) and puts
 * hypot(A(i,j), A(k,j)) in A(i,j).  c and s never leave the device.
 * Indices are checked (i, k < m; j < ncols). */
int gputils_sgivens_annihilate(int64_t m, int64_t ncols, float *A, int64_t lda, int64_t i, int64_t k, int64_t j,
                               void *stream);
int gputils_dgivens_annihilate(int64_t m, int64_t ncols, double *A, int64_t lda, int64_t i, int64_t k, int64_t j,
                               void *stream);

/* ---------------------------------------------------------------- GEMM --- */
/* C_i <- alpha A_i B_i + beta C_i; A_i m x k, B_i k x n, C_i m x n
 * (cublas?gemmBatched / ?gemm, tensor.cuh:1286-1338).  C may alias B
 * exactly (same pointer, ld and stride), as Nullspace::project does
 * (tensor.cuh:2084): every matrix is then read completely before it is
 * overwritten.  Other partial overlaps are not supported.
 * mode (float only): GPUTILS_GEMM_FP32 = fp32 products and fp32
 * accumulation on MFMA/VALU; GPUTILS_GEMM_BF16X3 = each operand split into
 * two bfloat16 terms, three bf16 MFMA products with fp32 accumulation. */
#define GPUTILS_GEMM_FP32 0
#define GPUTILS_GEMM_BF16X3 1
int gputils_sgemm_strided_batched(int64_t m, int64_t n, int64_t k, float alpha,
                                  const float *A, int64_t lda, int64_t strideA,
                                  const float *B, int64_t ldb, int64_t strideB, float beta,
                                  float *C, int64_t ldc, int64_t strideC, int64_t batch,
                                  int mode, void *stream);
int gputils_dgemm_strided_batched(int64_t m, int64_t n, int64_t k, double alpha,
                                  const double *A, int64_t lda, int64_t strideA,
                                  const double *B, int64_t ldb, int64_t strideB, double beta,
                                  double *C, int64_t ldc, int64_t strideC, int64_t batch,
                                  void *stream);

/* ------------------------------------------------------------ Cholesky --- */
/* Lower Cholesky A_i = L_i L_i^T in place; the strictly upper triangle is
 * not referenced.  info[i] = 0 or the order of the first non-positive
 * leading minor (cusolverDn?potrfBatched, tensor.cuh:2136-2159). */
int gputils_spotrf_batched(int64_t n, float *A, int64_t lda, int64_t strideA, int *info,
                           int64_t batch, void *stream);
int gputils_dpotrf_batched(int64_t n, double *A, int64_t lda, int64_t strideA, int *info,
                           int64_t batch, void *stream);
/* Solve L_i L_i^T X_i = B_i in place for nrhs >= 1 columns
 * (cusolverDn?potrsBatched, tensor.cuh:2161-2197, which only allows
 * nrhs = 1).  info is written with 0 (argument checks happen on the host). */
int gputils_spotrs_batched(int64_t n, int64_t nrhs, const float *L, int64_t lda, int64_t strideL,
                           float *B, int64_t ldb, int64_t strideB, int *info, int64_t batch,
                           void *stream);
int gputils_dpotrs_batched(int64_t n, int64_t nrhs, const double *L, int64_t lda, int64_t strideL,
                           double *B, int64_t ldb, int64_t strideB, int *info, int64_t batch,
                           void *stream);
/* Fused factor + solve in FP64: A_i <- L_i, B_i <- A_i^{-1} B_i in one
 * launch (the pair factorise(); solve(b) of CholeskyBatchFactoriser).
 * info[i] = 0 or the order of the first non-positive leading minor; a
 * system whose factorisation fails keeps its B_i (as LAPACK dposv) for
 * n > 128, where info must not be NULL. */
int gputils_dposv_batched(int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t strideA,
                          double *B, int64_t ldb, int64_t strideB, int *info, int64_t batch,
                          void *stream);
/* Mixed-precision solve (no reference counterpart: a new entry point next
 * to CholeskyBatchFactoriser).  Factor fl32(A_i) on the GPU, solve in
 * FP32 and refine in FP64 until, for every right-hand side,
 *   ||b - A x||_inf <= tol * (||A||_inf ||x||_inf + ||b||_inf),
 * or max_iter refinement steps were made.  A_i and B_i are read only
 * (FP64, lower triangle of A_i used); X_i receives the FP64 solution.
 * iters[i] = refinement steps used, or -1 if the FP32 factorisation failed,
 * refinement did not converge, or a norm in the stopping test was not
 * finite; those matrices are then solved by a second, FP64 launch (as
 * LAPACK DSPOSV falls back), which also reads A_i and B_i only, and
 * info[i] holds the FP64 status (0 or the failing minor).
 * n <= 128 runs the fused mixed kernel (work may be NULL).  Larger n is
 * solved in FP64 only (iters[i] = -1 for all i) on a copy of A in `work`,
 * gputils_dposv_mixed_workspace_bytes(n, batch) bytes of device memory. */
int64_t gputils_dposv_mixed_workspace_bytes(int64_t n, int64_t batch);
int gputils_dposv_mixed_batched(int64_t n, int64_t nrhs, const double *A, int64_t lda,
                                int64_t strideA, const double *B, int64_t ldb, int64_t strideB,
                                double *X, int64_t ldx, int64_t strideX, double tol, int max_iter,
                                int *iters, int *info, void *work, int64_t batch, void *stream);
/* gputils_dposv_mixed_batched without the work argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  No reference counterpart.  It
 * allocates nothing: for n > 128 the FP64 copies of A go to a 16 MiB
 * array of the library that each GPU holds for the life of the process,
 * so n*n*batch*8 must not exceed 16 MiB (larger calls return -16 and
 * launch nothing); n <= 128 needs no scratch.  Results are those of
 * gputils_dposv_mixed_batched.  Calls on one GPU share the array: calls
 * that overlap in time on one GPU (different streams, or threads) race.
 * Every call overwrites what it uses before reading it. */
int gputils_dposv_mixed_batched_noalloc(int64_t n, int64_t nrhs, const double *A, int64_t lda,
                                        int64_t strideA, const double *B, int64_t ldb, int64_t strideB,
                                        double *X, int64_t ldx, int64_t strideX, double tol, int max_iter,
                                        int *iters, int *info, int64_t batch, void *stream);

/* ------------------------------------------------------- QR / least sq --- */
/* Batched least squares min ||A_i x - b_i|| for m >= n (cublas?gelsBatched
 * behind DTensor::leastSquaresBatched, tensor.cuh:1340-1394; any nrhs).
 * A_i is overwritten with its Householder QR factors (LAPACK storage) and
 * the first n rows of each column of B_i with the solution. */
int gputils_sgels_batched(int64_t m, int64_t n, int64_t nrhs, float *A, int64_t lda, int64_t strideA,
                          float *B, int64_t ldb, int64_t strideB, int *info, int64_t batch, void *stream);
int gputils_dgels_batched(int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda, int64_t strideA,
                          double *B, int64_t ldb, int64_t strideB, int *info, int64_t batch, void *stream);
/* Householder QR, LAPACK storage: R on and above the diagonal, reflector
 * vectors below, tau_i (n values) at tau + i*strideTau (cusolverDn?geqrf of
 * QRFactoriser::factorise, tensor.cuh:1866-1889). */
int gputils_sgeqrf_batched(int64_t m, int64_t n, float *A, int64_t lda, int64_t strideA, float *tau,
                           int64_t strideTau, int *info, int64_t batch, void *stream);
int gputils_dgeqrf_batched(int64_t m, int64_t n, double *A, int64_t lda, int64_t strideA, double *tau,
                           int64_t strideTau, int *info, int64_t batch, void *stream);
/* C_i <- Q_i^T C_i (trans = 1) or Q_i C_i (trans = 0) with Q_i from the
 * first k reflectors of a geqrf result (cusolverDn?ormqr side L,
 * tensor.cuh:1896-1902 and 1946-1952). */
int gputils_sormqr_batched(int trans, int64_t m, int64_t ncols, int64_t k, const float *A, int64_t lda,
                           int64_t strideA, const float *tau, int64_t strideTau, float *C, int64_t ldc,
                           int64_t strideC, int64_t batch, void *stream);
int gputils_dormqr_batched(int trans, int64_t m, int64_t ncols, int64_t k, const double *A, int64_t lda,
                           int64_t strideA, const double *tau, int64_t strideTau, double *C, int64_t ldc,
                           int64_t strideC, int64_t batch, void *stream);
/* B_i <- R_i^{-1} B_i, R_i upper triangular non-unit n x n (cublas?trsm
 * left/upper/N/non-unit, tensor.cuh:1903-1907). */
int gputils_strsm_upper_batched(int64_t n, int64_t nrhs, const float *R, int64_t ldr, int64_t strideR,
                                float *B, int64_t ldb, int64_t strideB, int64_t batch, void *stream);
int gputils_dtrsm_upper_batched(int64_t n, int64_t nrhs, const double *R, int64_t ldr, int64_t strideR,
                                double *B, int64_t ldb, int64_t strideB, int64_t batch, void *stream);

/* Ragged least squares (no reference counterpart: the reference needs
 * uniform sizes).  A plan holds the per-system sizes and offsets, bucketed
 * by working-set size and sorted by cost (m n^2, largest first) in device
 * memory; executing it allocates nothing.  System s is the m[s] x n[s]
 * column-major FP64 matrix at A + offA[s] (ld = m[s]) and the m[s]-vector at
 * B + offB[s]; 1 <= n[s] <= m[s] <= 512, n[s] <= 64.  Execution solves every
 * system with an FP32 Householder QR (FP64 norms and dot products), then
 * refine_steps FP64 refinements of the augmented system [I A; A^T 0][r; x] =
 * [b; 0] (Bjorck), which also converge for inconsistent systems; A is read
 * only and the solution overwrites the first n[s] entries of each b. */
typedef struct gputils_gels_plan *gputils_gels_plan_t;
int gputils_gels_ragged_plan_create(gputils_gels_plan_t *plan, int64_t count, const int32_t *m,
                                    const int32_t *n, const int64_t *offA, const int64_t *offB);
int gputils_gels_ragged_plan_destroy(gputils_gels_plan_t plan);
int gputils_dgels_ragged(gputils_gels_plan_t plan, const double *A, double *B, int refine_steps,
                         void *stream);
/* Launch shape per bucket.  A plan has one bucket per LDS size class in use
 * (FP32 m*n <= 2048, 4096, 8192, 16384, 32768; largest first), one launch
 * each.  Its shape: the factorisation method (0: A staged in LDS and
 * factorised column by column, one barrier per column; 1: panels of
 * 64 / waves columns held in registers, one barrier per panel), the
 * wavefronts per workgroup (2, 4, 8 or 16 for method 0; 8 or 16 for
 * method 1), the workgroups per CU (the grid is min(systems in the
 * bucket, CUs * per_cu)) and the scheduling: static (queue = 0: the
 * systems are assigned to the grid's workgroups longest-processing-time
 * first by a cost model) or a work queue (queue = 1: each workgroup takes
 * the next system, most expensive first, from a device counter that the
 * last workgroup resets, so nothing is cleared between launches).  A new
 * plan uses method 0, 8 wavefronts, as many workgroups per CU as fit (LDS,
 * 32 wavefronts per CU) and static scheduling.  For one method every shape gives bit-identical results; the two
 * methods round the FP32 factorisation differently (the refined solutions
 * agree to the refinement's accuracy).
 *
 * describe: writes {maxe, systems, method, waves, per_cu, grid, queue} for
 * the first min(capacity, buckets) buckets to desc (7 int32 each) and the
 * bucket count to *nbuckets (if not NULL).
 * configure: sets one bucket's shape.
 * tune: the autotuner.  For every bucket it times each shape (both
 * methods, their wavefront counts, every per_cu the CU can hold, both
 * schedulings) on the
 * caller's A and a device copy of B (B itself is not written), with
 * refine_steps refinements, on stream, and keeps the fastest; then it
 * chooses between concurrent and sequential buckets (below).  It
 * allocates the copy of B and so is a plan-time call like plan_create.
 * configure and tune synchronise the device first and rewrite the plan's
 * device tables: no execution of the plan may be pending or run
 * concurrently.  While a bucket uses the work queue, executions of one
 * plan must not overlap each other either (they would share its counters;
 * executions on one stream never overlap). */
int gputils_gels_ragged_plan_describe(gputils_gels_plan_t plan, int32_t *desc, int capacity, int *nbuckets);
int gputils_gels_ragged_plan_configure(gputils_gels_plan_t plan, int bucket, int method, int waves, int per_cu,
                                       int queue);
int gputils_gels_ragged_plan_tune(gputils_gels_plan_t plan, const double *A, const double *B, int refine_steps,
                                  void *stream);
/* Whether a plan's buckets run concurrently (1: each bucket after the
 * first on a stream of the plan, forked from and joined back to the
 * caller's stream by events, so the call stays ordered on it) or one after
 * another on the caller's stream (0).  A new plan runs them concurrently;
 * tune times both with the tuned shapes and keeps the faster.  The results
 * are the same either way (the buckets write disjoint systems). */
int gputils_gels_ragged_plan_set_concurrency(gputils_gels_plan_t plan, int concurrent);
int gputils_gels_ragged_plan_get_concurrency(gputils_gels_plan_t plan, int *concurrent);

/* ------------------------------------------------------------ graphs --- */
/* Capture the work enqueued on a (non-null) stream into an executable HIP
 * graph and replay it (no reference counterpart; used for repeated
 * Nullspace::project calls).  Between begin and end only asynchronous work
 * may be enqueued on that stream. */
typedef struct gputils_graph *gputils_graph_t;
int gputils_graph_capture_begin(void *stream);
int gputils_graph_capture_end(void *stream, gputils_graph_t *graph);
int gputils_graph_launch(gputils_graph_t graph, void *stream);
int gputils_graph_destroy(gputils_graph_t graph);

/* ------------------------------------------------------------ SVD / rank --- */
/* A_i = U_i diag(S_i) VT_i for tall m x n matrices (m >= n >= 1):
 * replaces the per-matrix cusolverDn?gesvd loop of Svd::factorise
 * (tensor.cuh:1624-1676, jobu 'A' when U != NULL else 'N', jobvt 'A').
 * S_i: n values, descending.  VT_i: n x n (ld ldvt), row j = v_j^T with the
 * largest-magnitude entry of v_j negative.  U_i: m x m (ld ldu) or NULL.
 * A is read only.  Householder QR and the final Jacobi sweeps in FP64 for
 * both types; for float data with 4 <= n <= 64 and VT == NULL the sweeps
 * start in FP32 and FP64 sweeps finish them from the FP32 result
 * (gputils_gesvd_set_fp32_sweeps turns that off).  info[i] = 0, or 1 if the FP64 sweeps did not converge;
 * sweeps (may be NULL) receives the FP64 sweep count (a final convergence
 * check made on W^T W instead of by a sweep is not counted) in bits 0..15 and the
 * FP32 one in bits 16..31.  work: gputils_gesvd_workspace_bytes(m, n, batch)
 * bytes of device memory (0 bytes, work may be NULL, when the matrix fits
 * in LDS).
 * VT may be NULL (jobvt 'N', no reference counterpart; used by the nullspace
 * construction): the sweeps then run on R^T and U_R is the accumulated
 * rotation matrix, which gives the same S and the same column spaces of U
 * (column signs may differ from the VT != NULL result). */
/* Whether float SVDs run their first sweeps in FP32 (1, the default) or
 * all sweeps in FP64 (0); returns the previous setting.  Process-wide; no
 * reference counterpart (a switch for measuring the two schemes). */
int gputils_gesvd_set_fp32_sweeps(int enable);
int64_t gputils_gesvd_workspace_bytes(int m, int n, int64_t batch);
int gputils_sgesvd_batched(int m, int n, const float *A, int64_t lda, int64_t strideA, float *S, int64_t strideS,
                           float *U, int64_t ldu, int64_t strideU, float *VT, int64_t ldvt, int64_t strideVT,
                           void *work, int *info, int *sweeps, int64_t batch, void *stream);
int gputils_dgesvd_batched(int m, int n, const double *A, int64_t lda, int64_t strideA, double *S, int64_t strideS,
                           double *U, int64_t ldu, int64_t strideU, double *VT, int64_t ldvt, int64_t strideVT,
                           void *work, int *info, int *sweeps, int64_t batch, void *stream);

/* rank[i] = #{j < n : S_i[j] > eps} (Svd::rank and k_countNonzeroSingularValues,
 * tensor.cuh:1484-1491, 1600-1609).  Unlike the reference, the count is
 * written, not accumulated, so repeated calls give the same result. */
int gputils_ssvd_rank_batched(int n, const float *S, int64_t strideS, float eps, unsigned *rank, int64_t batch,
                              void *stream);
int gputils_dsvd_rank_batched(int n, const double *S, int64_t strideS, double eps, unsigned *rank, int64_t batch,
                              void *stream);

/* Nullspace construction from the left singular vectors of A_i^T
 * (Nullspace ctor loop, tensor.cuh:2058-2078): with nul = n - rank[i],
 * N_i[:, 0:nul] = U_i[:, rank[i]:n], remaining columns zero, and
 * P_i = N_i N_i^T; all matrices n x n. */
int gputils_snullspace_build_batched(int n, const float *U, int64_t ldu, int64_t strideU, const unsigned *rank,
                                     float *N, int64_t ldn, int64_t strideN, float *P, int64_t ldp, int64_t strideP,
                                     int64_t batch, void *stream);
int gputils_dnullspace_build_batched(int n, const double *U, int64_t ldu, int64_t strideU, const unsigned *rank,
                                     double *N, int64_t ldn, int64_t strideN, double *P, int64_t ldp,
                                     int64_t strideP, int64_t batch, void *stream);

/* Nullspace::project (tensor.cuh:2083-2085, b_i <- P_i b_i with
 * P_i = N_i N_i^T) formed as b_i <- N_i (N_i^T b_i) from the first
 * n - rank[i] columns of N_i (as gputils_?nullspace_build_batched leaves
 * them): the same projection up to rounding, reading n (n - rank[i])
 * elements of N_i instead of the n^2 of P_i.  b_i: n contiguous elements,
 * batch stride strideB; rank[i] >= n gives b_i = 0. */
int gputils_snullspace_project_batched(int n, const float *N, int64_t ldn, int64_t strideN, const unsigned *rank,
                                       float *b, int64_t strideB, int64_t batch, void *stream);
int gputils_dnullspace_project_batched(int n, const double *N, int64_t ldn, int64_t strideN, const unsigned *rank,
                                       double *b, int64_t strideB, int64_t batch, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GPUTILS_HIP_H */
