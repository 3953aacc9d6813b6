/*
 * lapack_loop.c -- host LAPACK batch loops for the CPU baselines that
 * bench.py reports beside the GPU numbers (never the thing measured on the
 * GPU, and never part of the shipped library).
 *
 * The reference hands each batched operation to cuSOLVER/cuBLAS
 * (include/tensor.cuh:2136-2197 potrf/potrs, 1340-1394 gels, 1624-1676
 * gesvd, 1286-1338 gemm); the CPU counterpart of those calls is LAPACK.
 * This file loops LAPACK over the matrices of a batch: the LAPACK is the
 * OpenBLAS build that scipy bundles, opened with dlopen at run time (its
 * routines carry a scipy_ prefix), pinned to one thread per call, with
 * OpenMP spreading the matrices over the host cores.  Fortran calling
 * convention: every argument by reference, one hidden size_t length per
 * character argument, after the others.
 *
 * Layouts follow the GPU side: matrix b of a batch is column-major at
 * base + b * stride.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

typedef void (*potrf_t)(const char *, const int *, double *, const int *, int *, size_t);
typedef void (*potrs_t)(const char *, const int *, const int *, const double *, const int *, double *,
                        const int *, int *, size_t);
typedef void (*gels_t)(const char *, const int *, const int *, const int *, double *, const int *, double *,
                       const int *, double *, const int *, int *, size_t);
typedef void (*gesvd_t)(const char *, const char *, const int *, const int *, double *, const int *, double *,
                        double *, const int *, double *, const int *, double *, const int *, int *, size_t,
                        size_t);
typedef void (*sgemm_t)(const char *, const char *, const int *, const int *, const int *, const float *,
                        const float *, const int *, const float *, const int *, const float *, float *,
                        const int *, size_t, size_t);
typedef void (*dgemm_t)(const char *, const char *, const int *, const int *, const int *, const double *,
                        const double *, const int *, const double *, const int *, const double *, double *,
                        const int *, size_t, size_t);
typedef void (*sgemv_t)(const char *, const int *, const int *, const float *, const float *, const int *,
                        const float *, const int *, const float *, float *, const int *, size_t);
typedef void (*set_threads_t)(int);

static struct {
    void *handle;
    potrf_t dpotrf;
    potrs_t dpotrs;
    gels_t dgels;
    gesvd_t dgesvd;
    sgemm_t sgemm;
    dgemm_t dgemm;
    sgemv_t sgemv;
    set_threads_t set_threads;
} L;

/* Opens the LAPACK library at path; 0 on success, -1 if it or a routine is missing. */
int ll_open(const char *path) {
    if (L.handle) return 0;
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    L.dpotrf = (potrf_t)dlsym(h, "scipy_dpotrf_");
    L.dpotrs = (potrs_t)dlsym(h, "scipy_dpotrs_");
    L.dgels = (gels_t)dlsym(h, "scipy_dgels_");
    L.dgesvd = (gesvd_t)dlsym(h, "scipy_dgesvd_");
    L.sgemm = (sgemm_t)dlsym(h, "scipy_sgemm_");
    L.dgemm = (dgemm_t)dlsym(h, "scipy_dgemm_");
    L.sgemv = (sgemv_t)dlsym(h, "scipy_sgemv_");
    L.set_threads = (set_threads_t)dlsym(h, "scipy_openblas_set_num_threads");
    if (!L.dpotrf || !L.dpotrs || !L.dgels || !L.dgesvd || !L.sgemm || !L.dgemm || !L.sgemv || !L.set_threads) {
        dlclose(h);
        return -1;
    }
    L.set_threads(1);
    L.handle = h;
    return 0;
}

/* X_b = A_b^{-1} B_b (dpotrf 'L' on a copy of A_b, then dpotrs); A, B read
 * only.  Returns the number of systems whose factorisation failed. */
int ll_posv(int64_t k, int n, int nrhs, const double *A, const double *B, double *X, int threads) {
    int failed = 0;
#pragma omp parallel num_threads(threads) reduction(+ : failed)
    {
        double *F = malloc(sizeof(double) * (size_t)n * n);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < k; b++) {
            int info = 0;
            memcpy(F, A + b * (int64_t)n * n, sizeof(double) * (size_t)n * n);
            double *x = X + b * (int64_t)n * nrhs;
            memcpy(x, B + b * (int64_t)n * nrhs, sizeof(double) * (size_t)n * nrhs);
            L.dpotrf("L", &n, F, &n, &info, 1);
            if (info != 0) {
                failed++;
                continue;
            }
            L.dpotrs("L", &n, &nrhs, F, &n, x, &n, &info, 1);
        }
        free(F);
    }
    return failed;
}

/* Least squares for systems of uneven sizes (dgels 'N', one right-hand
 * side): system s is m[s] x n[s] at A + offA[s], b at B + offB[s]; the
 * solution overwrites X[offB[s] : offB[s] + n[s]] (X holds a copy of B on
 * entry).  A is read only.  Returns the number of failed systems. */
int ll_gels_ragged(int64_t count, const int32_t *m, const int32_t *n, const int64_t *offA, const int64_t *offB,
                   const double *A, double *X, int threads) {
    int failed = 0;
    int maxe = 0, maxm = 0;
    for (int64_t s = 0; s < count; s++) {
        if (m[s] * n[s] > maxe) maxe = m[s] * n[s];
        if (m[s] > maxm) maxm = m[s];
    }
#pragma omp parallel num_threads(threads) reduction(+ : failed)
    {
        double *F = malloc(sizeof(double) * (size_t)(maxe > 0 ? maxe : 1));
        int lwork = 64 * (maxm + 64) + 64;
        double *work = malloc(sizeof(double) * (size_t)lwork);
#pragma omp for schedule(dynamic, 8)
        for (int64_t s = 0; s < count; s++) {
            const int ms = m[s], ns = n[s], one = 1;
            int info = 0;
            memcpy(F, A + offA[s], sizeof(double) * (size_t)ms * ns);
            L.dgels("N", &ms, &ns, &one, F, &ms, X + offB[s], &ms, work, &lwork, &info, 1);
            if (info != 0) failed++;
        }
        free(work);
        free(F);
    }
    return failed;
}

/* C_b = A_b B_b for k FP32 products (A m x kk, B kk x n). */
void ll_sgemm(int64_t k, int m, int n, int kk, const float *A, const float *B, float *C, int threads) {
    const float one = 1.0f, zero = 0.0f;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t b = 0; b < k; b++)
        L.sgemm("N", "N", &m, &n, &kk, &one, A + b * (int64_t)m * kk, &m, B + b * (int64_t)kk * n, &kk, &zero,
                C + b * (int64_t)m * n, &m, 1, 1);
}

/* Nullspace build of k fat m x n matrices (m <= n) as the reference's
 * constructor does it: dgesvd of A_b^T (n x m, jobu 'A', jobvt 'N'),
 * rank = #{sigma > eps}, N = U[:, rank:], P = N N^T (dgemm).  A is
 * column-major m x n per matrix; P receives n x n per matrix.  Returns
 * the number of failed SVDs. */
int ll_nullspace(int64_t k, int m, int n, const double *A, double eps, double *P, int threads) {
    int failed = 0;
#pragma omp parallel num_threads(threads) reduction(+ : failed)
    {
        double *At = malloc(sizeof(double) * (size_t)n * m);
        double *S = malloc(sizeof(double) * (size_t)m);
        double *U = malloc(sizeof(double) * (size_t)n * n);
        double dummy = 0.0, wq = 0.0;
        int lwork = -1, info = 0, ldvt = 1;
        L.dgesvd("A", "N", &n, &m, At, &n, S, U, &n, &dummy, &ldvt, &wq, &lwork, &info, 1, 1);
        lwork = (int)wq + 1;
        double *work = malloc(sizeof(double) * (size_t)lwork);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < k; b++) {
            const double *a = A + b * (int64_t)m * n;
            for (int i = 0; i < m; i++)
                for (int j = 0; j < n; j++) At[j + (int64_t)i * n] = a[i + (int64_t)j * m];
            L.dgesvd("A", "N", &n, &m, At, &n, S, U, &n, &dummy, &ldvt, work, &lwork, &info, 1, 1);
            if (info != 0) failed++;
            int rank = 0;
            for (int i = 0; i < m; i++) rank += S[i] > eps;
            int nul = n - rank;
            const double one = 1.0, zero = 0.0;
            double *p = P + b * (int64_t)n * n;
            if (nul > 0)
                L.dgemm("N", "T", &n, &n, &nul, &one, U + (int64_t)rank * n, &n, U + (int64_t)rank * n, &n, &zero,
                        p, &n, 1, 1);
            else
                memset(p, 0, sizeof(double) * (size_t)n * n);
        }
        free(work);
        free(U);
        free(S);
        free(At);
    }
    return failed;
}

/* calls rounds of x_b <- P_b x_b (FP32 sgemv, n x n) over k matrices: the
 * CPU counterpart of repeated Nullspace::project calls. */
void ll_project(int64_t k, int n, const float *P, float *x, int calls, int threads) {
    const float one = 1.0f, zero = 0.0f;
    const int inc = 1;
#pragma omp parallel num_threads(threads)
    {
        float *y = malloc(sizeof(float) * (size_t)n);
        for (int c = 0; c < calls; c++) {
#pragma omp for schedule(static)
            for (int64_t b = 0; b < k; b++) {
                L.sgemv("N", &n, &n, &one, P + b * (int64_t)n * n, &n, x + b * (int64_t)n, &inc, &zero, y, &inc, 1);
                memcpy(x + b * (int64_t)n, y, sizeof(float) * (size_t)n);
            }
        }
        free(y);
    }
}

int ll_max_threads(void) { return omp_get_max_threads(); }
