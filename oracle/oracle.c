/*
 * oracle.c -- FP64 host restatement of GPUtils' batched dense linear algebra.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): only tests/, the smoke check and
 * bench.py's cpu_baseline leg load this file's shared object.  The shipped
 * library never links it and has no CPU fallback.
 *
 * The reference (include/tensor.cuh of GPUtils) delegates every operation on
 * the hot path to cuBLAS/cuSOLVER; those libraries implement the LAPACK/BLAS
 * definitions that are restated below in plain, unblocked C.  Each function
 * names the reference call site it stands for.
 *
 * How each function is pinned (checked by tests/test_oracle.py):
 *   orc_dgemm        exact-integer golden of the reference suite + numpy matmul
 *   orc_dpotrf_lower golden L of the reference suite + scipy LAPACK dpotrf
 *   orc_dpotrs_lower golden x of the reference suite + scipy LAPACK dpotrs
 *   orc_dgeqrf       scipy LAPACK dgeqrf (R, v, tau element-wise)
 *   orc_dormqr       scipy LAPACK dormqr, and Q^T Q = I
 *   orc_dtrsm_lun    back substitution checked against scipy solve_triangular
 *   orc_dgels        golden residual 80.003169364198072 + scipy LAPACK dgels
 *   orc_dgesvd       golden sigma/U/V of the reference suite + scipy dgesvd
 *   orc_rank         golden ranks 3,2,1 of the reference suite
 *   orc_nullspace    A N = 0, N^T N = I, P = I - U_r U_r^T (closed form)
 *   orc_drot         scipy BLAS drot, and the rotated values the reference
 *                    suite expects of applyLeft/RightGivensRotation
 *   orc_dgivens_annihilate  golden entries of the reference suite, A(k,j) = 0,
 *                    A(i,j) = hypot, and the two rows' norms preserved
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((i) + (int64_t)(j) * (ld))

/* ---------------------------------------------------------------- GEMM --- */
/* tensor.cuh:1286-1311 (double) / 1313-1338 (float): C_i <- beta C_i + alpha A_i B_i. */
void orc_dgemm(int64_t m, int64_t n, int64_t k, double alpha, const double *A,
               int64_t lda, const double *B, int64_t ldb, double beta,
               double *C, int64_t ldc) {
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int64_t p = 0; p < k; ++p) acc += A[IDX(i, p, lda)] * B[IDX(p, j, ldb)];
            double c = (beta == 0.0) ? 0.0 : beta * C[IDX(i, j, ldc)];
            C[IDX(i, j, ldc)] = c + alpha * acc;
        }
    }
}

/* ------------------------------------------------------------ Cholesky --- */
/* tensor.cuh:2136-2146: cusolverDn?potrfBatched(LOWER).  LAPACK dpotf2 'L'
 * (left-looking Crout form): for each column j,
 *   d = a_jj - sum_{p<j} l_jp^2; fail with info=j+1 unless d > 0;
 *   l_jj = sqrt(d); l_ij = (a_ij - sum_{p<j} l_ip l_jp) / l_jj, i > j. */
int orc_dpotrf_lower(int64_t n, double *A, int64_t lda) {
    for (int64_t j = 0; j < n; ++j) {
        double d = A[IDX(j, j, lda)];
        for (int64_t p = 0; p < j; ++p) d -= A[IDX(j, p, lda)] * A[IDX(j, p, lda)];
        if (!(d > 0.0)) { /* also catches NaN */
            A[IDX(j, j, lda)] = d;
            return (int)(j + 1);
        }
        double ljj = sqrt(d);
        A[IDX(j, j, lda)] = ljj;
        for (int64_t i = j + 1; i < n; ++i) {
            double s = A[IDX(i, j, lda)];
            for (int64_t p = 0; p < j; ++p) s -= A[IDX(i, p, lda)] * A[IDX(j, p, lda)];
            A[IDX(i, j, lda)] = s / ljj;
        }
    }
    return 0;
}

/* tensor.cuh:2162-2178: cusolverDn?potrsBatched(LOWER): L y = b, L^T x = y. */
void orc_dpotrs_lower(int64_t n, int64_t nrhs, const double *L, int64_t lda,
                      double *B, int64_t ldb) {
    for (int64_t r = 0; r < nrhs; ++r) {
        double *b = B + r * ldb;
        for (int64_t i = 0; i < n; ++i) {
            double s = b[i];
            for (int64_t p = 0; p < i; ++p) s -= L[IDX(i, p, lda)] * b[p];
            b[i] = s / L[IDX(i, i, lda)];
        }
        for (int64_t i = n - 1; i >= 0; --i) {
            double s = b[i];
            for (int64_t p = i + 1; p < n; ++p) s -= L[IDX(p, i, lda)] * b[p];
            b[i] = s / L[IDX(i, i, lda)];
        }
    }
}

/* ------------------------------------------------------------------ QR --- */
/* LAPACK dlarfg: given alpha and x, find beta, tau, v with
 * H [alpha; x] = [beta; 0], H = I - tau [1; v][1; v]^T.
 * beta = -sign(alpha) * ||[alpha; x]||, tau = (beta - alpha)/beta,
 * v = x / (alpha - beta).  H = I (tau = 0) when x = 0.  (LAPACK's rescaling
 * loop for |beta| < safmin is not restated; inputs here never reach it.) */
static double householder(int64_t len, double *alpha, double *x) {
    double xnorm2 = 0.0;
    for (int64_t i = 0; i < len; ++i) xnorm2 += x[i] * x[i];
    if (xnorm2 == 0.0) return 0.0;
    double a = *alpha;
    double norm = sqrt(a * a + xnorm2);
    double beta = (a >= 0.0) ? -norm : norm;
    double tau = (beta - a) / beta;
    double scale = 1.0 / (a - beta);
    for (int64_t i = 0; i < len; ++i) x[i] *= scale;
    *alpha = beta;
    return tau;
}

/* Apply H = I - tau v v^T (v[0] = 1 implicit, v[1:] = vtail) to the rows
 * r0..r0+len-1 of C (ncols columns). */
static void apply_reflector(int64_t len, const double *vtail, double tau,
                            double *C, int64_t ldc, int64_t ncols) {
    if (tau == 0.0) return;
    for (int64_t j = 0; j < ncols; ++j) {
        double *c = C + j * ldc;
        double w = c[0];
        for (int64_t i = 1; i < len; ++i) w += vtail[i - 1] * c[i];
        w *= tau;
        c[0] -= w;
        for (int64_t i = 1; i < len; ++i) c[i] -= w * vtail[i - 1];
    }
}

/* tensor.cuh:1866-1876: cusolverDn?geqrf == LAPACK dgeqr2 (unblocked). */
void orc_dgeqrf(int64_t m, int64_t n, double *A, int64_t lda, double *tau) {
    int64_t kmax = (m < n) ? m : n;
    for (int64_t j = 0; j < kmax; ++j) {
        double *ajj = &A[IDX(j, j, lda)];
        tau[j] = householder(m - j - 1, ajj, ajj + 1);
        if (j + 1 < n)
            apply_reflector(m - j, ajj + 1, tau[j], &A[IDX(j, j + 1, lda)], lda, n - j - 1);
    }
}

/* tensor.cuh:1896-1902 (Q^T b) and 1946-1952 (Q I): LAPACK dormqr side 'L'.
 * Q = H_0 H_1 ... H_{k-1}; Q^T applies H_0 first, Q applies H_{k-1} first. */
void orc_dormqr(int trans, int64_t m, int64_t ncols, int64_t k, const double *A,
                int64_t lda, const double *tau, double *C, int64_t ldc) {
    for (int64_t t = 0; t < k; ++t) {
        int64_t j = trans ? t : (k - 1 - t);
        apply_reflector(m - j, &A[IDX(j + 1, j, lda)], tau[j], C + j, ldc, ncols);
    }
}

/* tensor.cuh:1903-1907: cublas?trsm(LEFT, UPPER, N, NON_UNIT): R X = B. */
void orc_dtrsm_lun(int64_t n, int64_t nrhs, const double *R, int64_t ldr,
                   double *B, int64_t ldb) {
    for (int64_t r = 0; r < nrhs; ++r) {
        double *b = B + r * ldb;
        for (int64_t i = n - 1; i >= 0; --i) {
            double s = b[i];
            for (int64_t p = i + 1; p < n; ++p) s -= R[IDX(i, p, ldr)] * b[p];
            b[i] = s / R[IDX(i, i, ldr)];
        }
    }
}

/* tensor.cuh:1340-1366: cublas?gelsBatched(N) for m >= n:
 * A = QR, c = Q^T b, solve R x = c[0:n]; x overwrites b[0:n]. */
void orc_dgels(int64_t m, int64_t n, int64_t nrhs, double *A, int64_t lda,
               double *B, int64_t ldb) {
    double *tau = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_dgeqrf(m, n, A, lda, tau);
    orc_dormqr(1, m, nrhs, n, A, lda, tau, B, ldb);
    orc_dtrsm_lun(n, nrhs, A, lda, B, ldb);
    free(tau);
}

/* ----------------------------------------------------------------- SVD --- */
/* Complete the orthonormal columns U[:, 0:r] (m x r) to an orthonormal basis
 * of R^m: Householder QR of U[:, 0:r] gives Q = H_0..H_{r-1}; its trailing
 * columns Q e_j (j >= r) span the orthogonal complement. */
static void complete_basis(int64_t m, int64_t r, double *U, int64_t ldu) {
    if (r >= m) return;
    double *W = (double *)malloc(sizeof(double) * (size_t)(m * (r > 0 ? r : 1)));
    double *tau = (double *)malloc(sizeof(double) * (size_t)(r > 0 ? r : 1));
    for (int64_t j = 0; j < r; ++j)
        for (int64_t i = 0; i < m; ++i) W[IDX(i, j, m)] = U[IDX(i, j, ldu)];
    if (r > 0) orc_dgeqrf(m, r, W, m, tau);
    for (int64_t j = r; j < m; ++j) {
        double *c = &U[IDX(0, j, ldu)];
        for (int64_t i = 0; i < m; ++i) c[i] = (i == j) ? 1.0 : 0.0;
        if (r > 0) orc_dormqr(0, m, 1, r, W, m, tau, c, ldu);
    }
    free(W);
    free(tau);
}

/* tensor.cuh:1624-1649: cusolverDn?gesvd(jobu 'A' or 'N', jobvt 'A') on a tall
 * matrix.  Restated as one-sided (Hestenes) Jacobi: rotate column pairs of
 * W = A until all pairs are orthogonal to |w_p . w_q| <= tol ||w_p|| ||w_q||,
 * accumulating V; then sigma_j = ||w_j||, sorted descending, u_j = w_j/sigma_j
 * for numerically non-zero sigma_j and the rest of U completes a basis.
 * (Demmel & Veselic 1992; Golub & Van Loan 4th ed. section 8.6.3.) */
int orc_dgesvd(int64_t m, int64_t n, double *A, int64_t lda, double *S,
               double *U, int64_t ldu, double *VT, int64_t ldvt) {
    double *V = (double *)calloc((size_t)(n * n > 0 ? n * n : 1), sizeof(double));
    for (int64_t i = 0; i < n; ++i) V[IDX(i, i, n)] = 1.0;
    const double tol = (double)m * DBL_EPSILON;
    int sweeps = 0;
    for (sweeps = 1; sweeps <= 60; ++sweeps) {
        int rotated = 0;
        for (int64_t p = 0; p < n - 1; ++p) {
            for (int64_t q = p + 1; q < n; ++q) {
                double *wp = &A[IDX(0, p, lda)], *wq = &A[IDX(0, q, lda)];
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                for (int64_t i = 0; i < m; ++i) {
                    alpha += wp[i] * wp[i];
                    beta += wq[i] * wq[i];
                    gamma += wp[i] * wq[i];
                }
                if (gamma == 0.0 || fabs(gamma) <= tol * sqrt(alpha * beta)) continue;
                rotated = 1;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int64_t i = 0; i < m; ++i) {
                    double xp = wp[i], xq = wq[i];
                    wp[i] = c * xp - s * xq;
                    wq[i] = s * xp + c * xq;
                }
                double *vp = &V[IDX(0, p, n)], *vq = &V[IDX(0, q, n)];
                for (int64_t i = 0; i < n; ++i) {
                    double xp = vp[i], xq = vq[i];
                    vp[i] = c * xp - s * xq;
                    vq[i] = s * xp + c * xq;
                }
            }
        }
        if (!rotated) break;
    }
    /* sigma and a descending order (stable selection sort on indices) */
    double *sig = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int64_t *ord = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t j = 0; j < n; ++j) {
        double s2 = 0.0;
        for (int64_t i = 0; i < m; ++i) s2 += A[IDX(i, j, lda)] * A[IDX(i, j, lda)];
        sig[j] = sqrt(s2);
        ord[j] = j;
    }
    for (int64_t a = 0; a < n; ++a) {
        int64_t best = a;
        for (int64_t b = a + 1; b < n; ++b)
            if (sig[ord[b]] > sig[ord[best]]) best = b;
        int64_t tmp = ord[a]; ord[a] = ord[best]; ord[best] = tmp;
    }
    double smax = (n > 0) ? sig[ord[0]] : 0.0;
    double thresh = smax * (double)m * DBL_EPSILON;
    int64_t r = 0;
    for (int64_t jj = 0; jj < n; ++jj) {
        int64_t j = ord[jj];
        S[jj] = sig[j];
        /* sign convention: largest-magnitude entry of v_j is negative for a
         * nonzero sigma_j; vectors of zero sigma keep their accumulated sign
         * (an untouched identity column stays e_j, as in LAPACK) */
        int64_t imax = 0;
        for (int64_t i = 1; i < n; ++i)
            if (fabs(V[IDX(i, j, n)]) > fabs(V[IDX(imax, j, n)])) imax = i;
        double sgn = (sig[j] > thresh && V[IDX(imax, j, n)] > 0.0) ? -1.0 : 1.0;
        for (int64_t i = 0; i < n; ++i) VT[IDX(jj, i, ldvt)] = sgn * V[IDX(i, j, n)];
        if (U && sig[j] > thresh) {
            for (int64_t i = 0; i < m; ++i) U[IDX(i, jj, ldu)] = sgn * A[IDX(i, j, lda)] / sig[j];
            r = jj + 1;
        }
    }
    if (U) complete_basis(m, r, U, ldu);
    free(sig);
    free(ord);
    free(V);
    return sweeps;
}

/* ------------------------------------------------------------- Givens --- */
/* tensor.cuh:1074-1104: cublas?rot(n, x, incx, y, incy, c, s), BLAS drot. */
void orc_drot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s) {
    for (int64_t k = 0; k < n; ++k) {
        const double xv = x[k * incx], yv = y[k * incy];
        x[k * incx] = c * xv + s * yv;
        y[k * incy] = c * yv - s * xv;
    }
}

/* tensor.cuh:2250-2258 (the rhypot kernel) and 2261-2311 (rot of rows i, k
 * with the device-side c and s). */
void orc_dgivens_annihilate(int64_t m, int64_t ncols, double *A, int64_t lda, int64_t i, int64_t k,
                            int64_t j) {
    (void)m;
    const double xij = A[IDX(i, j, lda)], xkj = A[IDX(k, j, lda)];
    const double r = 1.0 / hypot(xij, xkj);
    orc_drot(ncols, A + i, lda, A + k, lda, xij * r, xkj * r);
}

/* tensor.cuh:1484-1491 (kernel) and 1600-1609: count sigma > eps. */
int orc_rank(int64_t n, const double *S, double eps) {
    int r = 0;
    for (int64_t i = 0; i < n; ++i) r += (S[i] > eps) ? 1 : 0;
    return r;
}

/* tensor.cuh:2046-2079: SVD of A^T (n x m), rank with eps, N = U[:, r:n]
 * placed in the first nullity columns of a zeroed n x n N, P = N N^T.
 * The basis of ker A is the orthogonal complement of U[:, 0:r]; any
 * orthonormal basis of it is valid (P does not depend on the choice). */
int orc_nullspace(int64_t m, int64_t n, const double *A, int64_t lda, double eps,
                  double *N, double *P) {
    double *At = (double *)malloc(sizeof(double) * (size_t)(n * m > 0 ? n * m : 1));
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) At[IDX(j, i, n)] = A[IDX(i, j, lda)];
    double *S = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double *U = (double *)calloc((size_t)(n * n), sizeof(double));
    double *VT = (double *)malloc(sizeof(double) * (size_t)(m * m > 0 ? m * m : 1));
    orc_dgesvd(n, m, At, n, S, U, n, VT, m);
    int r = orc_rank(m, S, eps);
    complete_basis(n, r, U, n); /* ker A = complement of U[:, 0:r] */
    int nullity = (int)(n - r);
    memset(N, 0, sizeof(double) * (size_t)(n * n));
    for (int64_t j = 0; j < nullity; ++j)
        for (int64_t i = 0; i < n; ++i) N[IDX(i, j, n)] = U[IDX(i, r + j, n)];
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t p = 0; p < nullity; ++p) s += N[IDX(i, p, n)] * N[IDX(j, p, n)];
            P[IDX(i, j, n)] = s;
        }
    free(At);
    free(S);
    free(U);
    free(VT);
    return nullity;
}

/* ------------------------------------------------------ batched loops --- */
int orc_dposv_batched(int64_t n, int64_t nrhs, double *A, double *B, int64_t batch,
                      int threads) {
    int bad = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4) reduction(+ : bad)
    for (int64_t b = 0; b < batch; ++b) {
        double *a = A + b * n * n;
        if (orc_dpotrf_lower(n, a, n) != 0) {
            bad += 1;
            continue;
        }
        orc_dpotrs_lower(n, nrhs, a, n, B + b * n * nrhs, n);
    }
    return bad;
}

void orc_dgemm_batched(int64_t m, int64_t n, int64_t k, const double *A, const double *B,
                       double *C, int64_t batch, int threads) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t b = 0; b < batch; ++b)
        orc_dgemm(m, n, k, 1.0, A + b * m * k, m, B + b * k * n, k, 0.0, C + b * m * n, m);
}

void orc_sgemm_batched(int64_t m, int64_t n, int64_t k, const float *A, const float *B,
                       float *C, int64_t batch, int threads) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t b = 0; b < batch; ++b) {
        const float *a = A + b * m * k, *bb = B + b * k * n;
        float *c = C + b * m * n;
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) {
                float acc = 0.0f;
                for (int64_t p = 0; p < k; ++p) acc += a[i + p * m] * bb[p + j * k];
                c[i + j * m] = acc;
            }
    }
}

void orc_dgels_ragged(int64_t count, const int32_t *m, const int32_t *n, const int64_t *offA,
                      const int64_t *offB, double *A, double *B, int threads) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t s = 0; s < count; ++s)
        orc_dgels(m[s], n[s], 1, A + offA[s], m[s], B + offB[s], m[s]);
}

/* Nullspace of every m x n matrix of a batch (OpenMP over matrices). */
int orc_nullspace_batched(int64_t m, int64_t n, const double *A, double eps, double *N, double *P,
                          int32_t *nullity, int64_t batch, int threads) {
    int bad = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : bad)
    for (int64_t b = 0; b < batch; ++b) {
        nullity[b] = orc_nullspace(m, n, A + b * m * n, m, eps, N + b * n * n, P + b * n * n);
        bad += (nullity[b] < 0) ? 1 : 0;
    }
    return bad;
}
