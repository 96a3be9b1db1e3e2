/*
 * tile_kernels.c -- FP64 tile kernels of the oracle (test infrastructure).
 *
 * Plain-C restatements of the BLAS/LAPACK semantics that DPLASMA's CPU task
 * bodies call:
 *   CORE_dpotrf -> LAPACKE_dpotrf_work        src/cores/core_zpotrf.c:68-74
 *   CORE_dtrsm  -> cblas_dtrsm                src/cores/core_ztrsm.c:80-93
 *   CORE_dsyrk  -> cblas_dsyrk (Lower, NoTrans) src/cores/core_zsyrk.c:80-91
 *   CORE_dgemm  -> cblas_dgemm                src/cores/core_zgemm.c:90-103
 * plus an optional path through a dlopen'ed CBLAS/LAPACKE (the scipy-bundled
 * OpenBLAS here) for the CPU baseline timing.  The enum values are the CBLAS
 * ones, which DPLASMA's constants.h shares (111/112, 121/122, 131/132, 141/142).
 */
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "oracle.h"

/* ---------------------------------------------------------------- POTRF */
/* Unblocked Cholesky (dot-product form of LAPACK dpotf2).  Returns 0, or the
 * 1-based column j whose pivot is not positive (or NaN); columns < j hold the
 * factor and the rest is untouched, as LAPACK leaves it. */
int orc_dpotrf(int uplo, int n, double *A, int lda)
{
    if (uplo == ORC_LOWER) {
        for (int j = 0; j < n; j++) {
            double *colj = A + (int64_t)j * lda;
            double ajj = colj[j];
            for (int p = 0; p < j; p++) {
                double l = A[(int64_t)p * lda + j];
                ajj -= l * l;
            }
            if (!(ajj > 0.0)) { colj[j] = ajj; return j + 1; }
            ajj = sqrt(ajj);
            colj[j] = ajj;
            /* A(j+1:n, j) = (A(j+1:n, j) - A(j+1:n, 0:j) A(j, 0:j)^T) / ajj */
            for (int p = 0; p < j; p++) {
                double l = A[(int64_t)p * lda + j];
                const double *colp = A + (int64_t)p * lda;
                for (int i = j + 1; i < n; i++) colj[i] -= colp[i] * l;
            }
            double r = 1.0 / ajj;
            for (int i = j + 1; i < n; i++) colj[i] *= r;
        }
    } else {
        for (int j = 0; j < n; j++) {
            double *colj = A + (int64_t)j * lda;
            double ajj = colj[j];
            for (int p = 0; p < j; p++) ajj -= colj[p] * colj[p];
            if (!(ajj > 0.0)) { colj[j] = ajj; return j + 1; }
            ajj = sqrt(ajj);
            colj[j] = ajj;
            double r = 1.0 / ajj;
            for (int i = j + 1; i < n; i++) {
                double *coli = A + (int64_t)i * lda;
                double s = coli[j];
                for (int p = 0; p < j; p++) s -= colj[p] * coli[p];
                coli[j] = s * r;
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- GEMM */
void orc_dgemm(int transa, int transb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc)
{
    for (int j = 0; j < n; j++) {
        double *c = C + (int64_t)j * ldc;
        if (beta == 0.0) for (int i = 0; i < m; i++) c[i] = 0.0;
        else if (beta != 1.0) for (int i = 0; i < m; i++) c[i] *= beta;
        for (int p = 0; p < k; p++) {
            double b = (transb == ORC_NOTRANS) ? B[(int64_t)j * ldb + p] : B[(int64_t)p * ldb + j];
            double t = alpha * b;
            if (transa == ORC_NOTRANS) {
                const double *a = A + (int64_t)p * lda;
                for (int i = 0; i < m; i++) c[i] += t * a[i];
            } else {
                for (int i = 0; i < m; i++) c[i] += t * A[(int64_t)i * lda + p];
            }
        }
    }
}

/* ---------------------------------------------------------------- SYRK */
/* C (lower) = alpha A A^T + beta C, A n x k. */
void orc_dsyrk_ln(int n, int k, double alpha, const double *A, int lda,
                  double beta, double *C, int ldc)
{
    for (int j = 0; j < n; j++) {
        double *c = C + (int64_t)j * ldc;
        if (beta == 0.0) for (int i = j; i < n; i++) c[i] = 0.0;
        else if (beta != 1.0) for (int i = j; i < n; i++) c[i] *= beta;
        for (int p = 0; p < k; p++) {
            const double *a = A + (int64_t)p * lda;
            double t = alpha * a[j];
            for (int i = j; i < n; i++) c[i] += t * a[i];
        }
    }
}

/* ---------------------------------------------------------------- TRSM */
static inline double tri(const double *A, int lda, int i, int j) { return A[(int64_t)j * lda + i]; }

/* op(A) X = alpha B (Left) or X op(A) = alpha B (Right); A triangular. */
void orc_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb)
{
    int unit = (diag == ORC_UNIT);
    /* effective lower-ness of op(A) */
    int lower = (uplo == ORC_LOWER) ^ (trans != ORC_NOTRANS);
#define OPA(i, j) ((trans == ORC_NOTRANS) ? tri(A, lda, (i), (j)) : tri(A, lda, (j), (i)))
    if (side == ORC_LEFT) {
        for (int j = 0; j < n; j++) {
            double *b = B + (int64_t)j * ldb;
            if (alpha != 1.0) for (int i = 0; i < m; i++) b[i] *= alpha;
            if (lower) {
                for (int i = 0; i < m; i++) {
                    double s = b[i];
                    for (int p = 0; p < i; p++) s -= OPA(i, p) * b[p];
                    b[i] = unit ? s : s / OPA(i, i);
                }
            } else {
                for (int i = m - 1; i >= 0; i--) {
                    double s = b[i];
                    for (int p = i + 1; p < m; p++) s -= OPA(i, p) * b[p];
                    b[i] = unit ? s : s / OPA(i, i);
                }
            }
        }
    } else {
        /* X op(A) = alpha B: column j of X from columns of B */
        if (alpha != 1.0)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < m; i++) B[(int64_t)j * ldb + i] *= alpha;
        if (lower) {
            /* op(A) lower: X(:,j) = (B(:,j) - sum_{p>j} X(:,p) op(A)(p,j)) / op(A)(j,j) */
            for (int j = n - 1; j >= 0; j--) {
                double *bj = B + (int64_t)j * ldb;
                for (int p = j + 1; p < n; p++) {
                    double a = OPA(p, j);
                    const double *bp = B + (int64_t)p * ldb;
                    for (int i = 0; i < m; i++) bj[i] -= bp[i] * a;
                }
                if (!unit) {
                    double r = 1.0 / OPA(j, j);
                    for (int i = 0; i < m; i++) bj[i] *= r;
                }
            }
        } else {
            for (int j = 0; j < n; j++) {
                double *bj = B + (int64_t)j * ldb;
                for (int p = 0; p < j; p++) {
                    double a = OPA(p, j);
                    const double *bp = B + (int64_t)p * ldb;
                    for (int i = 0; i < m; i++) bj[i] -= bp[i] * a;
                }
                if (!unit) {
                    double r = 1.0 / OPA(j, j);
                    for (int i = 0; i < m; i++) bj[i] *= r;
                }
            }
        }
    }
#undef OPA
}

/* ------------------------------------------------- optional loaded BLAS */
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double, const double *, int,
                              const double *, int, double, double *, int);
typedef void (*cblas_dsyrk_t)(int, int, int, int, int, double, const double *, int,
                              double, double *, int);
typedef void (*cblas_dtrsm_t)(int, int, int, int, int, int, int, double, const double *,
                              int, double *, int);
typedef int (*lapacke_dpotrf_work_t)(int, char, int, double *, int);
typedef void (*set_threads_t)(int);

static cblas_dgemm_t p_dgemm;
static cblas_dsyrk_t p_dsyrk;
static cblas_dtrsm_t p_dtrsm;
static lapacke_dpotrf_work_t p_dpotrf;
static set_threads_t p_set_threads;

int orc_blas_load(const char *path, const char *prefix)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    char name[128];
    if (!h) return -1;
#define SYM(var, type, base) do { snprintf(name, sizeof name, "%s%s", prefix, base); \
        var = (type)dlsym(h, name); } while (0)
    SYM(p_dgemm, cblas_dgemm_t, "cblas_dgemm");
    SYM(p_dsyrk, cblas_dsyrk_t, "cblas_dsyrk");
    SYM(p_dtrsm, cblas_dtrsm_t, "cblas_dtrsm");
    SYM(p_dpotrf, lapacke_dpotrf_work_t, "LAPACKE_dpotrf_work");
    SYM(p_set_threads, set_threads_t, "openblas_set_num_threads");
#undef SYM
    if (!p_dgemm || !p_dsyrk || !p_dtrsm || !p_dpotrf) {
        p_dgemm = 0; p_dsyrk = 0; p_dtrsm = 0; p_dpotrf = 0;
        return -2;
    }
    return 0;
}

int orc_blas_loaded(void) { return p_dgemm != 0; }

void orc_blas_set_threads(int n) { if (p_set_threads) p_set_threads(n); }

/* Dispatchers used by the DAG executor: same argument lists as CORE_d*. */
int orc_k_potrf(int use_blas, int n, double *A, int lda)
{
    if (use_blas && p_dpotrf) return p_dpotrf(102 /*col major*/, 'L', n, A, lda);
    return orc_dpotrf(ORC_LOWER, n, A, lda);
}

void orc_k_trsm_rltn(int use_blas, int m, int n, const double *T, int ldt, double *C, int ldc)
{
    if (use_blas && p_dtrsm)
        p_dtrsm(102, ORC_RIGHT, ORC_LOWER, ORC_TRANS, ORC_NONUNIT, m, n, 1.0, T, ldt, C, ldc);
    else
        orc_dtrsm(ORC_RIGHT, ORC_LOWER, ORC_TRANS, ORC_NONUNIT, m, n, 1.0, T, ldt, C, ldc);
}

void orc_k_syrk_ln(int use_blas, int n, int k, const double *A, int lda, double *C, int ldc)
{
    if (use_blas && p_dsyrk)
        p_dsyrk(102, ORC_LOWER, ORC_NOTRANS, n, k, -1.0, A, lda, 1.0, C, ldc);
    else
        orc_dsyrk_ln(n, k, -1.0, A, lda, 1.0, C, ldc);
}

void orc_k_gemm_nt(int use_blas, int m, int n, int k, const double *A, int lda,
                   const double *B, int ldb, double *C, int ldc)
{
    if (use_blas && p_dgemm)
        p_dgemm(102, ORC_NOTRANS, ORC_TRANS, m, n, k, -1.0, A, lda, B, ldb, 1.0, C, ldc);
    else
        orc_dgemm(ORC_NOTRANS, ORC_TRANS, m, n, k, -1.0, A, lda, B, ldb, 1.0, C, ldc);
}

void orc_k_gemm(int use_blas, int ta, int tb, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb, double beta,
                double *C, int ldc)
{
    if (use_blas && p_dgemm)
        p_dgemm(102, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        orc_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void orc_k_trsm(int use_blas, int side, int uplo, int trans, int diag, int m, int n,
                double alpha, const double *A, int lda, double *B, int ldb)
{
    if (use_blas && p_dtrsm)
        p_dtrsm(102, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
    else
        orc_dtrsm(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}
