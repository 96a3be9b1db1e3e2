/*
 * check.c -- DPLASMA's residual checks restated on full LAPACK-layout arrays
 * (test infrastructure).
 *
 *   orc_check_dpotrf : src/dplasma_zcheck.c:63-133.  LLt = uplo part of A
 *       (dlaset + dlacpy :88-89), LLt = L L^T or U^T U (dtrmm :92-94),
 *       LLt -= A0 on the uplo triangle (dtradd :97-98), Inf norms of the
 *       symmetric matrices defined by that triangle (dlansy :100-102),
 *       result = Rnorm / (Anorm N eps) (:104), eps = dlamch('e').
 *   orc_check_daxmb  : src/dplasma_zcheck.c:179-224.  Anorm = dlansy(Inf),
 *       Bnorm, Xnorm = dlange(Inf) (:193-195), r = b - A0 x (dsymm :198),
 *       result = ||r|| / ((Anorm Xnorm + Bnorm) N eps) (:202).
 * Callers compare the result with the reference's threshold, 60 (:116-122,
 * :214).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* LAPACK dlamch('E'): relative machine epsilon for rounding arithmetic,
 * 2^-53 in IEEE double. */
double orc_dlamch_eps(void) { return DBL_EPSILON * 0.5; }

static inline double symv_at(int uplo, const double *A, int lda, int i, int j)
{
    if (uplo == ORC_LOWER) return (i >= j) ? A[(int64_t)j * lda + i] : A[(int64_t)i * lda + j];
    return (i <= j) ? A[(int64_t)j * lda + i] : A[(int64_t)i * lda + j];
}

double orc_dlansy_inf(int uplo, int N, const double *A, int lda)
{
    double *rs = (double *)calloc((size_t)N, sizeof(double));
    /* walk the stored triangle once, adding each entry to both row sums */
    for (int j = 0; j < N; j++) {
        const double *c = A + (int64_t)j * lda;
        if (uplo == ORC_LOWER) {
            rs[j] += fabs(c[j]);
            for (int i = j + 1; i < N; i++) { double v = fabs(c[i]); rs[i] += v; rs[j] += v; }
        } else {
            for (int i = 0; i < j; i++) { double v = fabs(c[i]); rs[i] += v; rs[j] += v; }
            rs[j] += fabs(c[j]);
        }
    }
    double mx = 0.0;
    for (int i = 0; i < N; i++) if (rs[i] > mx || isnan(rs[i])) mx = rs[i];
    free(rs);
    return mx;
}

double orc_dlange_inf(int M, int N, const double *A, int lda)
{
    double *rs = (double *)calloc((size_t)(M > 0 ? M : 1), sizeof(double));
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) rs[i] += fabs(A[(int64_t)j * lda + i]);
    double mx = 0.0;
    for (int i = 0; i < M; i++) if (rs[i] > mx || isnan(rs[i])) mx = rs[i];
    free(rs);
    return mx;
}

double orc_check_dpotrf(int uplo, int N, const double *A, int lda,
                        const double *A0, int lda0, double *Anorm_out, double *Rnorm_out)
{
    /* R holds the uplo triangle of L L^T - A0 (or U^T U - A0). */
    double *R = (double *)calloc((size_t)N * N, sizeof(double));
    if (uplo == ORC_LOWER) {
        /* R(i,j) = sum_{p<=j} L(i,p) L(j,p), i >= j */
        for (int j = 0; j < N; j++) {
            double *r = R + (int64_t)j * N;
            for (int p = 0; p <= j; p++) {
                const double *lp = A + (int64_t)p * lda;
                double ljp = lp[j];
                for (int i = j; i < N; i++) r[i] += lp[i] * ljp;
            }
            const double *a0 = A0 + (int64_t)j * lda0;
            for (int i = j; i < N; i++) r[i] -= a0[i];
        }
    } else {
        /* R(i,j) = sum_{p<=i} U(p,i) U(p,j), i <= j */
        for (int j = 0; j < N; j++) {
            double *r = R + (int64_t)j * N;
            const double *uj = A + (int64_t)j * lda;
            for (int i = 0; i <= j; i++) {
                const double *ui = A + (int64_t)i * lda;
                double s = 0.0;
                for (int p = 0; p <= i; p++) s += ui[p] * uj[p];
                r[i] = s - A0[(int64_t)j * lda0 + i];
            }
        }
    }
    double Anorm = orc_dlansy_inf(uplo, N, A0, lda0);
    double Rnorm = orc_dlansy_inf(uplo, N, R, N);
    free(R);
    if (Anorm_out) *Anorm_out = Anorm;
    if (Rnorm_out) *Rnorm_out = Rnorm;
    return Rnorm / (Anorm * N * orc_dlamch_eps());
}

double orc_check_daxmb(int uplo, int N, int NRHS, const double *A0, int lda0,
                       const double *b, int ldb, const double *x, int ldx)
{
    double Anorm = orc_dlansy_inf(uplo, N, A0, lda0);
    double Bnorm = orc_dlange_inf(N, NRHS, b, ldb);
    double Xnorm = orc_dlange_inf(N, NRHS, x, ldx);
    double *r = (double *)malloc(sizeof(double) * (size_t)N * NRHS);
    for (int c = 0; c < NRHS; c++) {
        double *rc = r + (int64_t)c * N;
        const double *xc = x + (int64_t)c * ldx;
        memcpy(rc, b + (int64_t)c * ldb, sizeof(double) * N);
        /* r -= A0 x using the stored triangle: column j contributes A0(:,j) x_j */
        for (int j = 0; j < N; j++) {
            const double *a = A0 + (int64_t)j * lda0;
            double xj = xc[j];
            if (uplo == ORC_LOWER) {
                double s = 0.0;
                rc[j] -= a[j] * xj;
                for (int i = j + 1; i < N; i++) { rc[i] -= a[i] * xj; s += a[i] * xc[i]; }
                rc[j] -= s;
            } else {
                double s = 0.0;
                for (int i = 0; i < j; i++) { rc[i] -= a[i] * xj; s += a[i] * xc[i]; }
                rc[j] -= s + a[j] * xj;
            }
        }
    }
    double Rnorm = orc_dlange_inf(N, NRHS, r, N);
    free(r);
    return Rnorm / ((Anorm * Xnorm + Bnorm) * N * orc_dlamch_eps());
}

/* check_zpoinv (src/dplasma_zcheck.c:264-314): ||I - Ainv A||_1 /
 * (||A||_1 ||Ainv||_1 N eps), A and Ainv symmetric with their uplo
 * triangle referenced (the reference forms Ainv A with zhemm(Left, uplo),
 * then zlanhe / zlange one-norms).  For a symmetric matrix the one-norm
 * equals the infinity norm. */
double orc_check_dpoinv(int uplo, int N, const double *A, int lda, const double *Ainv, int ldi,
                        double *Anorm_out, double *Ainvnorm_out, double *Rnorm_out)
{
    double *Af = (double *)malloc(sizeof(double) * (size_t)N * N);
    double *If = (double *)malloc(sizeof(double) * (size_t)N * N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            const int lo = (uplo == ORC_LOWER) ? (i >= j) : (i <= j);
            Af[(int64_t)j * N + i] = lo ? A[(int64_t)j * lda + i] : A[(int64_t)i * lda + j];
            If[(int64_t)j * N + i] = lo ? Ainv[(int64_t)j * ldi + i] : Ainv[(int64_t)i * ldi + j];
        }
    double Rnorm = 0.0;
    double *col = (double *)malloc(sizeof(double) * (size_t)N);
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) col[i] = (i == j) ? 1.0 : 0.0;
        for (int p = 0; p < N; p++) {
            const double a = Af[(int64_t)j * N + p];
            const double *ip = If + (int64_t)p * N;
            for (int i = 0; i < N; i++) col[i] -= ip[i] * a;
        }
        double cs = 0.0;
        for (int i = 0; i < N; i++) cs += fabs(col[i]);
        if (cs > Rnorm || cs != cs) Rnorm = cs;
    }
    double Anorm = orc_dlansy_inf(uplo, N, A, lda);
    double Ainvnorm = orc_dlansy_inf(uplo, N, Ainv, ldi);
    free(Af); free(If); free(col);
    if (Anorm_out) *Anorm_out = Anorm;
    if (Ainvnorm_out) *Ainvnorm_out = Ainvnorm;
    if (Rnorm_out) *Rnorm_out = Rnorm;
    return Rnorm / ((Anorm * Ainvnorm) * N * orc_dlamch_eps());
}
