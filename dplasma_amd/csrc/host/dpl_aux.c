/*
 * dpl_aux.c -- generators, copies, residual checks and data placement.
 *
 *   dplasma_dplgsy   src/zplgsy_wrapper.c:184-206 (check of uplo -> -3);
 *                    entries generated on the GPU (see mi_dplgsy_tiles)
 *   dplasma_dplrnt   src/zplrnt_wrapper.c:192-206, diagdom adds max(M,N) to
 *                    the diagonal (:80-93)
 *   dplasma_dlacpy   src/zlacpy_wrapper.c (uplo part of A into B)
 *   check_dpotrf     src/dplasma_zcheck.c:63-133
 *   check_daxmb      src/dplasma_zcheck.c:179-224
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "dpl_internal.h"

int dpl_run_plgsy(parsec_context_t *ctx, double bump, int uplo, parsec_tiled_matrix_t *A,
                  unsigned long long seed)
{
    const char *f = "dplasma_dplgsy";
    if (A->m != A->n || A->mb != A->nb) { dpl_error(f, "square matrix and tiles required"); return -2; }
    int sym_uplo = dpl_is_sym_desc(A) ? parsec_shim_uplo(A) : dplasmaLower;
    dpl_sym_dev_t s;
    int rc = dpl_sym_acquire(ctx, A, sym_uplo, 0, &s, NULL);
    if (rc) return rc;
    /* the lower image of a symmetric matrix is also the image of its upper
     * triangle (transposed tiles), so one generator serves both */
    if ((rc = mi_dplgsy_tiles(ctx->mi, bump, s.N, s.NB, s.d, seed))) {
        dpl_sym_release(ctx, A, &s, 0, NULL);
        return dpl_mi_fail(f, rc);
    }
    if (dpl_is_sym_desc(A) || uplo == dplasmaLower) return dpl_sym_release(ctx, A, &s, 1, NULL);
    if (uplo == dplasmaUpper) {
        /* general descriptor, upper tiles only */
        rc = dpl_sym_release(ctx, A, &s, 0, NULL);
        if (rc) return rc;
        if ((rc = dpl_sym_acquire(ctx, A, dplasmaUpper, 0, &s, NULL))) return rc;
        if ((rc = mi_dplgsy_tiles(ctx->mi, bump, s.N, s.NB, s.d, seed))) {
            dpl_sym_release(ctx, A, &s, 0, NULL);
            return dpl_mi_fail(f, rc);
        }
        return dpl_sym_release(ctx, A, &s, 1, NULL);
    }
    /* UpperLower on a general descriptor: both triangles */
    rc = dpl_sym_release(ctx, A, &s, 1, NULL);
    if (rc) return rc;
    if ((rc = dpl_sym_acquire(ctx, A, dplasmaUpper, 0, &s, NULL))) return rc;
    if ((rc = mi_dplgsy_tiles(ctx->mi, bump, s.N, s.NB, s.d, seed))) {
        dpl_sym_release(ctx, A, &s, 0, NULL);
        return dpl_mi_fail(f, rc);
    }
    return dpl_sym_release(ctx, A, &s, 1, NULL);
}

parsec_taskpool_t *dplasma_dplgsy_New(double bump, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                                      unsigned long long int seed)
{
    parsec_taskpool_t *tp = (parsec_taskpool_t *)calloc(1, sizeof(parsec_taskpool_t));
    tp->kind = DPL_TP_PLGSY;
    tp->bump = bump;
    tp->uplo = uplo;
    tp->A = A;
    tp->seed = seed;
    return tp;
}

void dplasma_dplgsy_Destruct(parsec_taskpool_t *tp) { free(tp); }

int dplasma_dplgsy(parsec_context_t *parsec, double bump, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                   unsigned long long int seed)
{
    if (uplo != dplasmaLower && uplo != dplasmaUpper && uplo != dplasmaUpperLower) {
        dpl_error("dplasma_dplgsy", "illegal value of type");
        return -3;
    }
    parsec_taskpool_t *tp = dplasma_dplgsy_New(bump, uplo, A, seed);
    parsec_context_add_taskpool(parsec, tp);
    int rc = parsec_context_wait(parsec);
    dplasma_dplgsy_Destruct(tp);
    return rc;
}

/* real precision: the Hermitian generator is the symmetric one */
int dplasma_dplghe(parsec_context_t *parsec, double bump, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                   unsigned long long int seed)
{
    return dplasma_dplgsy(parsec, bump, uplo, A, seed);
}

int dplasma_dplrnt(parsec_context_t *parsec, int diagdom, parsec_tiled_matrix_t *A,
                   unsigned long long int seed)
{
    const char *f = "dplasma_dplrnt";
    dpl_ge_dev_t g;
    int rc = dpl_ge_acquire(parsec, A, 0, &g, NULL);
    if (rc) return rc;
    if ((rc = mi_dplrnt(parsec->mi, g.M, g.N, g.d, g.ld, g.M, 0, 0, seed))) {
        dpl_ge_release(parsec, A, &g, 0, NULL);
        return dpl_mi_fail(f, rc);
    }
    if (diagdom) {
        int maxmn = A->m > A->n ? A->m : A->n;
        if ((rc = mi_dadd_diag(parsec->mi, A->m < A->n ? A->m : A->n, g.d, g.ld, (double)maxmn))) {
            dpl_ge_release(parsec, A, &g, 0, NULL);
            return dpl_mi_fail(f, rc);
        }
    }
    return dpl_ge_release(parsec, A, &g, 1, NULL);
}

/* host-side tile copy of the uplo part (both descriptors in host memory) */
int dplasma_dlacpy(parsec_context_t *parsec, dplasma_enum_t uplo, const parsec_tiled_matrix_t *A,
                   parsec_tiled_matrix_t *B)
{
    if (A->mt != B->mt || A->nt != B->nt || A->mb != B->mb || A->nb != B->nb) {
        dpl_error("dplasma_dlacpy", "descriptors must share the tiling");
        return -2;
    }
    dpl_sync_host(parsec, A);
    for (int n = 0; n < A->nt; n++)
        for (int m = 0; m < A->mt; m++) {
            if (uplo == dplasmaLower && m < n) continue;
            if (uplo == dplasmaUpper && m > n) continue;
            const double *a = (const double *)parsec_shim_tile_addr(A, m, n);
            double *b = (double *)parsec_shim_tile_addr(B, m, n);
            if (!a || !b) continue;
            int tm = (m == A->mt - 1) ? A->m - m * A->mb : A->mb;
            int tn = (n == A->nt - 1) ? A->n - n * A->nb : A->nb;
            int lda = (A->storage == PARSEC_MATRIX_LAPACK && !dpl_is_sym_desc(A)) ? A->llm : A->mb;
            int ldb = (B->storage == PARSEC_MATRIX_LAPACK && !dpl_is_sym_desc(B)) ? B->llm : B->mb;
            for (int j = 0; j < tn; j++) {
                int i0 = 0, i1 = tm;
                if (m == n && uplo == dplasmaLower) i0 = j;
                if (m == n && uplo == dplasmaUpper) i1 = (j + 1 < tm) ? j + 1 : tm;
                if (i1 > i0) memcpy(b + (int64_t)j * ldb + i0, a + (int64_t)j * lda + i0, sizeof(double) * (i1 - i0));
            }
        }
    dpl_release(parsec, B);   /* host copy of B changed: drop a stale image */
    return 0;
}

/* ---------------------------------------------------------------- checks */
static double eps_d(void) { return DBL_EPSILON * 0.5; }   /* LAPACK dlamch('e') */

int check_dpotrf(parsec_context_t *parsec, int loud, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                 parsec_tiled_matrix_t *A0)
{
    dpl_sym_dev_t sl, sa;
    int N = A->n;
    if (dpl_sym_acquire(parsec, A, uplo, 1, &sl, NULL)) return 1;
    if (dpl_sym_acquire(parsec, A0, uplo, 1, &sa, NULL)) { dpl_sym_release(parsec, A, &sl, 0, NULL); return 1; }
    double Anorm = 0.0, Rnorm = 0.0;
    memset(&parsec->last_check, 0, sizeof parsec->last_check);
    int rc = mi_dcheck_potrf(parsec->mi, sl.N, sl.NB, sl.d, sa.d, &Anorm, &Rnorm);
    dpl_sym_release(parsec, A0, &sa, 0, NULL);
    dpl_sym_release(parsec, A, &sl, 0, NULL);
    if (rc) { dpl_mi_fail("check_dpotrf", rc); return 1; }
    double result = Rnorm / (Anorm * N * eps_d());
    parsec->last_check.Anorm = Anorm;
    parsec->last_check.Rnorm = Rnorm;
    parsec->last_check.result = result;
    if (loud > 2) {
        printf("============\n");
        printf("Checking the Cholesky factorization \n");
        if (loud > 3) printf("-- ||A||_oo = %e, ||L'L-A||_oo = %e\n", Anorm, Rnorm);
        printf("-- ||L'L-A||_oo/(||A||_oo.N.eps) = %e \n", result);
    }
    int bad = isnan(Rnorm) || isinf(Rnorm) || isnan(result) || isinf(result) || (result > 60.0);
    if (loud) printf(bad ? "-- Factorization is suspicious ! \n" : "-- Factorization is CORRECT ! \n");
    return bad ? 1 : 0;
}

/* check_dpoinv (src/dplasma_zcheck.c:264-314): ||I - Ainv A||_1 /
 * (||A||_1 ||Ainv||_1 N eps) < 10, computed on the GPU.  A and Ainv are
 * symmetric descriptors with their uplo part stored (the reference takes a
 * fully generated A). */
int check_dpoinv(parsec_context_t *parsec, int loud, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                 parsec_tiled_matrix_t *Ainv)
{
    dpl_sym_dev_t sa, si;
    int N = A->n;
    if (dpl_sym_acquire(parsec, A, uplo, 1, &sa, NULL)) return 1;
    if (dpl_sym_acquire(parsec, Ainv, uplo, 1, &si, NULL)) { dpl_sym_release(parsec, A, &sa, 0, NULL); return 1; }
    double Anorm = 0.0, Ainvnorm = 0.0, Rnorm = 0.0;
    memset(&parsec->last_check, 0, sizeof parsec->last_check);
    int rc = mi_dcheck_poinv(parsec->mi, sa.N, sa.NB, sa.d, si.d, &Anorm, &Ainvnorm, &Rnorm);
    dpl_sym_release(parsec, Ainv, &si, 0, NULL);
    dpl_sym_release(parsec, A, &sa, 0, NULL);
    if (rc) { dpl_mi_fail("check_dpoinv", rc); return 1; }
    double result = Rnorm / ((Anorm * Ainvnorm) * N * eps_d());
    parsec->last_check.Anorm = Anorm;
    parsec->last_check.Ainvnorm = Ainvnorm;
    parsec->last_check.Rnorm = Rnorm;
    parsec->last_check.result = result;
    if (loud > 2)
        printf("  ||A||_one = %e, ||A^(-1)||_one = %e, ||I - A * A^(-1)||_one = %e, result = %e\n", Anorm,
               Ainvnorm, Rnorm, result);
    return (isinf(Ainvnorm) || isnan(result) || isinf(result) || result > 10.0) ? 1 : 0;
}

int check_daxmb(parsec_context_t *parsec, int loud, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                parsec_tiled_matrix_t *b, parsec_tiled_matrix_t *x)
{
    dpl_sym_dev_t sa;
    dpl_ge_dev_t gb, gx;
    int N = b->m;
    if (dpl_sym_acquire(parsec, A, uplo, 1, &sa, NULL)) return 1;
    if (dpl_ge_acquire(parsec, b, 1, &gb, NULL)) { dpl_sym_release(parsec, A, &sa, 0, NULL); return 1; }
    if (dpl_ge_acquire(parsec, x, 1, &gx, NULL)) {
        dpl_ge_release(parsec, b, &gb, 0, NULL);
        dpl_sym_release(parsec, A, &sa, 0, NULL);
        return 1;
    }
    double An = 0, Bn = 0, Xn = 0, Rn = 0;
    memset(&parsec->last_check, 0, sizeof parsec->last_check);
    int rc = mi_dcheck_axmb(parsec->mi, sa.N, sa.NB, sa.d, gb.N, gb.d, gb.ld, gx.d, gx.ld, &An, &Bn, &Xn, &Rn);
    dpl_ge_release(parsec, x, &gx, 0, NULL);
    dpl_ge_release(parsec, b, &gb, 0, NULL);
    dpl_sym_release(parsec, A, &sa, 0, NULL);
    if (rc) { dpl_mi_fail("check_daxmb", rc); return 1; }
    double result = Rn / ((An * Xn + Bn) * N * eps_d());
    parsec->last_check.Anorm = An;
    parsec->last_check.Bnorm = Bn;
    parsec->last_check.Xnorm = Xn;
    parsec->last_check.Rnorm = Rn;
    parsec->last_check.result = result;
    if (loud > 2) {
        printf("============\n");
        printf("Checking the Residual of the solution \n");
        if (loud > 3)
            printf("-- ||A||_oo = %e, ||X||_oo = %e, ||B||_oo= %e, ||A X - B||_oo = %e\n", An, Xn, Bn, Rn);
        printf("-- ||Ax-B||_oo/((||A||_oo||x||_oo+||B||_oo).N.eps) = %e \n", result);
    }
    int bad = isnan(Xn) || isinf(Xn) || isnan(result) || isinf(result) || (result > 60.0);
    if (loud) printf(bad ? "-- Solution is suspicious ! \n" : "-- Solution is CORRECT ! \n");
    return bad ? 1 : 0;
}

/* ----------------------------------------------------------- placement */
int dplasma_advise_data_on_device(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                                  parsec_tiled_matrix_unary_op_t operation, void *op_args)
{
    (void)operation;
    (void)op_args;
    return dpl_make_resident(parsec, A, uplo);
}

int dplasma_data_sync_host(parsec_context_t *parsec, parsec_tiled_matrix_t *A)
{
    return dpl_sync_host(parsec, A);
}

int dplasma_data_release(parsec_context_t *parsec, parsec_tiled_matrix_t *A)
{
    return dpl_release(parsec, A);
}

int dplasma_mi_last_check(parsec_context_t *parsec, dplasma_mi_check_t *c)
{
    if (!parsec || !c) return -1;
    *c = parsec->last_check;
    return 0;
}

int dplasma_mi_last_timing(parsec_context_t *parsec, dplasma_mi_timing_t *t)
{
    *t = parsec->last;
    return 0;
}

struct mi_ctx_s *dplasma_mi_context(parsec_context_t *parsec) { return parsec->mi; }
