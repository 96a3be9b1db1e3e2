/*
 * dpl_potrf.c -- dplasma_dpotrf / dpotrs / dposv / dtrsm on the GPU.
 *
 * Argument checks and return conventions follow the reference wrappers:
 *   dplasma_zpotrf      src/zpotrf_wrapper.c:345-371 (New :193-258)
 *   dplasma_zpotrs      src/zpotrs_wrapper.c:56-97
 *   dplasma_zposv       src/zposv_wrapper.c:63-113
 *   dplasma_ztrsm       src/ztrsm_wrapper.c:262-309 (New :93-162)
 * The operand images live in HBM for the whole call (or across calls when
 * advised); the work itself is the kernel library's mi_dpotrf_tiles /
 * mi_dtrsm_LL / mi_dpotrs_tiles.
 */
#include <stdlib.h>
#include <string.h>
#include "dpl_internal.h"

/* ----------------------------------------------------------- executors */
int dpl_run_potrf(parsec_context_t *ctx, int uplo, parsec_tiled_matrix_t *A, int hmb, int *info)
{
    const char *f = "dplasma_dpotrf";
    dpl_sym_dev_t s;
    memset(&ctx->last, 0, sizeof ctx->last);
    int rc = dpl_sym_acquire(ctx, A, uplo, 1, &s, &ctx->last.stage_in_ms);
    if (rc) return rc;
    int iinfo = 0;
    rc = hmb > 0 ? mi_dpotrf_tiles_rec(ctx->mi, s.N, s.NB, s.d, hmb, &iinfo)
                 : mi_dpotrf_tiles(ctx->mi, s.N, s.NB, s.d, &iinfo);
    if (rc) {
        dpl_sym_release(ctx, A, &s, 0, NULL);
        return dpl_mi_fail(f, rc);
    }
    mi_last_span_ms(ctx->mi, &ctx->last.total_ms);
    if (info && *info == 0) *info = iinfo;
    return dpl_sym_release(ctx, A, &s, 1, &ctx->last.stage_out_ms);
}

/* map (uplo, trans) of the stored triangle onto the lower image:
 * Upper holds U = L^T, so U X = B is L^T X = B and U^T X = B is L X = B */
static int lower_trans(int uplo, int trans)
{
    int t = (trans == dplasmaNoTrans) ? dplasmaNoTrans : dplasmaTrans;
    if (uplo == dplasmaUpper) t = (t == dplasmaNoTrans) ? dplasmaTrans : dplasmaNoTrans;
    return t;
}

int dpl_run_trsm(parsec_context_t *ctx, int side, int uplo, int trans, int diag, double alpha,
                 const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B)
{
    const char *f = "dplasma_dtrsm";
    if (side != dplasmaLeft) {
        dpl_error(f, "side = Right is not provided by this library");
        return -101;
    }
    dpl_sym_dev_t s;
    dpl_ge_dev_t g;
    memset(&ctx->last, 0, sizeof ctx->last);
    int rc = dpl_sym_acquire(ctx, A, uplo, 1, &s, &ctx->last.stage_in_ms);
    if (rc) return rc;
    if (B->mb != A->mb && B->mt > 1) {
        dpl_sym_release(ctx, A, &s, 0, NULL);
        dpl_error(f, "B must use the row tile size of A");
        return -6;
    }
    if ((rc = dpl_ge_acquire(ctx, B, 1, &g, &ctx->last.stage_in_ms))) {
        dpl_sym_release(ctx, A, &s, 0, NULL);
        return rc;
    }
    rc = mi_dtrsm_LL(ctx->mi, lower_trans(uplo, trans), diag, alpha, s.N, s.NB, s.d, g.N, g.d, g.ld);
    if (rc) rc = dpl_mi_fail(f, rc);
    else mi_last_span_ms(ctx->mi, &ctx->last.total_ms);
    int r2 = dpl_ge_release(ctx, B, &g, rc == 0, &ctx->last.stage_out_ms);
    dpl_sym_release(ctx, A, &s, 0, NULL);
    return rc ? rc : r2;
}

/* --------------------------------------------------------------- potrf */
parsec_taskpool_t *dplasma_dpotrf_New(dplasma_enum_t uplo, parsec_tiled_matrix_t *A, int *info)
{
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dpotrf_New", "illegal value of uplo");
        return NULL;
    }
    *info = 0;
    parsec_taskpool_t *tp = (parsec_taskpool_t *)calloc(1, sizeof(parsec_taskpool_t));
    tp->kind = DPL_TP_POTRF;
    tp->uplo = uplo;
    tp->A = A;
    tp->info = info;
    return tp;
}

void dplasma_dpotrf_Destruct(parsec_taskpool_t *tp) { free(tp); }

void dplasma_dpotrf_setrecursive(parsec_taskpool_t *tp, int hmb)
{
    /* diagonal tiles larger than hmb are factored as tile Choleskys of their
     * own with hmb x hmb sub-tiles (mi_dpotrf_tiles_rec) */
    if (tp && hmb > 0) tp->hmb = hmb;
}

int dplasma_dpotrf(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A)
{
    int info = 0;
    parsec_taskpool_t *tp = dplasma_dpotrf_New(uplo, A, &info);
    if (tp) {
        parsec_context_add_taskpool(parsec, tp);
        int rc = parsec_context_wait(parsec);
        dplasma_dpotrf_Destruct(tp);
        if (rc) return rc;
    }
    return info;
}

int dplasma_dpotrf_rec(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A, int hmb)
{
    int info = 0;
    parsec_taskpool_t *tp = dplasma_dpotrf_New(uplo, A, &info);
    if (tp) {
        dplasma_dpotrf_setrecursive(tp, hmb);
        parsec_context_add_taskpool(parsec, tp);
        int rc = parsec_context_wait(parsec);
        dplasma_dpotrf_Destruct(tp);
        if (rc) return rc;
    }
    return info;
}

/* ---------------------------------------------------------------- trsm */
parsec_taskpool_t *dplasma_dtrsm_New(dplasma_enum_t side, dplasma_enum_t uplo, dplasma_enum_t trans,
                                     dplasma_enum_t diag, double alpha, const parsec_tiled_matrix_t *A,
                                     parsec_tiled_matrix_t *B)
{
    if (side != dplasmaLeft) {
        dpl_error("dplasma_dtrsm_New", "side = Right is not provided by this library");
        return NULL;
    }
    parsec_taskpool_t *tp = (parsec_taskpool_t *)calloc(1, sizeof(parsec_taskpool_t));
    tp->kind = DPL_TP_TRSM;
    tp->side = side;
    tp->uplo = uplo;
    tp->trans = trans;
    tp->diag = diag;
    tp->alpha = alpha;
    tp->Ac = A;
    tp->B = B;
    return tp;
}

void dplasma_dtrsm_Destruct(parsec_taskpool_t *tp) { free(tp); }

int dplasma_dtrsm(parsec_context_t *parsec, dplasma_enum_t side, dplasma_enum_t uplo, dplasma_enum_t trans,
                  dplasma_enum_t diag, double alpha, const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B)
{
    if (side != dplasmaLeft && side != dplasmaRight) {
        dpl_error("dplasma_dtrsm", "illegal value of side");
        return -1;
    }
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dtrsm", "illegal value of uplo");
        return -2;
    }
    if (trans != dplasmaConjTrans && trans != dplasmaNoTrans && trans != dplasmaTrans) {
        dpl_error("dplasma_dtrsm", "illegal value of trans");
        return -3;
    }
    if (diag != dplasmaUnit && diag != dplasmaNonUnit) {
        dpl_error("dplasma_dtrsm", "illegal value of diag");
        return -4;
    }
    if ((A->m != A->n) || ((side == dplasmaLeft) && (A->n != B->m)) ||
        ((side == dplasmaRight) && (A->n != B->n))) {
        dpl_error("dplasma_dtrsm", "illegal matrix A");
        return -6;
    }
    parsec_taskpool_t *tp = dplasma_dtrsm_New(side, uplo, trans, diag, alpha, A, B);
    if (!tp) return -101;
    parsec_context_add_taskpool(parsec, tp);
    int rc = parsec_context_wait(parsec);
    dplasma_dtrsm_Destruct(tp);
    return rc;
}

/* ------------------------------------------------------------- potrs/posv */
int dplasma_dpotrs(parsec_context_t *parsec, dplasma_enum_t uplo, const parsec_tiled_matrix_t *A,
                   parsec_tiled_matrix_t *B)
{
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dpotrs", "illegal value of uplo");
        return -1;
    }
    /* B must have N rows: the reference's two dplasma_ztrsm calls each
     * reject this shape with -6 and leave B untouched (ztrsm_wrapper.c:
     * 281-285), but zpotrs_wrapper.c:88-94 drops their codes and returns 0;
     * here the -6 is returned (B untouched as there). */
    if (A->m != A->n || B->m != A->n) {
        dpl_error("dplasma_dpotrs", "illegal matrix A or B (B must have as many rows as A)");
        return -6;
    }
    /* L L^T X = B as two triangular solves on one staged image of A
     * (zpotrs_wrapper.c:88-94: Upper = ConjTrans then NoTrans, Lower =
     * NoTrans then ConjTrans; both are L then L^T on the lower image) */
    dpl_sym_dev_t s;
    dpl_ge_dev_t g;
    memset(&parsec->last, 0, sizeof parsec->last);
    int rc = dpl_sym_acquire(parsec, A, uplo, 1, &s, &parsec->last.stage_in_ms);
    if (rc) return rc;
    if ((rc = dpl_ge_acquire(parsec, B, 1, &g, &parsec->last.stage_in_ms))) {
        dpl_sym_release(parsec, A, &s, 0, NULL);
        return rc;
    }
    rc = mi_dpotrs_tiles(parsec->mi, s.N, s.NB, s.d, g.N, g.d, g.ld);
    if (rc) rc = dpl_mi_fail("dplasma_dpotrs", rc);
    else mi_last_span_ms(parsec->mi, &parsec->last.total_ms);
    int r2 = dpl_ge_release(parsec, B, &g, rc == 0, &parsec->last.stage_out_ms);
    dpl_sym_release(parsec, A, &s, 0, NULL);
    return rc ? rc : r2;
}

int dplasma_dposv(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                  parsec_tiled_matrix_t *B)
{
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dposv", "illegal value of uplo");
        return -1;
    }
    /* checked before the factorization: the reference factors A and then
     * its solves fail with -6 (B untouched) while it returns info; here A
     * and B are both left untouched and -6 is returned */
    if (A->m != A->n || B->m != A->n) {
        dpl_error("dplasma_dposv", "illegal matrix A or B (B must have as many rows as A)");
        return -6;
    }
    dpl_sym_dev_t s;
    dpl_ge_dev_t g;
    memset(&parsec->last, 0, sizeof parsec->last);
    int rc = dpl_sym_acquire(parsec, A, uplo, 1, &s, &parsec->last.stage_in_ms);
    if (rc) return rc;
    int info = 0;
    double ms_f = 0.0, ms_s = 0.0;
    rc = mi_dpotrf_tiles(parsec->mi, s.N, s.NB, s.d, &info);
    if (rc) {
        dpl_sym_release(parsec, A, &s, 0, NULL);
        return dpl_mi_fail("dplasma_dposv", rc);
    }
    mi_last_span_ms(parsec->mi, &ms_f);
    if (info == 0) {
        if ((rc = dpl_ge_acquire(parsec, B, 1, &g, &parsec->last.stage_in_ms))) {
            dpl_sym_release(parsec, A, &s, 1, &parsec->last.stage_out_ms);
            return rc;
        }
        rc = mi_dpotrs_tiles(parsec->mi, s.N, s.NB, s.d, g.N, g.d, g.ld);
        if (rc) rc = dpl_mi_fail("dplasma_dposv", rc);
        else mi_last_span_ms(parsec->mi, &ms_s);
        int r2 = dpl_ge_release(parsec, B, &g, rc == 0, &parsec->last.stage_out_ms);
        if (!rc) rc = r2;
    }
    parsec->last.total_ms = ms_f + ms_s;
    int r3 = dpl_sym_release(parsec, A, &s, 1, &parsec->last.stage_out_ms);
    if (rc) return rc;
    if (r3) return r3;
    return info;
}

/* Mixed-precision solve with dsposv semantics (extension of dplasma_dposv,
 * src/zposv_wrapper.c:63-113): A is written back only when the FP64
 * fallback factored it (*iter < 0); B is read; X receives the solution. */
int dplasma_dposv_mixed(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                        parsec_tiled_matrix_t *B, parsec_tiled_matrix_t *X, int prec, int *iter)
{
    int it = 0;
    if (iter) *iter = 0;
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dposv_mixed", "illegal value of uplo");
        return -1;
    }
    if (prec != DPLASMA_MIXED_FP32 && prec != DPLASMA_MIXED_BF16X3) {
        dpl_error("dplasma_dposv_mixed", "illegal value of prec");
        return -1;
    }
    if (B->m != A->m || X->m != A->m || X->n != B->n) {
        dpl_error("dplasma_dposv_mixed", "B and X must be N x NRHS with N the order of A");
        return -1;
    }
    dpl_sym_dev_t s;
    dpl_ge_dev_t gb, gx;
    memset(&parsec->last, 0, sizeof parsec->last);
    int rc = dpl_sym_acquire(parsec, A, uplo, 1, &s, &parsec->last.stage_in_ms);
    if (rc) return rc;
    if ((rc = dpl_ge_acquire(parsec, B, 1, &gb, &parsec->last.stage_in_ms))) {
        dpl_sym_release(parsec, A, &s, 0, NULL);
        return rc;
    }
    if ((rc = dpl_ge_acquire(parsec, X, 0, &gx, &parsec->last.stage_in_ms))) {
        dpl_ge_release(parsec, B, &gb, 0, NULL);
        dpl_sym_release(parsec, A, &s, 0, NULL);
        return rc;
    }
    int info = 0;
    rc = mi_dsposv_tiles(parsec->mi, s.N, s.NB, s.d, gb.N, gb.d, gb.ld, gx.d, gx.ld, prec, &it, &info);
    if (rc) rc = dpl_mi_fail("dplasma_dposv_mixed", rc);
    else mi_last_span_ms(parsec->mi, &parsec->last.total_ms);
    int r1 = dpl_ge_release(parsec, X, &gx, rc == 0, &parsec->last.stage_out_ms);
    int r2 = dpl_ge_release(parsec, B, &gb, 0, NULL);
    int r3 = dpl_sym_release(parsec, A, &s, rc == 0 && it < 0, &parsec->last.stage_out_ms);
    if (iter) *iter = it;
    if (rc) return rc;
    if (r1) return r1;
    if (r2) return r2;
    if (r3) return r3;
    return info;
}

int dplasma_dsposv(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                   parsec_tiled_matrix_t *B, parsec_tiled_matrix_t *X, int *iter)
{
    return dplasma_dposv_mixed(parsec, uplo, A, B, X, DPLASMA_MIXED_FP32, iter);
}

/* Reverse-mode derivative of dplasma_dpotrf (extension, no counterpart in
 * the reference): L is the factor returned by dplasma_dpotrf (uplo part),
 * Lbar holds df/dL on entry (same uplo part) and the symmetric df/dA on
 * exit.  An Upper descriptor holds U = L^T and df/dU = (df/dL)^T: its lower
 * image is the lower problem, and the symmetric result maps back unchanged. */
int dplasma_dpotrf_adjoint(parsec_context_t *parsec, dplasma_enum_t uplo, const parsec_tiled_matrix_t *L,
                           parsec_tiled_matrix_t *Lbar)
{
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dpotrf_adjoint", "illegal value of uplo");
        return -1;
    }
    if (L->m != Lbar->m || L->n != Lbar->n || L->mb != Lbar->mb || L->nb != Lbar->nb) {
        dpl_error("dplasma_dpotrf_adjoint", "L and Lbar must have the same shape and tiling");
        return -2;
    }
    dpl_sym_dev_t sl, sg;
    memset(&parsec->last, 0, sizeof parsec->last);
    int rc = dpl_sym_acquire(parsec, L, uplo, 1, &sl, &parsec->last.stage_in_ms);
    if (rc) return rc;
    if ((rc = dpl_sym_acquire(parsec, Lbar, uplo, 1, &sg, &parsec->last.stage_in_ms))) {
        dpl_sym_release(parsec, L, &sl, 0, NULL);
        return rc;
    }
    rc = mi_dpotrf_adjoint_tiles(parsec->mi, sl.N, sl.NB, sl.d, sg.d);
    if (rc) rc = dpl_mi_fail("dplasma_dpotrf_adjoint", rc);
    else mi_last_span_ms(parsec->mi, &parsec->last.total_ms);
    int r2 = dpl_sym_release(parsec, Lbar, &sg, rc == 0, &parsec->last.stage_out_ms);
    dpl_sym_release(parsec, L, &sl, 0, NULL);
    return rc ? rc : r2;
}

/* dplasma_zpotri (src/zpotri_wrapper.c:54-82): A holds the Cholesky factor
 * of dplasma_dpotrf (uplo part) and receives the uplo part of A^-1.  The
 * reference runs trtri then lauum; here the inverse is the adjoint of
 * logdet A with respect to A (L^-T L^-1 from the same tile kernels).
 * Returns -1 for an illegal uplo, i > 0 when L(i,i) is exactly zero (dtrtri's
 * info; A unchanged), 0 otherwise. */
int dplasma_dpotri(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A)
{
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dpotri", "illegal value of uplo");
        return -1;
    }
    dpl_sym_dev_t s;
    memset(&parsec->last, 0, sizeof parsec->last);
    int rc = dpl_sym_acquire(parsec, A, uplo, 1, &s, &parsec->last.stage_in_ms);
    if (rc) return rc;
    int info = 0;
    rc = mi_dpotri_tiles(parsec->mi, s.N, s.NB, s.d, &info);
    if (rc) rc = dpl_mi_fail("dplasma_dpotri", rc);
    else mi_last_span_ms(parsec->mi, &parsec->last.total_ms);
    int r2 = dpl_sym_release(parsec, A, &s, rc == 0 && info == 0, &parsec->last.stage_out_ms);
    return rc ? rc : (r2 ? r2 : info);
}

/* dplasma_zpoinv (src/zpoinv_wrapper.c:109-135): factor and invert in place;
 * returns dplasma_dpotrf's info (the inverse is formed only when it is 0),
 * -1 for an illegal uplo. */
int dplasma_dpoinv(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A)
{
    if (uplo != dplasmaUpper && uplo != dplasmaLower) {
        dpl_error("dplasma_dpoinv", "illegal value of uplo");
        return -1;
    }
    dpl_sym_dev_t s;
    memset(&parsec->last, 0, sizeof parsec->last);
    int rc = dpl_sym_acquire(parsec, A, uplo, 1, &s, &parsec->last.stage_in_ms);
    if (rc) return rc;
    int info = 0;
    double ms_f = 0.0, ms_i = 0.0;
    rc = mi_dpotrf_tiles(parsec->mi, s.N, s.NB, s.d, &info);
    if (!rc) mi_last_span_ms(parsec->mi, &ms_f);
    if (!rc && info == 0) {
        int tinfo = 0;
        rc = mi_dpotri_tiles(parsec->mi, s.N, s.NB, s.d, &tinfo);
        if (!rc) mi_last_span_ms(parsec->mi, &ms_i);
        if (!rc && tinfo) info = tinfo;
    }
    if (rc) rc = dpl_mi_fail("dplasma_dpoinv", rc);
    parsec->last.total_ms = ms_f + ms_i;
    int r2 = dpl_sym_release(parsec, A, &s, rc == 0, &parsec->last.stage_out_ms);
    return rc ? rc : (r2 ? r2 : info);
}
