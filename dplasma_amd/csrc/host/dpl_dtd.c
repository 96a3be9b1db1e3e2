/*
 * dpl_dtd.c -- the insert-task (DTD) front-end of the tile Cholesky: the
 * PaRSEC DTD calls of tests/testing_zpotrf_dtd.c:91-302 and the task
 * classes of src/dtd_wrappers/z{potrf,trsm,herk,gemm}.c (d precision), over
 * the task-insertion engine of the kernel library (mi_dtd_* in
 * dplasma_mi.h).
 *
 * The reference's classes carry CPU chores (CORE_d*, src/cores) and CUDA
 * chores (cuSOLVER/cuBLAS); here every task runs on the GPU, on the tile
 * image of the descriptor staged into HBM at the taskpool's first task.
 * Each insertion is checked against the tile-Cholesky pattern of the tester
 * (which tiles a task may read and write, its side/uplo/trans/diag and
 * scalars) and mapped onto the lower image: for an Upper descriptor, whose
 * image holds L = U^T, tile A(i, j) of the caller is image tile (j, i).
 */
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include "dpl_internal.h"

typedef enum { DTD_POTRF = 1, DTD_TRSM, DTD_SYRK, DTD_GEMM } dtd_class_t;

struct parsec_task_class_s {
    dtd_class_t kind;
    int nargs;
};

struct parsec_dtd_tile_s {
    const parsec_tiled_matrix_t *A;
    int m, n;
};

typedef struct dpl_dtd_s {
    parsec_context_t *ctx;
    const parsec_tiled_matrix_t *A;   /* the descriptor bound at the first task */
    int uplo;
    dpl_sym_dev_t img;
    int bound;
    mi_dtd_t *eng;
    int *info;                        /* the info argument of the POTRF tasks */
    long ntasks;
    int err;
} dpl_dtd_t;

/* ------------------------------------------------ per-collection tile table */
typedef struct dtd_coll_s {
    const parsec_data_collection_t *dc;
    parsec_dtd_tile_t *tiles;
    int ntiles;
    struct dtd_coll_s *next;
} dtd_coll_t;
static dtd_coll_t *g_colls;

static dtd_coll_t *coll_of(parsec_data_collection_t *dc, int create)
{
    for (dtd_coll_t *c = g_colls; c; c = c->next)
        if (c->dc == dc) return c;
    if (!create) return NULL;
    const parsec_tiled_matrix_t *A = (const parsec_tiled_matrix_t *)dc;
    dtd_coll_t *c = (dtd_coll_t *)calloc(1, sizeof *c);
    c->dc = dc;
    c->ntiles = A->lmt * A->lnt;
    c->tiles = (parsec_dtd_tile_t *)calloc((size_t)(c->ntiles > 0 ? c->ntiles : 1), sizeof(parsec_dtd_tile_t));
    for (int n = 0; n < A->lnt; n++)
        for (int m = 0; m < A->lmt; m++) {
            parsec_dtd_tile_t *t = &c->tiles[m + (size_t)n * A->lmt];
            t->A = A;
            t->m = m;
            t->n = n;
        }
    c->next = g_colls;
    g_colls = c;
    return c;
}

int parsec_dtd_data_collection_init(parsec_data_collection_t *dc)
{
    return coll_of(dc, 1) ? 0 : -1;
}

void parsec_dtd_data_collection_fini(parsec_data_collection_t *dc)
{
    for (dtd_coll_t **p = &g_colls; *p; p = &(*p)->next)
        if ((*p)->dc == dc) {
            dtd_coll_t *c = *p;
            *p = c->next;
            free(c->tiles);
            free(c);
            return;
        }
}

parsec_data_key_t parsec_shim_data_key(const parsec_tiled_matrix_t *A, int m, int n)
{
    return (parsec_data_key_t)m + (parsec_data_key_t)n * (parsec_data_key_t)A->lmt;
}

parsec_dtd_tile_t *parsec_dtd_tile_of(parsec_data_collection_t *dc, parsec_data_key_t key)
{
    dtd_coll_t *c = coll_of(dc, 1);
    if (!c || key >= (parsec_data_key_t)c->ntiles) return NULL;
    return &c->tiles[key];
}

/* ------------------------------------------------------------ taskpool */
parsec_taskpool_t *parsec_dtd_taskpool_new(void)
{
    parsec_taskpool_t *tp = (parsec_taskpool_t *)calloc(1, sizeof(parsec_taskpool_t));
    tp->kind = DPL_TP_DTD;
    tp->dtd = (dpl_dtd_t *)calloc(1, sizeof(dpl_dtd_t));
    return tp;
}

void dpl_dtd_attach(parsec_taskpool_t *tp, parsec_context_t *ctx) { tp->dtd->ctx = ctx; }

long parsec_shim_dtd_ntasks(const parsec_taskpool_t *tp)
{
    return (tp && tp->kind == DPL_TP_DTD) ? tp->dtd->ntasks : 0;
}

static parsec_task_class_t *new_class(dtd_class_t kind, int nargs)
{
    parsec_task_class_t *tc = (parsec_task_class_t *)calloc(1, sizeof *tc);
    tc->kind = kind;
    tc->nargs = nargs;
    return tc;
}

parsec_task_class_t *parsec_dtd_create_dpotrf_task_class(parsec_taskpool_t *tp, int tile_full, int devices)
{
    (void)tp; (void)tile_full; (void)devices;
    return new_class(DTD_POTRF, 5);
}
parsec_task_class_t *parsec_dtd_create_dtrsm_task_class(parsec_taskpool_t *tp, int tile_full, int devices)
{
    (void)tp; (void)tile_full; (void)devices;
    return new_class(DTD_TRSM, 11);
}
parsec_task_class_t *parsec_dtd_create_dsyrk_task_class(parsec_taskpool_t *tp, int tile_full, int devices)
{
    (void)tp; (void)tile_full; (void)devices;
    return new_class(DTD_SYRK, 10);
}
parsec_task_class_t *parsec_dtd_create_dgemm_task_class(parsec_taskpool_t *tp, int tile_full, int devices)
{
    (void)tp; (void)tile_full; (void)devices;
    return new_class(DTD_GEMM, 13);
}

void parsec_dtd_task_class_release(parsec_taskpool_t *tp, parsec_task_class_t *tc)
{
    (void)tp;
    free(tc);
}

static int fail(dpl_dtd_t *d, const char *msg)
{
    dpl_error("parsec_dtd_insert_task_with_task_class", msg);
    if (!d->err) d->err = -2;
    return -2;
}

/* binds the taskpool to the descriptor of its first task */
static int bind(dpl_dtd_t *d, const parsec_tiled_matrix_t *A)
{
    if (d->bound) return d->A == A ? 0 : fail(d, "all tasks of a taskpool must use one descriptor");
    if (!d->ctx) return fail(d, "taskpool not added to a context (parsec_context_add_taskpool)");
    if (!dpl_is_sym_desc(A)) return fail(d, "the tile Cholesky tasks need a symmetric tile descriptor");
    if (A->mb != A->nb || A->m != A->n) return fail(d, "square tiles and a square matrix are required");
    d->A = A;
    d->uplo = parsec_shim_uplo(A);
    int rc = dpl_sym_acquire(d->ctx, A, d->uplo, 1, &d->img, NULL);
    if (rc) return rc;
    if ((rc = mi_dtd_create(d->ctx->mi, d->img.N, d->img.NB, d->img.d, &d->eng))) {
        dpl_sym_release(d->ctx, A, &d->img, 0, NULL);
        return dpl_mi_fail("parsec_dtd_insert_task_with_task_class", rc);
    }
    d->bound = 1;
    return 0;
}

/* image coordinates of the caller's tile A(m, n) */
static void img_mn(const dpl_dtd_t *d, const parsec_dtd_tile_t *t, int *r, int *c)
{
    if (d->uplo == dplasmaUpper) { *r = t->n; *c = t->m; }
    else { *r = t->m; *c = t->n; }
}

static int is_trans(int t) { return t == dplasmaTrans || t == dplasmaConjTrans; }

int parsec_dtd_insert_task_with_task_class(parsec_taskpool_t *tp, parsec_task_class_t *tc, int priority,
                                           int device_type, ...)
{
    (void)priority; (void)device_type;
    if (!tp || tp->kind != DPL_TP_DTD || !tc) {
        dpl_error("parsec_dtd_insert_task_with_task_class", "not a DTD taskpool / task class");
        return -1;
    }
    dpl_dtd_t *d = tp->dtd;
    void *arg[16];
    int n = 0;
    va_list ap;
    va_start(ap, device_type);
    /* exactly the class's (flag, pointer) pairs, then PARSEC_DTD_ARG_END */
    for (n = 0; n < tc->nargs && n < 16; n++) {
        (void)va_arg(ap, int);
        arg[n] = va_arg(ap, void *);
    }
    int end = va_arg(ap, int);
    va_end(ap);
    if (end != PARSEC_DTD_ARG_END) return fail(d, "argument list does not match the task class");
#define IV(i) (*(const int *)arg[i])
#define DV(i) (*(const double *)arg[i])
#define TL(i) ((const parsec_dtd_tile_t *)arg[i])
    int rc = 0, r, c;
    switch (tc->kind) {
    case DTD_POTRF: {
        /* uplo, m, A(k,k), lda, info */
        const parsec_dtd_tile_t *A = TL(2);
        if (!A || (rc = bind(d, A->A))) return rc ? rc : fail(d, "missing tile");
        img_mn(d, A, &r, &c);
        if (IV(0) != d->uplo || r != c) return fail(d, "POTRF: uplo or tile outside the pattern");
        d->info = (int *)arg[4];
        rc = mi_dtd_potrf(d->eng, r);
        break;
    }
    case DTD_TRSM: {
        /* side, uplo, transA, diag, m, n, alpha, A(k,k), lda, B, ldb */
        const parsec_dtd_tile_t *A = TL(7), *B = TL(9);
        if (!A || !B || (rc = bind(d, A->A))) return rc ? rc : fail(d, "missing tile");
        if (B->A != d->A) return fail(d, "all tasks of a taskpool must use one descriptor");
        const int side_ok = d->uplo == dplasmaLower ? IV(0) == dplasmaRight : IV(0) == dplasmaLeft;
        if (!side_ok || IV(1) != d->uplo || !is_trans(IV(2)) || IV(3) != dplasmaNonUnit || DV(6) != 1.0)
            return fail(d, "TRSM: side/uplo/trans/diag/alpha outside the tile-Cholesky pattern");
        int ka, kc, mb, kb;
        img_mn(d, A, &ka, &kc);
        img_mn(d, B, &mb, &kb);
        if (ka != kc || kb != ka || mb <= kb) return fail(d, "TRSM: tiles outside the tile-Cholesky pattern");
        rc = mi_dtd_trsm(d->eng, mb, kb);
        break;
    }
    case DTD_SYRK: {
        /* uplo, trans, n, k, alpha, A, lda, beta, C, ldc */
        const parsec_dtd_tile_t *A = TL(5), *C = TL(8);
        if (!A || !C || (rc = bind(d, A->A))) return rc ? rc : fail(d, "missing tile");
        if (C->A != d->A) return fail(d, "all tasks of a taskpool must use one descriptor");
        const int tr_ok = d->uplo == dplasmaLower ? IV(1) == dplasmaNoTrans : is_trans(IV(1));
        if (IV(0) != d->uplo || !tr_ok || DV(4) != -1.0 || DV(7) != 1.0)
            return fail(d, "SYRK: uplo/trans/alpha/beta outside the tile-Cholesky pattern");
        int ma, ka, mc, nc;
        img_mn(d, A, &ma, &ka);
        img_mn(d, C, &mc, &nc);
        if (mc != nc || ma != mc || ka >= ma) return fail(d, "SYRK: tiles outside the tile-Cholesky pattern");
        rc = mi_dtd_syrk(d->eng, ka, ma);
        break;
    }
    case DTD_GEMM: {
        /* transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc */
        const parsec_dtd_tile_t *A = TL(6), *B = TL(8), *C = TL(11);
        if (!A || !B || !C || (rc = bind(d, A->A))) return rc ? rc : fail(d, "missing tile");
        if (B->A != d->A || C->A != d->A) return fail(d, "all tasks of a taskpool must use one descriptor");
        const int tr_ok = d->uplo == dplasmaLower ? (IV(0) == dplasmaNoTrans && is_trans(IV(1)))
                                                  : (is_trans(IV(0)) && IV(1) == dplasmaNoTrans);
        if (!tr_ok || DV(5) != -1.0 || DV(10) != 1.0)
            return fail(d, "GEMM: trans/alpha/beta outside the tile-Cholesky pattern");
        /* image: C(i, j) -= L(i, k) L(j, k)^T, i > j > k, with L(i, k) from
         * the operand on C's row side and L(j, k) from the other */
        int ar, ak, br, bk, cr, cc;
        img_mn(d, A, &ar, &ak);
        img_mn(d, B, &br, &bk);
        img_mn(d, C, &cr, &cc);
        if (d->uplo == dplasmaUpper) {       /* C^T = B^T A: rows from B */
            int t;
            t = ar; ar = br; br = t;
            t = ak; ak = bk; bk = t;
        }
        if (ak != bk || ar != cr || br != cc || !(cr > cc && cc > ak))
            return fail(d, "GEMM: tiles outside the tile-Cholesky pattern");
        rc = mi_dtd_gemm(d->eng, cr, cc, ak);
        break;
    }
    default: return fail(d, "unknown task class");
    }
#undef IV
#undef DV
#undef TL
    if (rc) {
        if (!d->err) d->err = rc;
        return dpl_mi_fail("parsec_dtd_insert_task_with_task_class", rc);
    }
    d->ntasks++;
    return 0;
}

void parsec_dtd_data_flush(parsec_taskpool_t *tp, parsec_dtd_tile_t *tile)
{
    (void)tp;
    (void)tile;   /* tiles stay in the device image until the whole collection is flushed */
}

/* waits for every inserted task, writes the image back and unbinds */
static int finish(dpl_dtd_t *d)
{
    if (!d->bound) return d->err;
    int info = 0;
    long nt = 0;
    int rc = mi_dtd_wait(d->eng, &info, &nt);
    if (rc) rc = dpl_mi_fail("parsec_taskpool_wait", rc);
    int r2 = dpl_sym_release(d->ctx, d->A, &d->img, rc == 0, NULL);
    mi_dtd_destroy(d->eng);
    d->eng = NULL;
    d->bound = 0;
    if (d->info && info) *d->info = info;
    if (rc) return rc;
    if (r2) return r2;
    return d->err;
}

void parsec_dtd_data_flush_all(parsec_taskpool_t *tp, parsec_data_collection_t *dc)
{
    if (!tp || tp->kind != DPL_TP_DTD) return;
    dpl_dtd_t *d = tp->dtd;
    if (d->bound && (const parsec_data_collection_t *)d->A == dc) (void)finish(d);
}

int parsec_taskpool_wait(parsec_taskpool_t *tp)
{
    if (!tp) return -1;
    if (tp->kind != DPL_TP_DTD) return 0;
    return finish(tp->dtd);
}

int dpl_dtd_finish(parsec_taskpool_t *tp) { return finish(tp->dtd); }

void dpl_dtd_free(parsec_taskpool_t *tp)
{
    if (!tp->dtd) return;
    (void)finish(tp->dtd);
    free(tp->dtd);
    tp->dtd = NULL;
}
