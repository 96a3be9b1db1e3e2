/*
 * parsec_shim.c -- the PaRSEC calls DPLASMA's testers and callers make,
 * provided for one rank driving one MI355X.
 *
 * A context wraps one GPU context of the kernel library.  Taskpools are
 * deferred operations: parsec_context_add_taskpool queues them and
 * parsec_context_wait runs them in order, blocking, which is what
 * dplasma_wait_until_completion (src/dplasmaaux.h:97-101) gives a single
 * rank.  Descriptors follow the field semantics of PaRSEC's two-dim block
 * cyclic matrices on one rank (P = Q = 1); the tile storage of the symmetric
 * descriptor is the uplo tiles column by column, each mb*nb with lda = mb.
 */
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "dpl_internal.h"

double dpl_now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

void dpl_error(const char *func, const char *msg)
{
    fprintf(stderr, "[dplasma] %s: %s\n", func, msg);
}

int dpl_mi_fail(const char *func, int rc)
{
    fprintf(stderr, "[dplasma] %s: GPU failure %d: %s\n", func, rc, mi_last_error());
    return rc < 0 ? rc : -1;
}

/* ------------------------------------------------------------ runtime */
static int parse_device(int *pargc, char **pargv[])
{
    int dev = 0;
    const char *e = getenv("PARSEC_MCA_device_hip_device");
    if (e) dev = atoi(e);
    if (pargc && pargv && *pargv) {
        for (int i = 0; i + 2 < *pargc; i++)
            if (!strcmp((*pargv)[i], "--mca") && !strcmp((*pargv)[i + 1], "device_hip_device"))
                dev = atoi((*pargv)[i + 2]);
    }
    return dev;
}

parsec_context_t *parsec_init(int nb_cores, int *pargc, char **pargv[])
{
    parsec_context_t *ctx = (parsec_context_t *)calloc(1, sizeof(parsec_context_t));
    ctx->nb_cores = nb_cores;
    ctx->device = parse_device(pargc, pargv);
    int rc = mi_ctx_create(ctx->device, &ctx->mi);
    if (rc) {
        fprintf(stderr, "[dplasma] parsec_init: no usable GPU (%s)\n", mi_last_error());
        free(ctx);
        return NULL;
    }
    return ctx;
}

int parsec_fini(parsec_context_t **pctx)
{
    if (!pctx || !*pctx) return 0;
    parsec_context_t *ctx = *pctx;
    dpl_release_all();
    mi_ctx_destroy(ctx->mi);
    free(ctx);
    *pctx = NULL;
    return 0;
}

int parsec_context_add_taskpool(parsec_context_t *ctx, parsec_taskpool_t *tp)
{
    if (!ctx || !tp) return -1;
    if (tp->kind == DPL_TP_DTD) dpl_dtd_attach(tp, ctx);
    if (ctx->npending >= DPL_MAX_PENDING) return -1;
    ctx->pending[ctx->npending++] = tp;
    return 0;
}

int parsec_context_start(parsec_context_t *ctx)
{
    (void)ctx;
    return 0;
}

int parsec_context_wait(parsec_context_t *ctx)
{
    int status = 0;
    for (int i = 0; i < ctx->npending; i++) {
        parsec_taskpool_t *tp = ctx->pending[i];
        switch (tp->kind) {
        case DPL_TP_POTRF: tp->rc = dpl_run_potrf(ctx, tp->uplo, tp->A, tp->hmb, tp->info); break;
        case DPL_TP_TRSM:
            tp->rc = dpl_run_trsm(ctx, tp->side, tp->uplo, tp->trans, tp->diag, tp->alpha, tp->Ac, tp->B);
            break;
        case DPL_TP_PLGSY: tp->rc = dpl_run_plgsy(ctx, tp->bump, tp->uplo, tp->A, tp->seed); break;
        case DPL_TP_DTD: tp->rc = dpl_dtd_finish(tp); break;
        default: tp->rc = -1;
        }
        if (tp->rc && !status) status = tp->rc;
    }
    ctx->npending = 0;
    return status;
}

void parsec_devices_release_memory(void) { dpl_release_all(); }

void parsec_devices_reset_load(parsec_context_t *ctx) { (void)ctx; }

void parsec_taskpool_free(parsec_taskpool_t *tp)
{
    if (!tp) return;
    if (tp->kind == DPL_TP_DTD) dpl_dtd_free(tp);
    free(tp);
}

void parsec_taskpool_sync_ids(void) {}

/* ---------------------------------------------------------- descriptors */
size_t parsec_datadist_getsizeoftype(parsec_matrix_type_t type)
{
    switch (type) {
    case PARSEC_MATRIX_BYTE: return 1;
    case PARSEC_MATRIX_INTEGER: return sizeof(int);
    case PARSEC_MATRIX_FLOAT: return sizeof(float);
    case PARSEC_MATRIX_DOUBLE: return sizeof(double);
    case PARSEC_MATRIX_COMPLEX_FLOAT: return 2 * sizeof(float);
    case PARSEC_MATRIX_COMPLEX_DOUBLE: return 2 * sizeof(double);
    }
    return 0;
}

static void tiled_init(parsec_tiled_matrix_t *t, parsec_matrix_type_t mtype,
                       parsec_matrix_storage_t storage, int dtype, int mb, int nb, int lm, int ln,
                       int i, int j, int m, int n)
{
    memset(t, 0, sizeof *t);
    t->super.nodes = 1;
    t->super.myrank = 0;
    t->mtype = mtype;
    t->storage = storage;
    t->dtype = dtype;
    t->mb = mb;
    t->nb = nb;
    t->bsiz = mb * nb;
    t->lm = lm;
    t->ln = ln;
    t->lmt = (lm + mb - 1) / mb;
    t->lnt = (ln + nb - 1) / nb;
    t->i = i;
    t->j = j;
    t->m = m;
    t->n = n;
    t->mt = (i + m - 1) / mb - i / mb + 1;
    t->nt = (j + n - 1) / nb - j / nb + 1;
    if (m == 0) t->mt = 0;
    if (n == 0) t->nt = 0;
    t->llm = lm;
    t->lln = ln;
    t->slm = m;
    t->sln = n;
}

int parsec_matrix_block_cyclic_init(parsec_matrix_block_cyclic_t *dc, parsec_matrix_type_t mtype,
                                    parsec_matrix_storage_t storage, int myrank, int mb, int nb,
                                    int lm, int ln, int i, int j, int m, int n, int P, int Q,
                                    int kp, int kq, int ip, int jq)
{
    if (P * Q != 1 || myrank != 0) {
        fprintf(stderr, "[dplasma] parsec_matrix_block_cyclic_init: single rank only (P=%d Q=%d)\n", P, Q);
    }
    tiled_init(&dc->super, mtype, storage, PARSEC_SHIM_BLOCK_CYCLIC, mb, nb, lm, ln, i, j, m, n);
    memset(&dc->grid, 0, sizeof dc->grid);
    dc->grid.rows = P;
    dc->grid.cols = Q;
    dc->grid.krows = kp;
    dc->grid.kcols = kq;
    dc->grid.ip = ip;
    dc->grid.jq = jq;
    dc->super.nb_local_tiles = dc->super.lmt * dc->super.lnt;
    if (storage == PARSEC_MATRIX_LAPACK) {
        /* one column-major array; bsiz keeps its meaning for allocations */
        dc->super.nb_local_tiles = dc->super.lmt * dc->super.lnt;
    }
    dc->mat = NULL;
    return 0;
}

int parsec_matrix_sym_block_cyclic_init(parsec_matrix_sym_block_cyclic_t *dc, parsec_matrix_type_t mtype,
                                        int myrank, int mb, int nb, int lm, int ln, int i, int j,
                                        int m, int n, int P, int Q, int uplo)
{
    if (P * Q != 1 || myrank != 0) {
        fprintf(stderr, "[dplasma] parsec_matrix_sym_block_cyclic_init: single rank only (P=%d Q=%d)\n", P, Q);
    }
    tiled_init(&dc->super, mtype, PARSEC_MATRIX_TILE, PARSEC_SHIM_SYM_BLOCK_CYCLIC, mb, nb, lm, ln,
               i, j, m, n);
    memset(&dc->grid, 0, sizeof dc->grid);
    dc->grid.rows = P;
    dc->grid.cols = Q;
    dc->grid.krows = 1;
    dc->grid.kcols = 1;
    dc->uplo = uplo;
    int MT = dc->super.lmt;
    dc->super.nb_local_tiles = MT * (MT + 1) / 2;
    dc->mat = NULL;
    return 0;
}

void parsec_tiled_matrix_destroy(parsec_tiled_matrix_t *t)
{
    if (!t) return;
    dpl_release(NULL, t);
    if (t->super.key) {
        free(t->super.key);
        t->super.key = NULL;
    }
}

void parsec_data_collection_set_key(parsec_data_collection_t *dc, const char *key)
{
    if (dc->key) free(dc->key);
    dc->key = key ? strdup(key) : NULL;
}

/* pinned allocations, remembered so parsec_data_free frees them the right way */
typedef struct pin_s { void *p; struct pin_s *next; } pin_t;
static pin_t *g_pins;

void *parsec_data_allocate(size_t size)
{
    void *p = NULL;
    if (size == 0) return NULL;
    if (mi_host_alloc(&p, size) == 0 && p) {
        pin_t *e = (pin_t *)malloc(sizeof(pin_t));
        e->p = p;
        e->next = g_pins;
        g_pins = e;
        return p;
    }
    return malloc(size);   /* no GPU runtime: pageable memory */
}

void parsec_data_free(void *p)
{
    if (!p) return;
    dpl_forget_mat(p);
    pin_t **pp = &g_pins;
    while (*pp) {
        if ((*pp)->p == p) {
            pin_t *e = *pp;
            *pp = e->next;
            free(e);
            mi_host_free(p);
            return;
        }
        pp = &(*pp)->next;
    }
    free(p);
}

/* ------------------------------------------------------------- helpers */
void *parsec_shim_mat(const parsec_tiled_matrix_t *A)
{
    if (!A) return NULL;
    if (A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC) return ((const parsec_matrix_sym_block_cyclic_t *)A)->mat;
    if (A->dtype == PARSEC_SHIM_BLOCK_CYCLIC) return ((const parsec_matrix_block_cyclic_t *)A)->mat;
    return NULL;
}

int parsec_shim_uplo(const parsec_tiled_matrix_t *A)
{
    if (A && A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC) return ((const parsec_matrix_sym_block_cyclic_t *)A)->uplo;
    return PARSEC_MATRIX_FULL;
}

int dpl_is_sym_desc(const parsec_tiled_matrix_t *A) { return A && A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC; }

size_t parsec_shim_sizeof(int which)
{
    switch (which) {
    case 0: return sizeof(parsec_tiled_matrix_t);
    case 1: return sizeof(parsec_matrix_block_cyclic_t);
    case 2: return sizeof(parsec_matrix_sym_block_cyclic_t);
    }
    return 0;
}

void *parsec_shim_tile_addr(const parsec_tiled_matrix_t *A, int m, int n)
{
    char *mat = (char *)parsec_shim_mat(A);
    size_t elt = parsec_datadist_getsizeoftype(A->mtype);
    if (!mat || m < 0 || n < 0 || m >= A->lmt || n >= A->lnt) return NULL;
    if (A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC) {
        int MT = A->lmt;
        int64_t idx;
        if (parsec_shim_uplo(A) == PARSEC_MATRIX_LOWER) {
            if (m < n) return NULL;
            idx = (int64_t)n * MT - ((int64_t)n * (n - 1)) / 2 + (m - n);
        } else {
            if (m > n) return NULL;
            idx = ((int64_t)n * (n + 1)) / 2 + m;
        }
        return mat + idx * (int64_t)A->bsiz * elt;
    }
    if (A->storage == PARSEC_MATRIX_LAPACK)
        return mat + ((int64_t)n * A->nb * A->llm + (int64_t)m * A->mb) * elt;
    return mat + ((int64_t)n * A->lmt + m) * (int64_t)A->bsiz * elt;
}
