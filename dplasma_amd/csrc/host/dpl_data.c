/*
 * dpl_data.c -- host <-> HBM staging and residency of tile matrices.
 *
 * The reference stages every tile a GPU task touches through PaRSEC's device
 * module and writes it back when the flow returns to the descriptor
 * (src/zpotrf_L.jdf:111-112,236-237; LAPACK-layout tiles through the 2-D
 * pitched copies of src/cuda/lapack_cuda_stage_in.c:18-162).  Here a whole
 * operand moves at once:
 *   - the uplo part of a symmetric/triangular operand becomes the packed
 *     lower tile image of dplasma_mi.h (one contiguous copy for a lower
 *     symmetric descriptor in tile storage; upper operands are transposed on
 *     the device so that every kernel works on L = U^T);
 *   - a general operand (right-hand sides) becomes a column-major array.
 * dplasma_advise_data_on_device() keeps an image resident in HBM across
 * calls; the host copy is then refreshed only by dplasma_data_sync_host(),
 * dplasma_data_release() or parsec_devices_release_memory().
 */
#include <stdlib.h>
#include <string.h>
#include "dpl_internal.h"

typedef struct res_s {
    const void *mat;
    parsec_context_t *ctx;
    const parsec_tiled_matrix_t *desc;
    int sym;              /* 1: packed lower image of the uplo part; 0: column-major */
    int uplo;
    double *d;
    int dev_valid, host_valid;
    struct res_s *next;
} res_t;

static res_t *g_res;

static size_t sym_image_bytes(const parsec_tiled_matrix_t *A)
{
    int MT = A->lmt;
    return (size_t)MT * (MT + 1) / 2 * (size_t)A->mb * A->mb * sizeof(double);
}

static int check_square(const char *f, const parsec_tiled_matrix_t *A)
{
    if (A->mtype != PARSEC_MATRIX_DOUBLE) { dpl_error(f, "only double precision descriptors"); return -2; }
    if (A->i != 0 || A->j != 0 || A->m != A->lm || A->n != A->ln) {
        dpl_error(f, "submatrix descriptors are not supported");
        return -2;
    }
    if (A->mb != A->nb || A->m != A->n) { dpl_error(f, "square tiles and matrix required"); return -2; }
    return 0;
}

static int tile_dim(int full, int b, int t, int nt) { return (t == nt - 1) ? full - t * b : b; }

/* host tile (m,n): address and pitch in bytes */
static const char *host_tile(const parsec_tiled_matrix_t *A, int m, int n, size_t *pitch)
{
    *pitch = (A->storage == PARSEC_MATRIX_LAPACK && A->dtype == PARSEC_SHIM_BLOCK_CYCLIC)
                 ? (size_t)A->llm * sizeof(double)
                 : (size_t)A->mb * sizeof(double);
    return (const char *)parsec_shim_tile_addr(A, m, n);
}

/* uplo tiles of A -> device packed image of the same triangle (lower: packed
 * lower layout; upper: packed upper layout) */
static int copy_triangle(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo, double *d,
                         int to_device)
{
    const int MT = A->lmt, NB = A->mb, N = A->m;
    if (A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC) {
        char *mat = (char *)parsec_shim_mat(A);
        size_t bytes = sym_image_bytes(A);
        if (to_device) return mi_memcpy_h2d(ctx->mi, d, mat, bytes);
        if (N % NB == 0) return mi_memcpy_d2h(ctx->mi, mat, d, bytes);
        /* ragged last tile row: its padding rows stay as the host has them */
        const size_t tb = (size_t)NB * NB * sizeof(double);
        const int tm = N - (MT - 1) * NB;
        for (int n = 0; n < MT; n++) {
            int64_t first = (uplo == dplasmaLower) ? (int64_t)n * MT - ((int64_t)n * (n - 1)) / 2
                                                   : ((int64_t)n * (n + 1)) / 2;
            int ntile = (uplo == dplasmaLower) ? MT - n : n + 1;
            int rc;
            if (uplo == dplasmaLower) {
                /* column n: tiles n..MT-2 are full, tile MT-1 is ragged */
                if (ntile > 1 && (rc = mi_memcpy_d2h(ctx->mi, mat + first * tb, d + first * (int64_t)NB * NB,
                                                      (ntile - 1) * tb)))
                    return rc;
                int64_t last = first + ntile - 1;
                int tn = (n == MT - 1) ? tm : NB;
                if ((rc = mi_memcpy2d_d2h(ctx->mi, mat + last * tb, NB * sizeof(double),
                                          d + last * (int64_t)NB * NB, NB * sizeof(double),
                                          tm * sizeof(double), tn)))
                    return rc;
            } else {
                /* column n holds tiles 0..n; only column MT-1 is ragged (in width) */
                int tn = (n == MT - 1) ? tm : NB;
                for (int m = 0; m <= n; m++) {
                    int64_t idx = first + m;
                    int tr = (m == MT - 1) ? tm : NB;
                    if ((rc = mi_memcpy2d_d2h(ctx->mi, mat + idx * tb, NB * sizeof(double),
                                              d + idx * (int64_t)NB * NB, NB * sizeof(double),
                                              tr * sizeof(double), tn)))
                        return rc;
                }
            }
        }
        return 0;
    }
    for (int n = 0; n < MT; n++) {
        int lo = (uplo == dplasmaLower) ? n : 0, hi = (uplo == dplasmaLower) ? MT : n + 1;
        for (int m = lo; m < hi; m++) {
            size_t hp;
            char *h = (char *)host_tile(A, m, n, &hp);
            int64_t idx = (uplo == dplasmaLower) ? (int64_t)n * MT - ((int64_t)n * (n - 1)) / 2 + (m - n)
                                                 : ((int64_t)n * (n + 1)) / 2 + m;
            double *dt = d + idx * (int64_t)NB * NB;
            size_t w = (size_t)tile_dim(N, NB, m, MT) * sizeof(double);
            size_t hgt = (size_t)tile_dim(N, NB, n, MT);
            int rc = to_device ? mi_memcpy2d_h2d(ctx->mi, dt, NB * sizeof(double), h, hp, w, hgt)
                               : mi_memcpy2d_d2h(ctx->mi, h, hp, dt, NB * sizeof(double), w, hgt);
            if (rc) return rc;
        }
    }
    return 0;
}

static int load_sym(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo, double *d)
{
    if (uplo == dplasmaLower) return copy_triangle(ctx, A, uplo, d, 1);
    double *tmp = NULL;
    int rc = mi_malloc(ctx->mi, (void **)&tmp, sym_image_bytes(A));
    if (rc) return rc;
    rc = copy_triangle(ctx, A, uplo, tmp, 1);
    if (!rc) rc = mi_dsym_upper_to_lower(ctx->mi, A->m, A->mb, tmp, d);
    mi_free(ctx->mi, tmp);
    return rc;
}

static int store_sym(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo, double *d)
{
    if (uplo == dplasmaLower) return copy_triangle(ctx, A, uplo, d, 0);
    double *tmp = NULL;
    int rc = mi_malloc(ctx->mi, (void **)&tmp, sym_image_bytes(A));
    if (rc) return rc;
    rc = mi_dsym_lower_to_upper(ctx->mi, A->m, A->mb, d, tmp);
    if (!rc) rc = copy_triangle(ctx, A, uplo, tmp, 0);
    mi_free(ctx->mi, tmp);
    return rc;
}

/* general descriptor <-> column-major device array (ld = M) */
static int copy_general(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, double *d, int to_device)
{
    const int M = A->m, N = A->n;
    if (A->storage == PARSEC_MATRIX_LAPACK) {
        char *mat = (char *)parsec_shim_mat(A);
        size_t hp = (size_t)A->llm * sizeof(double);
        return to_device ? mi_memcpy2d_h2d(ctx->mi, d, M * sizeof(double), mat, hp, M * sizeof(double), N)
                         : mi_memcpy2d_d2h(ctx->mi, mat, hp, d, M * sizeof(double), M * sizeof(double), N);
    }
    for (int n = 0; n < A->nt; n++)
        for (int m = 0; m < A->mt; m++) {
            size_t hp;
            char *h = (char *)host_tile(A, m, n, &hp);
            double *dt = d + (int64_t)n * A->nb * M + (int64_t)m * A->mb;
            size_t w = (size_t)tile_dim(M, A->mb, m, A->mt) * sizeof(double);
            size_t hgt = (size_t)tile_dim(N, A->nb, n, A->nt);
            int rc = to_device ? mi_memcpy2d_h2d(ctx->mi, dt, M * sizeof(double), h, hp, w, hgt)
                               : mi_memcpy2d_d2h(ctx->mi, h, hp, dt, M * sizeof(double), w, hgt);
            if (rc) return rc;
        }
    return 0;
}

static res_t *find(const void *mat, int sym)
{
    for (res_t *e = g_res; e; e = e->next)
        if (e->mat == mat && e->sym == sym) return e;
    return NULL;
}

static int sym_uplo_of(const parsec_tiled_matrix_t *A, int uplo)
{
    if (A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC) return parsec_shim_uplo(A);
    return uplo;
}

int dpl_sym_acquire(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo, int load,
                    dpl_sym_dev_t *out, double *ms)
{
    const char *f = "dpl_sym_acquire";
    int rc = check_square(f, A);
    if (rc) return rc;
    if (uplo != dplasmaLower && uplo != dplasmaUpper) { dpl_error(f, "uplo must be Lower or Upper"); return -2; }
    if (A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC && parsec_shim_uplo(A) != uplo) {
        dpl_error(f, "uplo differs from the triangle the symmetric descriptor stores");
        return -2;
    }
    if (!parsec_shim_mat(A)) { dpl_error(f, "descriptor has no storage"); return -2; }
    memset(out, 0, sizeof *out);
    out->N = A->m;
    out->NB = A->mb;
    out->MT = A->lmt;
    out->uplo = uplo;
    double t0 = dpl_now();
    int copied = 0;
    res_t *e = find(parsec_shim_mat(A), 1);
    if (e) {
        if (e->uplo != uplo) { dpl_error(f, "resident image holds the other triangle"); return -2; }
        if (!e->dev_valid && load) {
            if ((rc = load_sym(ctx, A, uplo, e->d))) return dpl_mi_fail(f, rc);
            e->dev_valid = 1;
            copied = 1;
        }
        out->d = e->d;
        out->entry = e;
    } else {
        if ((rc = mi_malloc(ctx->mi, (void **)&out->d, sym_image_bytes(A)))) return dpl_mi_fail(f, rc);
        out->owned = 1;
        if (load && (rc = load_sym(ctx, A, uplo, out->d))) {
            mi_free(ctx->mi, out->d);
            return dpl_mi_fail(f, rc);
        }
        copied = load;
    }
    if (ms && copied) *ms += (dpl_now() - t0) * 1e3;
    return 0;
}

int dpl_sym_release(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, dpl_sym_dev_t *s, int dirty,
                    double *ms)
{
    int rc = 0;
    double t0 = dpl_now();
    if (s->entry) {
        res_t *e = (res_t *)s->entry;
        if (dirty) { e->host_valid = 0; e->dev_valid = 1; }
    } else if (s->d) {
        if (dirty) rc = store_sym(ctx, A, s->uplo, s->d);
        mi_free(ctx->mi, s->d);
    }
    s->d = NULL;
    if (ms) *ms += (dpl_now() - t0) * 1e3;
    return rc ? dpl_mi_fail("dpl_sym_release", rc) : 0;
}

int dpl_ge_acquire(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int load, dpl_ge_dev_t *out,
                   double *ms)
{
    const char *f = "dpl_ge_acquire";
    int rc;
    if (A->mtype != PARSEC_MATRIX_DOUBLE || A->dtype != PARSEC_SHIM_BLOCK_CYCLIC) {
        dpl_error(f, "a double precision block-cyclic descriptor is required");
        return -2;
    }
    if (A->i != 0 || A->j != 0 || A->m != A->lm || A->n != A->ln) {
        dpl_error(f, "submatrix descriptors are not supported");
        return -2;
    }
    if (!parsec_shim_mat(A)) { dpl_error(f, "descriptor has no storage"); return -2; }
    memset(out, 0, sizeof *out);
    out->M = A->m;
    out->N = A->n;
    out->ld = A->m > 0 ? A->m : 1;
    double t0 = dpl_now();
    int copied = 0;
    res_t *e = find(parsec_shim_mat(A), 0);
    if (e) {
        if (!e->dev_valid && load) {
            if ((rc = copy_general(ctx, A, e->d, 1))) return dpl_mi_fail(f, rc);
            e->dev_valid = 1;
            copied = 1;
        }
        out->d = e->d;
        out->entry = e;
    } else {
        size_t bytes = (size_t)A->m * A->n * sizeof(double);
        if ((rc = mi_malloc(ctx->mi, (void **)&out->d, bytes ? bytes : 8))) return dpl_mi_fail(f, rc);
        out->owned = 1;
        if (load && (rc = copy_general(ctx, A, out->d, 1))) {
            mi_free(ctx->mi, out->d);
            return dpl_mi_fail(f, rc);
        }
        copied = load;
    }
    if (ms && copied) *ms += (dpl_now() - t0) * 1e3;
    return 0;
}

int dpl_ge_release(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, dpl_ge_dev_t *g, int dirty,
                   double *ms)
{
    int rc = 0;
    double t0 = dpl_now();
    if (g->entry) {
        res_t *e = (res_t *)g->entry;
        if (dirty) { e->host_valid = 0; e->dev_valid = 1; }
    } else if (g->d) {
        if (dirty) rc = copy_general(ctx, A, g->d, 0);
        mi_free(ctx->mi, g->d);
    }
    g->d = NULL;
    if (ms) *ms += (dpl_now() - t0) * 1e3;
    return rc ? dpl_mi_fail("dpl_ge_release", rc) : 0;
}

int dpl_make_resident(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo)
{
    const char *f = "dplasma_advise_data_on_device";
    int sym = (A->dtype == PARSEC_SHIM_SYM_BLOCK_CYCLIC) ||
              ((uplo == dplasmaLower || uplo == dplasmaUpper) && A->m == A->n);
    int u = sym_uplo_of(A, uplo);
    if (!parsec_shim_mat(A)) { dpl_error(f, "descriptor has no storage"); return -2; }
    res_t *e = find(parsec_shim_mat(A), sym);
    if (e) return 0;
    int rc;
    e = (res_t *)calloc(1, sizeof(res_t));
    e->mat = parsec_shim_mat(A);
    e->ctx = ctx;
    e->desc = A;
    e->sym = sym;
    e->uplo = u;
    if (sym) {
        if ((rc = check_square(f, A))) { free(e); return rc; }
        if ((rc = mi_malloc(ctx->mi, (void **)&e->d, sym_image_bytes(A)))) { free(e); return dpl_mi_fail(f, rc); }
        rc = load_sym(ctx, A, u, e->d);
    } else {
        size_t bytes = (size_t)A->m * A->n * sizeof(double);
        if ((rc = mi_malloc(ctx->mi, (void **)&e->d, bytes ? bytes : 8))) { free(e); return dpl_mi_fail(f, rc); }
        rc = copy_general(ctx, A, e->d, 1);
    }
    if (rc) {
        mi_free(ctx->mi, e->d);
        free(e);
        return dpl_mi_fail(f, rc);
    }
    e->dev_valid = 1;
    e->host_valid = 1;
    e->next = g_res;
    g_res = e;
    return 0;
}

static int writeback(res_t *e, const parsec_tiled_matrix_t *A)
{
    if (!e->dev_valid || e->host_valid) return 0;
    int rc = e->sym ? store_sym(e->ctx, A, e->uplo, e->d) : copy_general(e->ctx, A, e->d, 0);
    if (!rc) e->host_valid = 1;
    return rc;
}

int dpl_sync_host(parsec_context_t *ctx, const parsec_tiled_matrix_t *A)
{
    (void)ctx;
    const void *mat = parsec_shim_mat(A);
    for (res_t *e = g_res; e; e = e->next)
        if (e->mat == mat) {
            int rc = writeback(e, A);
            if (rc) return dpl_mi_fail("dplasma_data_sync_host", rc);
        }
    return 0;
}

static void drop_where(const void *mat, const parsec_tiled_matrix_t *A, int do_writeback)
{
    res_t **pp = &g_res;
    while (*pp) {
        res_t *e = *pp;
        if (mat == NULL || e->mat == mat) {
            if (do_writeback) {
                int rc = writeback(e, A ? A : e->desc);
                if (rc) dpl_mi_fail("writeback", rc);
            }
            mi_free(e->ctx->mi, e->d);
            *pp = e->next;
            free(e);
        } else {
            pp = &e->next;
        }
    }
}

int dpl_release(parsec_context_t *ctx, const parsec_tiled_matrix_t *A)
{
    (void)ctx;
    const void *mat = parsec_shim_mat(A);
    if (mat) drop_where(mat, A, 1);
    return 0;
}

void dpl_release_all(void) { drop_where(NULL, NULL, 1); }

/* parsec_data_free of a resident host buffer: the image goes away unwritten */
void dpl_forget_mat(const void *mat)
{
    if (mat) drop_where(mat, NULL, 0);
}
