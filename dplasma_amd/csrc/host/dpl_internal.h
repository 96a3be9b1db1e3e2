/*
 * dpl_internal.h -- internals of the C host library libdplasma.so: the
 * context, deferred operations ("taskpools") and HBM residency of tile
 * matrices.  The host library reaches the GPU only through dplasma_mi.h.
 */
#ifndef DPL_INTERNAL_H
#define DPL_INTERNAL_H

#include <stdio.h>
#include "dplasma.h"
#include "dplasma_mi.h"

#define DPL_MAX_PENDING 64

typedef enum {
    DPL_TP_POTRF = 1,
    DPL_TP_TRSM  = 2,
    DPL_TP_PLGSY = 3,
    DPL_TP_DTD   = 4
} dpl_tp_kind_t;

struct dpl_dtd_s;

struct parsec_taskpool_s {
    dpl_tp_kind_t kind;
    int uplo, side, trans, diag;
    double alpha, bump;
    unsigned long long seed;
    parsec_tiled_matrix_t *A;
    const parsec_tiled_matrix_t *Ac;
    parsec_tiled_matrix_t *B;
    int *info;
    int hmb;
    int rc;               /* runtime status of the last execution */
    struct dpl_dtd_s *dtd;       /* DTD taskpools: the insertion state */
};

struct parsec_context_s {
    mi_ctx_t *mi;
    int device;
    int nb_cores;
    parsec_taskpool_t *pending[DPL_MAX_PENDING];
    int npending;
    dplasma_mi_timing_t last;
    dplasma_mi_check_t last_check;
};

/* device image of the uplo part of a matrix: packed lower tiles (see
 * dplasma_mi.h), transposed when uplo == Upper */
typedef struct dpl_sym_dev_s {
    double *d;
    int N, NB, MT;
    int uplo;
    int owned;            /* 1: temporary, freed by dpl_sym_release */
    void *entry;          /* residency entry when resident */
} dpl_sym_dev_t;

/* device image of a general matrix as column-major (ld = M) */
typedef struct dpl_ge_dev_s {
    double *d;
    int M, N, ld;
    int owned;
    void *entry;
} dpl_ge_dev_t;

void dpl_error(const char *func, const char *msg);
int  dpl_mi_fail(const char *func, int rc);

/* staging (dpl_data.c) */
int  dpl_sym_acquire(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo, int load,
                     dpl_sym_dev_t *out, double *ms);
/* write back (if dirty and not resident) and free temporaries */
int  dpl_sym_release(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, dpl_sym_dev_t *s,
                     int dirty, double *ms);
int  dpl_ge_acquire(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int load,
                    dpl_ge_dev_t *out, double *ms);
int  dpl_ge_release(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, dpl_ge_dev_t *g,
                    int dirty, double *ms);
int  dpl_make_resident(parsec_context_t *ctx, const parsec_tiled_matrix_t *A, int uplo);
int  dpl_sync_host(parsec_context_t *ctx, const parsec_tiled_matrix_t *A);
int  dpl_release(parsec_context_t *ctx, const parsec_tiled_matrix_t *A);
void dpl_release_all(void);
void dpl_forget_mat(const void *mat);
int  dpl_is_sym_desc(const parsec_tiled_matrix_t *A);

/* operations (dpl_potrf.c) */
int dpl_run_potrf(parsec_context_t *ctx, int uplo, parsec_tiled_matrix_t *A, int hmb, int *info);
int dpl_run_trsm(parsec_context_t *ctx, int side, int uplo, int trans, int diag, double alpha,
                 const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B);
int dpl_run_plgsy(parsec_context_t *ctx, double bump, int uplo, parsec_tiled_matrix_t *A,
                  unsigned long long seed);

double dpl_now(void);

/* DTD taskpools (dpl_dtd.c) */
void dpl_dtd_attach(parsec_taskpool_t *tp, parsec_context_t *ctx);
int  dpl_dtd_finish(parsec_taskpool_t *tp);
void dpl_dtd_free(parsec_taskpool_t *tp);

#endif
