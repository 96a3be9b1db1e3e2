/*
 * parsec_shim.h -- the slice of the PaRSEC runtime API that DPLASMA's
 * Cholesky testers and callers touch, re-provided for a single rank on one
 * MI355X.  PaRSEC itself is not part of this library: there is no task
 * runtime, no MPI and no device-module layer.  A context owns one GPU (its
 * streams, kernels and HBM tile mirrors); a "taskpool" is a deferred
 * operation that runs on the GPU when the context is waited on.
 *
 * Source compatibility, not binary compatibility, with PaRSEC is the goal:
 * the field and function names below are the ones DPLASMA reads
 * (tests/common.h:185-193, tests/testing_zpotrf.c:42-45,138-139,
 * src/dplasma_zcheck.c:79-86, src/include/dplasma/types_lapack.h:100-170).
 */
#ifndef DPLASMA_MI_PARSEC_SHIM_H
#define DPLASMA_MI_PARSEC_SHIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types (parsec matrix.h names) */
typedef enum parsec_matrix_type_e {
    PARSEC_MATRIX_BYTE           = 0,
    PARSEC_MATRIX_INTEGER        = 1,
    PARSEC_MATRIX_FLOAT          = 2,
    PARSEC_MATRIX_DOUBLE         = 3,
    PARSEC_MATRIX_COMPLEX_FLOAT  = 4,
    PARSEC_MATRIX_COMPLEX_DOUBLE = 5
} parsec_matrix_type_t;

/* storage of the local tiles */
typedef enum parsec_matrix_storage_e {
    PARSEC_MATRIX_LAPACK = 0,   /* one column-major array, leading dim llm */
    PARSEC_MATRIX_TILE   = 1    /* tiles of mb*nb elements, ld = mb */
} parsec_matrix_storage_t;

/* triangle selectors; the values equal dplasma's Upper/Lower/UpperLower
 * (constants.h:82-84) because the testers pass dplasma enums here. */
#define PARSEC_MATRIX_UPPER 121
#define PARSEC_MATRIX_LOWER 122
#define PARSEC_MATRIX_FULL  123

/* descriptor kinds of this shim */
#define PARSEC_SHIM_BLOCK_CYCLIC      0x5ca1
#define PARSEC_SHIM_SYM_BLOCK_CYCLIC  0x5ca2

typedef struct parsec_context_s  parsec_context_t;
typedef struct parsec_taskpool_s parsec_taskpool_t;

typedef struct parsec_data_collection_s {
    int   nodes;          /* always 1 */
    int   myrank;         /* always 0 */
    char *key;            /* parsec_data_collection_set_key */
} parsec_data_collection_t;

typedef struct parsec_tiled_matrix_s {
    parsec_data_collection_t super;
    parsec_matrix_type_t     mtype;
    parsec_matrix_storage_t  storage;
    int    dtype;          /* PARSEC_SHIM_* descriptor kind */
    int    mb, nb;         /* tile size */
    int    bsiz;           /* mb * nb */
    int    lm, ln;         /* full matrix size */
    int    lmt, lnt;       /* full matrix size in tiles */
    int    i, j;           /* origin of the submatrix */
    int    m, n;           /* submatrix size */
    int    mt, nt;         /* submatrix size in tiles */
    int    llm, lln;       /* local sizes (== lm, ln on one rank) */
    int    slm, sln;
    int    nb_local_tiles; /* tiles stored by this rank */
} parsec_tiled_matrix_t;

typedef struct parsec_grid_2Dcyclic_s {
    int rank, rows, cols, krows, kcols, ip, jq, rrank, crank;
} parsec_grid_2Dcyclic_t;

typedef struct parsec_matrix_block_cyclic_s {
    parsec_tiled_matrix_t  super;
    parsec_grid_2Dcyclic_t grid;
    void *mat;
} parsec_matrix_block_cyclic_t;

typedef struct parsec_matrix_sym_block_cyclic_s {
    parsec_tiled_matrix_t  super;
    parsec_grid_2Dcyclic_t grid;
    void *mat;
    int   uplo;            /* PARSEC_MATRIX_LOWER or PARSEC_MATRIX_UPPER */
} parsec_matrix_sym_block_cyclic_t;

/* ---------------------------------------------------------------- runtime */
/* nb_cores is accepted for compatibility (host threads are not used for
 * compute).  The GPU is the HIP device named by the MCA-style argument
 * "--mca device_hip_device <n>" after "--", or env PARSEC_MCA_device_hip_device,
 * default 0.  Returns NULL when no GPU is usable. */
parsec_context_t *parsec_init(int nb_cores, int *pargc, char **pargv[]);
int  parsec_fini(parsec_context_t **pcontext);
int  parsec_context_add_taskpool(parsec_context_t *context, parsec_taskpool_t *tp);
int  parsec_context_start(parsec_context_t *context);
int  parsec_context_wait(parsec_context_t *context);
/* Write back device-authoritative tile mirrors and free them (the testers
 * call it before each timed run, tests/testing_zpotrf.c:61). */
void parsec_devices_release_memory(void);
void parsec_devices_reset_load(parsec_context_t *context);
void parsec_taskpool_free(parsec_taskpool_t *tp);
void parsec_taskpool_sync_ids(void);

/* ------------------------------------------------------------- descriptors */
int parsec_matrix_block_cyclic_init(parsec_matrix_block_cyclic_t *dc,
                                    parsec_matrix_type_t mtype,
                                    parsec_matrix_storage_t storage,
                                    int myrank, int mb, int nb, int lm, int ln,
                                    int i, int j, int m, int n,
                                    int P, int Q, int kp, int kq, int ip, int jq);
int parsec_matrix_sym_block_cyclic_init(parsec_matrix_sym_block_cyclic_t *dc,
                                        parsec_matrix_type_t mtype,
                                        int myrank, int mb, int nb, int lm, int ln,
                                        int i, int j, int m, int n,
                                        int P, int Q, int uplo);
void  parsec_tiled_matrix_destroy(parsec_tiled_matrix_t *tdesc);
void  parsec_data_collection_set_key(parsec_data_collection_t *dc, const char *key);
/* Host memory for local tiles (page-locked, so stage-in/out run at full
 * PCIe rate). */
void *parsec_data_allocate(size_t size);
void  parsec_data_free(void *p);
size_t parsec_datadist_getsizeoftype(parsec_matrix_type_t type);

/* Address of tile (m,n) in the descriptor's host storage, or NULL when the
 * descriptor does not store it (upper tiles of a lower symmetric one). */
void *parsec_shim_tile_addr(const parsec_tiled_matrix_t *A, int m, int n);
/* Host base pointer ("mat") of a shim descriptor. */
void *parsec_shim_mat(const parsec_tiled_matrix_t *A);
int   parsec_shim_uplo(const parsec_tiled_matrix_t *A);
/* sizeof the descriptor structs (0: tiled matrix, 1: block cyclic, 2: sym
 * block cyclic), for bindings that mirror the layouts. */
size_t parsec_shim_sizeof(int which);

#ifdef __cplusplus
}
#endif
#endif
