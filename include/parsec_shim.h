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

/* ------------------------------------------------------- task insertion (DTD) */
/* The insert-task interface of PaRSEC (interfaces/dtd/insert_function.h) as
 * tests/testing_zpotrf_dtd.c uses it, over the tile tasks of this library
 * (the task classes come from dplasma.h, parsec_dtd_create_d*_task_class).
 * A DTD taskpool binds to one symmetric tile descriptor at its first task:
 * the uplo part is staged into HBM once, every inserted task is enqueued on
 * the GPU at insertion (ordered only by the tiles it reads and writes), and
 * parsec_dtd_data_flush_all / parsec_taskpool_wait / parsec_context_wait
 * wait for the tasks and write the tiles back to host memory.
 * parsec_dtd_insert_task_with_task_class takes (flag, pointer) pairs in the
 * order of the task class's arguments, ended by PARSEC_DTD_ARG_END: values
 * by address, tiles from PARSEC_DTD_TILE_OF.  Flags are accepted and not
 * interpreted (the class fixes each argument's access mode).  Returns 0, or
 * a negative code for a task outside the tile-Cholesky pattern (printed on
 * stderr; nothing is enqueued). */
typedef struct parsec_task_class_s parsec_task_class_t;
typedef struct parsec_dtd_tile_s parsec_dtd_tile_t;
typedef uint64_t parsec_data_key_t;

#define PARSEC_DTD_ARG_END    0
#define PARSEC_DTD_EMPTY_FLAG 0
#define PARSEC_VALUE          0x00010000
#define PARSEC_SCRATCH        0x00020000
#define PASSED_BY_REF         (-1)
#define PARSEC_INPUT          0x00100000
#define PARSEC_OUTPUT         0x00200000
#define PARSEC_INOUT          (PARSEC_INPUT | PARSEC_OUTPUT)
#define PARSEC_AFFINITY       0x01000000
#define PARSEC_PUSHOUT        0x02000000
#define PARSEC_DEV_CPU        0x01
#define PARSEC_DEV_CUDA       0x02
#define PARSEC_DEV_HIP        0x04
#define PARSEC_DEV_ALL        0x7f

parsec_taskpool_t *parsec_dtd_taskpool_new(void);
int  parsec_dtd_data_collection_init(parsec_data_collection_t *dc);
void parsec_dtd_data_collection_fini(parsec_data_collection_t *dc);
/* key of tile (m, n) of a shim descriptor: m + n * lmt */
parsec_data_key_t parsec_shim_data_key(const parsec_tiled_matrix_t *A, int m, int n);
parsec_dtd_tile_t *parsec_dtd_tile_of(parsec_data_collection_t *dc, parsec_data_key_t key);
/* as in PaRSEC, DC names a descriptor variable dc for which a pointer
 * __dc<DC> is in scope */
#define PARSEC_DTD_TILE_OF(DC, I, J)                                                        \
    parsec_dtd_tile_of((parsec_data_collection_t *)(__dc##DC),                              \
                       parsec_shim_data_key((const parsec_tiled_matrix_t *)(__dc##DC), (I), (J)))
int  parsec_dtd_insert_task_with_task_class(parsec_taskpool_t *tp, parsec_task_class_t *tc, int priority,
                                            int device_type, ...);
void parsec_dtd_data_flush(parsec_taskpool_t *tp, parsec_dtd_tile_t *tile);
void parsec_dtd_data_flush_all(parsec_taskpool_t *tp, parsec_data_collection_t *dc);
int  parsec_taskpool_wait(parsec_taskpool_t *tp);
void parsec_dtd_task_class_release(parsec_taskpool_t *tp, parsec_task_class_t *tc);
/* tasks inserted into a DTD taskpool so far */
long parsec_shim_dtd_ntasks(const parsec_taskpool_t *tp);

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
