/*
 * dplasma.h -- DPLASMA's tile Cholesky factor-and-solve entry points, served
 * by hand-written gfx950 (MI355X) kernels with the tiles resident in HBM.
 *
 * Drop-in for the reference prototypes (src/include/dplasma/dplasma_z.h after
 * the z->d / z->s precision substitution) on a single rank with one GPU.
 * Each prototype cites the reference declaration it replaces.  Return values
 * and argument checks follow the reference wrappers (cited at each function
 * in the implementation); documented deviations:
 *   - dplasma_dpotrf reports info > 0 for a non-SPD matrix also when the
 *     factorization ran on the GPU (the reference's HIP/CUDA bodies drop the
 *     device info, src/zpotrf_L.jdf:135,163);
 *   - runtime failures (a HIP error, out of HBM) return a negative code
 *     instead of aborting the run (src/zpotrf_L.jdf:164,284,411,567).
 */
#ifndef DPLASMA_MI_DPLASMA_H
#define DPLASMA_MI_DPLASMA_H

#include "parsec_shim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- constants (src/include/dplasma/constants.h:31-99) */
#define DPLASMA_SUCCESS                 0
#define DPLASMA_ERR_NOT_INITIALIZED  -101
#define DPLASMA_ERR_NOT_SUPPORTED    -103
#define DPLASMA_ERR_ILLEGAL_VALUE    -104
#define DPLASMA_ERR_OUT_OF_RESOURCES -106

typedef int dplasma_enum_t;

#define dplasmaNoTrans       111
#define dplasmaTrans         112
#define dplasmaConjTrans     113
#define dplasmaUpper         121
#define dplasmaLower         122
#define dplasmaUpperLower    123
#define dplasmaNonUnit       131
#define dplasmaUnit          132
#define dplasmaLeft          141
#define dplasmaRight         142
#define dplasmaOneNorm       171
#define dplasmaInfNorm       175
#define dplasmaMaxNorm       177

/* ---------------- Cholesky (dplasma_z.h:81-84,172,239,273) */
int dplasma_dpotrf(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A);
int dplasma_dpotrf_rec(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A, int hmb);
parsec_taskpool_t *dplasma_dpotrf_New(dplasma_enum_t uplo, parsec_tiled_matrix_t *A, int *INFO);
void dplasma_dpotrf_Destruct(parsec_taskpool_t *o);
void dplasma_dpotrf_setrecursive(parsec_taskpool_t *o, int hmb);
int dplasma_dpotrs(parsec_context_t *parsec, dplasma_enum_t uplo,
                   const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B);
int dplasma_dposv(parsec_context_t *parsec, dplasma_enum_t uplo,
                  parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B);

/* ---------------- inverse from the Cholesky factor (dplasma_z.h:79,85) */
int dplasma_dpotri(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A);
int dplasma_dpoinv(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A);

/* ---------------- mixed precision (extension: the precision growth path of
 * dplasma_dposv, src/zposv_wrapper.c:63-113; no counterpart in the reference).
 * Solves A X = B for SPD A (FP64 descriptors) with a low-precision factor
 * refined in FP64, with LAPACK dsposv semantics: B unchanged; A unchanged
 * when *iter >= 0 (refinement steps taken), overwritten by its FP64 factor
 * when *iter < 0 (-2 overflow rounding A, -3 low-precision factorization
 * failed, -4 refinement stalled, -31 no convergence in 30 steps).  Returns
 * info as dplasma_dposv (or -1 for an illegal uplo / mismatched B, X). */
#define DPLASMA_MIXED_FP32   0   /* FP32 factor on FP32 MFMA */
#define DPLASMA_MIXED_BF16X3 1   /* FP32 factor, trailing updates on BF16 MFMA, two-term split */
int dplasma_dsposv(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                   parsec_tiled_matrix_t *B, parsec_tiled_matrix_t *X, int *iter);
int dplasma_dposv_mixed(parsec_context_t *parsec, dplasma_enum_t uplo, parsec_tiled_matrix_t *A,
                        parsec_tiled_matrix_t *B, parsec_tiled_matrix_t *X, int prec, int *iter);

/* ---------------- differentiation (extension: the reverse of dplasma_dpotrf,
 * src/zpotrf_L.jdf:91-633; no counterpart in the reference).  L: the factor
 * dplasma_dpotrf returned (uplo part); Lbar: df/dL on entry (uplo part, the
 * transpose convention of U = L^T for Upper), the symmetric df/dA on exit
 * (uplo part; df = sum_ij Abar_ij dA_ij).  Returns 0, -1 (uplo), -2 (shape). */
int dplasma_dpotrf_adjoint(parsec_context_t *parsec, dplasma_enum_t uplo, const parsec_tiled_matrix_t *L,
                           parsec_tiled_matrix_t *Lbar);

/* ---------------- DTD task classes (src/dtd_wrappers/dplasma_z_dtd.h:60-63;
 * zherk -> dsyrk).  Arguments of each class, in insertion order:
 *   dpotrf  uplo, m, A(k,k) [inout], lda, info [scratch: int *]
 *   dtrsm   side, uplo, transA, diag, m, n, alpha, A [in], lda, B [inout], ldb
 *   dsyrk   uplo, trans, n, k, alpha, A [in], lda, beta, C [inout], ldc
 *   dgemm   transA, transB, m, n, k, alpha, A [in], lda, B [in], ldb, beta,
 *           C [inout], ldc
 * (values by address; alpha/beta are double).  Supported: the tasks of the
 * tile Cholesky of a symmetric tile descriptor as
 * tests/testing_zpotrf_dtd.c:124-302 inserts them for both uplo (alpha -1,
 * beta 1 for the updates, alpha 1 for TRSM).  tile_full and devices are
 * accepted and ignored: every task runs on the GPU. */
parsec_task_class_t *parsec_dtd_create_dpotrf_task_class(parsec_taskpool_t *dtd_tp, int tile_full, int devices);
parsec_task_class_t *parsec_dtd_create_dtrsm_task_class(parsec_taskpool_t *dtd_tp, int tile_full, int devices);
parsec_task_class_t *parsec_dtd_create_dsyrk_task_class(parsec_taskpool_t *dtd_tp, int tile_full, int devices);
parsec_task_class_t *parsec_dtd_create_dgemm_task_class(parsec_taskpool_t *dtd_tp, int tile_full, int devices);

/* ---------------- triangular solve (dplasma_z.h:51,153,221) */
int dplasma_dtrsm(parsec_context_t *parsec, dplasma_enum_t side, dplasma_enum_t uplo,
                  dplasma_enum_t trans, dplasma_enum_t diag, double alpha,
                  const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B);
parsec_taskpool_t *dplasma_dtrsm_New(dplasma_enum_t side, dplasma_enum_t uplo,
                                     dplasma_enum_t trans, dplasma_enum_t diag, double alpha,
                                     const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B);
void dplasma_dtrsm_Destruct(parsec_taskpool_t *o);

/* ---------------- generators and copies (dplasma_z.h:97,107-109,193-195,260-262) */
int dplasma_dplgsy(parsec_context_t *parsec, double bump, dplasma_enum_t uplo,
                   parsec_tiled_matrix_t *A, unsigned long long int seed);
int dplasma_dplghe(parsec_context_t *parsec, double bump, dplasma_enum_t uplo,
                   parsec_tiled_matrix_t *A, unsigned long long int seed);
int dplasma_dplrnt(parsec_context_t *parsec, int diagdom, parsec_tiled_matrix_t *A,
                   unsigned long long int seed);
parsec_taskpool_t *dplasma_dplgsy_New(double bump, dplasma_enum_t uplo,
                                      parsec_tiled_matrix_t *A, unsigned long long int seed);
void dplasma_dplgsy_Destruct(parsec_taskpool_t *o);
int dplasma_dlacpy(parsec_context_t *parsec, dplasma_enum_t uplo,
                   const parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *B);

/* ---------------- checks (dplasma_z.h:279-280); computed on the GPU */
int check_dpotrf(parsec_context_t *parsec, int loud, dplasma_enum_t uplo,
                 parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *A0);
int check_daxmb(parsec_context_t *parsec, int loud, dplasma_enum_t uplo,
                parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *b, parsec_tiled_matrix_t *x);
/* src/dplasma_zcheck.c:264-314; A and Ainv symmetric descriptors (uplo part) */
int check_dpoinv(parsec_context_t *parsec, int loud, dplasma_enum_t uplo,
                 parsec_tiled_matrix_t *A, parsec_tiled_matrix_t *Ainv);

/* ---------------- data placement (src/dplasmaaux.h:147-151) */
/* Makes the uplo part of A resident in HBM until parsec_devices_release_memory()
 * or dplasma_data_release(): operations on A then work on the device copy and
 * skip the host<->device staging; the host copy is refreshed by
 * dplasma_data_sync_host() and by the release calls.  The operation/op_args
 * parameters of the reference (a per-tile device map for multi-GPU) are
 * accepted and ignored: one GPU holds every tile. */
typedef void *parsec_tiled_matrix_unary_op_t;
int dplasma_advise_data_on_device(parsec_context_t *parsec, dplasma_enum_t uplo,
                                  parsec_tiled_matrix_t *A,
                                  parsec_tiled_matrix_unary_op_t operation, void *op_args);

/* ================================================================
 * Extensions of this library (no counterpart in the reference)
 * ================================================================ */

/* Copies a resident matrix back to its host storage (no-op otherwise). */
int dplasma_data_sync_host(parsec_context_t *parsec, parsec_tiled_matrix_t *A);
/* Writes back (if needed) and drops the device copy of A. */
int dplasma_data_release(parsec_context_t *parsec, parsec_tiled_matrix_t *A);

/* Kernel-level timing of the last blocking call on this context, in ms
 * (GPU events on the streams the kernels ran on). */
typedef struct dplasma_mi_timing_s {
    double total_ms;      /* first kernel start -> last kernel end */
    double stage_in_ms;   /* host -> device copies */
    double stage_out_ms;  /* device -> host copies */
} dplasma_mi_timing_t;
int dplasma_mi_last_timing(parsec_context_t *parsec, dplasma_mi_timing_t *t);

/* Norms behind the last check_dpotrf / check_daxmb / check_dpoinv on this
 * context (the quantities src/dplasma_zcheck.c:100-104,193-202,290-300
 * prints), so that a caller can compare them, not only the pass/fail
 * answer.  Unused fields are 0. */
typedef struct dplasma_mi_check_s {
    double Anorm;      /* ||A0||_oo (check_dpotrf, check_daxmb), ||A||_1 (check_dpoinv) */
    double Rnorm;      /* ||L L^T - A0||_oo, ||b - A0 x||_oo, ||I - Ainv A||_1 */
    double Bnorm;      /* ||b||_oo (check_daxmb) */
    double Xnorm;      /* ||x||_oo (check_daxmb) */
    double Ainvnorm;   /* ||Ainv||_1 (check_dpoinv) */
    double result;     /* the ratio compared with the threshold */
} dplasma_mi_check_t;
int dplasma_mi_last_check(parsec_context_t *parsec, dplasma_mi_check_t *c);

/* The kernel-library context behind a parsec context (for the profiler of
 * dplasma_mi.h). */
struct mi_ctx_s *dplasma_mi_context(parsec_context_t *parsec);

#ifdef __cplusplus
}
#endif
#endif
