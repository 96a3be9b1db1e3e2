/*
 * oracle.h -- CPU restatement of DPLASMA's tile Cholesky factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only the pytest suite under tests/, the smoke() check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * liboracle.so, and only as the checker (or as the timed CPU baseline).
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the DPLASMA source tree).  Parity pins:
 *   - orc_rnd64_jump         : against the reference's own random.h compiled
 *                              from its sources (oracle/_ref), and against
 *                              brute-force LCG stepping in the tests.
 *   - orc_dplgsy_* / dplrnt  : exact float32 rounding recomputed in Python
 *                              with rational arithmetic; symmetry, bounds,
 *                              diagonal-bump invariants; committed fixtures.
 *   - orc_flops_*            : against the reference's flops.h compiled from
 *                              its sources (oracle/_ref) and closed forms.
 *   - tile kernels           : brute force on tiny inputs, A = L0 L0^T with
 *                              small-integer L0 (exact), identity / diagonal
 *                              special cases, numpy.linalg at small sizes.
 *   - orc_check_*            : hand-computed small cases, planted errors.
 *   - DAG executor           : equals the sequential right-looking loop on the
 *                              same kernels (same per-tile update order, so
 *                              bitwise equal), info cases.
 */
#ifndef DPLASMA_ORACLE_H
#define DPLASMA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dplasma enum values, src/include/dplasma/constants.h:78-96 */
#define ORC_NOTRANS   111
#define ORC_TRANS     112
#define ORC_UPPER     121
#define ORC_LOWER     122
#define ORC_UPPERLOWER 123
#define ORC_NONUNIT   131
#define ORC_UNIT      132
#define ORC_LEFT      141
#define ORC_RIGHT     142

/* ---------------- generators (src/cores/random.h, core_zplgsy.c, core_zplrnt.c) */
uint64_t orc_rnd64_jump(uint64_t n, uint64_t seed);
float    orc_rnd_to_entry(uint64_t ran);
void orc_dplgsy_tile(double bump, int m, int n, double *A, int lda,
                     int gM, int m0, int n0, uint64_t seed);
void orc_dplrnt_tile(int m, int n, double *A, int lda,
                     int gM, int m0, int n0, uint64_t seed);
void orc_splgsy_tile(float bump, int m, int n, float *A, int lda,
                     int gM, int m0, int n0, uint64_t seed);
/* full column-major (LAPACK layout) helpers built from the tile generators */
void orc_dplgsy_full(double bump, int N, double *A, int lda, uint64_t seed);
void orc_dplrnt_full(int M, int N, double *A, int lda, uint64_t seed);
/* tile-storage helpers: the symmetric descriptor stores the uplo tiles only,
 * column by column (tile (m,n) of a lower descriptor at index
 * n*MT - n*(n-1)/2 + (m-n)), each tile MB*NB elements with ld = MB. */
int64_t orc_sym_tile_index(int uplo, int MT, int m, int n);
void orc_dplgsy_symtiles(double bump, int uplo, int N, int NB, double *T, uint64_t seed);
void orc_dplrnt_tiles(int M, int N, int MB, int NB, double *T, uint64_t seed);
void orc_symtiles_to_full(int uplo, int N, int NB, const double *T, double *A, int lda);
void orc_full_to_symtiles(int uplo, int N, int NB, const double *A, int lda, double *T);
void orc_tiles_to_full(int M, int N, int MB, int NB, const double *T, double *A, int lda);
void orc_full_to_tiles(int M, int N, int MB, int NB, const double *A, int lda, double *T);

/* ---------------- flop counts (src/flops.h:86-92,215,231,249,272,282) */
double orc_flops_dpotrf(double n);
double orc_flops_dpotrs(double n, double nrhs);
double orc_flops_dgemm(double m, double n, double k);
double orc_flops_dsyrk(double k, double n);
double orc_flops_dtrsm(int side, double m, double n);

/* ---------------- FP64 tile kernels (src/cores/core_z{potrf,trsm,syrk,gemm}.c) */
int  orc_dpotrf(int uplo, int n, double *A, int lda);             /* returns LAPACK info */
void orc_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha,
               const double *A, int lda, double *B, int ldb);
void orc_dsyrk_ln(int n, int k, double alpha, const double *A, int lda,
                  double beta, double *C, int ldc);
void orc_dgemm(int transa, int transb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc);

/* Optional optimized BLAS (function pointers resolved with dlopen/dlsym from a
 * CBLAS/LAPACKE library such as the scipy-bundled OpenBLAS).  Returns 0 when
 * loaded.  prefix: symbol prefix ("scipy_" for scipy's OpenBLAS). */
int  orc_blas_load(const char *path, const char *prefix);
int  orc_blas_loaded(void);
void orc_blas_set_threads(int n);

/* ---------------- tile DAG executor (src/zpotrf_L.jdf, src/zpotrf_wrapper.c) */
/* Factor the lower symmetric tile matrix T (layout of orc_sym_tile_index) in
 * place with the four task classes POTRF/TRSM/SYRK/GEMM, nthreads workers,
 * per-tile updates in ascending k and the JDF's priorities.  use_blas selects
 * the loaded BLAS for the tile kernels (else the plain C kernels).
 * pri_change < 0 means the default (MT).  Returns info (0, or k*NB+iinfo). */
int orc_dag_dpotrf_L(int N, int NB, double *T, int nthreads, int use_blas, int pri_change);
/* Same DAG restricted to one thread and run in program order (right-looking
 * loop); used to pin the executor: both apply identical per-tile update
 * sequences, so the results are bitwise equal. */
int orc_seq_dpotrf_L(int N, int NB, double *T, int use_blas);
/* Tile solve: B (M x NRHS, tiles MB x NB, ld MB) <- A^{-1} B with A = L L^T held
 * in lower symmetric tiles (src/zpotrs_wrapper.c:91-94 -> ztrsm_LLN.jdf,
 * ztrsm_LLT.jdf). */
void orc_tile_dpotrs_L(int N, int NB, const double *T, int NRHS, double *B);
/* Tile triangular solve Left/Lower, trans in {NoTrans, Trans}, diag, alpha. */
void orc_tile_dtrsm_LL(int trans, int diag, double alpha, int N, int NB,
                       const double *T, int NRHS, double *B);

/* ---------------- checks (src/dplasma_zcheck.c:63-133,179-224) */
double orc_dlamch_eps(void);
double orc_dlansy_inf(int uplo, int N, const double *A, int lda);
double orc_dlange_inf(int M, int N, const double *A, int lda);
/* ||A0 - L L^T||_oo / (||A0||_oo N eps); L read from the uplo triangle of A
 * (full LAPACK layout), A0 symmetric given by its uplo triangle. */
double orc_check_dpotrf(int uplo, int N, const double *A, int lda,
                        const double *A0, int lda0, double *Anorm, double *Rnorm);
/* ||b - A0 x||_oo / ((||A0||_oo ||x||_oo + ||b||_oo) N eps) */
double orc_check_dpoinv(int uplo, int N, const double *A, int lda, const double *Ainv, int ldi,
                        double *Anorm_out, double *Ainvnorm_out, double *Rnorm_out);
double orc_check_daxmb(int uplo, int N, int NRHS, const double *A0, int lda0,
                       const double *b, int ldb, const double *x, int ldx);

/* wall clock in seconds (gettimeofday semantics of tests/common_timing.h:20-28) */
double orc_wtime(void);

#ifdef __cplusplus
}
#endif
#endif
