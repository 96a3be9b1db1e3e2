/*
 * dplasma_mi.h -- thin C ABI of the gfx950 (MI355X) kernel library
 * libdplasma_mi.so.  The C host code behind dplasma.h reaches HIP only
 * through these functions: plain pointers (device pointers where named d*),
 * sizes and an opaque context; every function returns 0 or a negative code
 * (mi_last_error() has the message).  Blocking functions synchronize the
 * context's streams before returning.
 *
 * Device layouts:
 *   packed lower tiles  tile (m,n), m >= n, of an MT x MT grid at tile index
 *                       n*MT - n*(n-1)/2 + (m-n); NB*NB elements per tile,
 *                       column-major, ld = NB; rows >= N - m*NB of the last
 *                       tile row are padding and never written.
 *   column-major        right-hand sides B: N x NRHS, leading dimension ldb.
 *
 * Each hot-path entry cites the reference task body or wrapper it replaces.
 *
 * Threading: a context owns its streams and every device workspace its
 * calls use (grown on demand, freed by mi_ctx_destroy), so different
 * contexts may be used at the same time from different threads.  Calls on
 * one context must not overlap (as for the reference, whose calls on one
 * parsec context are serialized by dplasma_wait_until_completion,
 * src/dplasmaaux.h:97-101).  The mi_set_* engine switches are process-wide
 * and must be set while no call is running.  mi_last_error() is per thread.
 */
#ifndef DPLASMA_MI_KERNELS_H
#define DPLASMA_MI_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mi_ctx_s mi_ctx_t;

/* ------------------------------------------------------------ context */
/* Per-device streams, the device info word and the profiler
 * (replaces the per-stream handles and the d_iinfo workspace of
 * src/dplasmaaux_hip.c:43-96 and src/zpotrf_wrapper.c:101-116,237-249). */
int  mi_ctx_create(int device, mi_ctx_t **ctx);
void mi_ctx_destroy(mi_ctx_t *ctx);
int  mi_device_count(int *count);
const char *mi_last_error(void);
int  mi_synchronize(mi_ctx_t *ctx);

/* ------------------------------------------------------------ memory */
int mi_malloc(mi_ctx_t *ctx, void **dptr, size_t bytes);
int mi_free(mi_ctx_t *ctx, void *dptr);
int mi_host_alloc(void **hptr, size_t bytes);     /* page-locked host memory */
int mi_host_free(void *hptr);
int mi_memcpy_h2d(mi_ctx_t *ctx, void *dst, const void *src, size_t bytes);
int mi_memcpy_d2h(mi_ctx_t *ctx, void *dst, const void *src, size_t bytes);
int mi_memcpy_d2d(mi_ctx_t *ctx, void *dst, const void *src, size_t bytes);
/* 2-D copies (rows of `width` bytes): the pitched stage-in/out of
 * src/cuda/lapack_cuda_stage_in.c:18-162. */
int mi_memcpy2d_h2d(mi_ctx_t *ctx, void *dst, size_t dpitch, const void *src, size_t spitch,
                    size_t width, size_t height);
int mi_memcpy2d_d2h(mi_ctx_t *ctx, void *dst, size_t dpitch, const void *src, size_t spitch,
                    size_t width, size_t height);
int mi_memset(mi_ctx_t *ctx, void *dptr, int value, size_t bytes);
/* Plain streaming copy kernel of n doubles (reads n*8 bytes, writes n*8
 * bytes): the reference point that scales HBM traffic counters. */
int mi_dcopy(mi_ctx_t *ctx, size_t n, const double *dsrc, double *ddst);

/* ------------------------------------------------------------ layout */
/* Packed upper tiles (tile (m,n), m <= n, at index n*(n+1)/2 + m) to packed
 * lower tiles holding the transposes, and back; used for uplo = Upper, where
 * U = L^T and every kernel below works on L. */
int mi_dsym_upper_to_lower(mi_ctx_t *ctx, int N, int NB, const double *dU, double *dL);
int mi_dsym_lower_to_upper(mi_ctx_t *ctx, int N, int NB, const double *dL, double *dU);
/* General tiles (MB x NB, ld MB, tile (m,n) at n*MT+m) <-> column-major. */
int mi_dtiles_to_colmajor(mi_ctx_t *ctx, int M, int N, int MB, int NB, const double *dT,
                          double *dA, int lda);
int mi_dcolmajor_to_tiles(mi_ctx_t *ctx, int M, int N, int MB, int NB, const double *dA,
                          int lda, double *dT);

/* ------------------------------------------------------------ generators */
/* dplasma_dplgsy on packed lower tiles (src/zplgsy_wrapper.c:24-46 ->
 * src/cores/core_zplgsy.c:72-150), diagonal tiles filled in full. */
int mi_dplgsy_tiles(mi_ctx_t *ctx, double bump, int N, int NB, double *dT, uint64_t seed);
/* dplasma_dplrnt into a column-major M x N block at (m0,n0) of a gM-row
 * matrix (src/zplrnt_wrapper.c:25-45 -> src/cores/core_zplrnt.c:68-91). */
int mi_dplrnt(mi_ctx_t *ctx, int M, int N, double *dA, int lda, int gM, int m0, int n0,
              uint64_t seed);
/* A(i,i) += alpha for i < n (the diagdom option of dplasma_dplrnt,
 * src/zplrnt_wrapper.c:80-93). */
int mi_dadd_diag(mi_ctx_t *ctx, int n, double *dA, int lda, double alpha);
/* SPD matrices of prescribed condition for the mixed-precision solve (no
 * counterpart in the reference, whose dplasma_dlatms, src/zlatms_wrapper.c,
 * uses a linear spectrum):
 *   mi_dgauss               n standard Gaussian values: splitmix64 words 2t,
 *                           2t+1 from seed, Box-Muller;
 *   mi_dorth_haar           Q (N x N, ld N) = orthogonal factor of the
 *                           Householder QR of the N x N Gaussian matrix of
 *                           seed (rocSOLVER dgeqrf + dorgqr, loaded at run
 *                           time);
 *   mi_dspd_from_orth_tiles A = Q diag(lambda) Q^T, lambda_j =
 *                           kappa^(-j/(N-1)), into packed lower tiles
 *                           (rocBLAS dsyrk; needs 2 N^2 doubles of scratch). */
int mi_dgauss(mi_ctx_t *ctx, int64_t n, double *dG, uint64_t seed);
int mi_dorth_haar(mi_ctx_t *ctx, int N, uint64_t seed, double *dQ);
int mi_dspd_from_orth_tiles(mi_ctx_t *ctx, int N, int NB, const double *dQ, double kappa, double *dT);

/* ------------------------------------------------------------ hot path */
/* In-place tile Cholesky, lower: the POTRF/TRSM/SYRK/GEMM task classes of
 * src/zpotrf_L.jdf:91-633 as fused per-step launches with one-panel
 * look-ahead.  For NB >= 512 (a multiple of 64) the diagonal tiles are
 * factored recursively with sub-tiles of NB / 2 (as mi_dpotrf_tiles_rec
 * with hmb = NB / 2), which is faster there.  *info receives k*NB + iinfo of
 * the first non-positive pivot (src/zpotrf_L.jdf:200-202), 0 on success.
 * Blocking. */
int mi_dpotrf_tiles(mi_ctx_t *ctx, int N, int NB, double *dT, int *info);
/* Single precision of the same factorization on FP32 tiles (the "s"
 * precision of src/zpotrf_wrapper.c:345-371): FP32 MFMA
 * (v_mfma_f32_16x16x4_f32) trailing update, FP32 diagonal factor and panel
 * solve, same info rule.  Blocking. */
int mi_spotrf_tiles(mi_ctx_t *ctx, int N, int NB, float *dT, int *info);
/* mi_dpotrf_tiles with a recursive diagonal factor: POTRF(k) of every
 * diagonal tile larger than hmb runs as a tile Cholesky of that tile with
 * hmb x hmb sub-tiles (the RECURSIVE body of src/zpotrf_L.jdf:168-192 that
 * dplasma_zpotrf_setrecursive / dplasma_zpotrf_rec select,
 * src/zpotrf_wrapper.c:48-55,425-451).  hmb is rounded down to a multiple of
 * 32 (at least 32); hmb >= NB changes nothing.  Same info.  Blocking. */
int mi_dpotrf_tiles_rec(mi_ctx_t *ctx, int N, int NB, double *dT, int hmb, int *info);
/* mi_dpotrf_tiles without any device allocation, for callers that cannot
 * allocate (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, results and
 * info as mi_dpotrf_tiles (bitwise the same factor, including its recursive
 * diagonal factor for NB >= 512).  Its scratch (the diagonal-block inverses
 * of two steps in flight and one diagonal tile's sub-tile image) is a fixed
 * 6.5 MiB device array, one per GPU, so NB <= 1024 (larger NB: -2 and an
 * error message).  Calls that overlap on one GPU race on that array,
 * whichever context they use: serialize them.  The context's profiler is not
 * used by this call.  Blocking. */
int mi_dpotrf_tiles_noalloc(mi_ctx_t *ctx, int N, int NB, double *dT, int *info);
/* Left/Lower tile triangular solve, B <- alpha op(L)^{-1} B, op = NoTrans
 * (111) or Trans (112), diag NonUnit (131) / Unit (132):
 * src/ztrsm_LLN.jdf, src/ztrsm_LLT.jdf.  Blocking. */
int mi_dtrsm_LL(mi_ctx_t *ctx, int trans, int diag, double alpha, int N, int NB,
                const double *dT, int NRHS, double *dB, int ldb);
/* B <- (L L^T)^{-1} B: src/zpotrs_wrapper.c:91-94.  Blocking. */
int mi_dpotrs_tiles(mi_ctx_t *ctx, int N, int NB, const double *dT, int NRHS, double *dB, int ldb);
/* Solve engine: 1 (default) = one wavefront launch per 16 right-hand sides
 * (workgroups of 128 rows, each handing its rows of X to the next ones
 * through uncached device memory) whenever NB % 32 == 0; 0 = one
 * diagonal-solve and one update launch per tile.  Process-wide. */
int mi_set_solve_engine(int which);
/* Bound on one hand-off wait between workgroups of the wavefront solve, in
 * seconds (default 10; 0 gives up at once).  A wait that exceeds it ends the
 * launches early: the blocking call then returns -5 and its outputs are not
 * valid.  Process-wide. */
int mi_set_solve_wait_limit(double seconds);
/* FP32 tiles and right-hand sides: the "s" precision of the two calls above. */
int mi_strsm_LL(mi_ctx_t *ctx, int trans, int diag, float alpha, int N, int NB,
                const float *dT, int NRHS, float *dB, int ldb);
int mi_spotrs_tiles(mi_ctx_t *ctx, int N, int NB, const float *dT, int NRHS, float *dB, int ldb);

/* LAPACK-layout entries: A is a column-major N x N array in device memory
 * (lda >= max(1, N)) of which only the uplo triangle (121 Upper, 122 Lower)
 * is read and written; B is column-major (ldb >= max(1, N)).  They replace
 * the reference's route to LAPACK-layout memory -- a block-cyclic descriptor
 * with PARSEC_MATRIX_LAPACK storage over the caller's array, as the
 * ScaLAPACK-style pdpotrf_w builds it
 * (src/scalapack_wrappers/dplasma_wrapper_pdpotrf.c:139-232), staged tile by
 * tile with 2-D pitched copies (src/cuda/lapack_cuda_stage_in.c:18-162) --
 * by one device staging pass each way into a packed lower tile image of
 * NB x NB tiles (allocated per call) and the tile path above.
 * mi_dpotrf_lapack: A = L L^T (Lower) or U^T U (Upper) in place, *info as
 * mi_dpotrf_tiles (the triangle is written back also when *info > 0).
 * mi_dpotrs_lapack: B <- A^{-1} B from that factor.  mi_dposv_lapack: both,
 * the solve only when *info == 0.  Blocking. */
int mi_dpotrf_lapack(mi_ctx_t *ctx, int uplo, int N, double *dA, int lda, int NB, int *info);
int mi_dpotrs_lapack(mi_ctx_t *ctx, int uplo, int N, int NRHS, const double *dA, int lda, double *dB,
                     int ldb, int NB);
int mi_dposv_lapack(mi_ctx_t *ctx, int uplo, int N, int NRHS, double *dA, int lda, double *dB, int ldb,
                    int NB, int *info);
/* The staging passes alone: the uplo triangle of column-major A to the
 * packed lower tile image dT (entries outside the triangle and beyond N
 * zero), and the lower triangle of dT back to A's uplo triangle. */
int mi_dsym_lapack_to_tiles(mi_ctx_t *ctx, int uplo, int N, const double *dA, int lda, int NB, double *dT);
int mi_dsym_tiles_to_lapack(mi_ctx_t *ctx, int uplo, int N, int NB, const double *dT, double *dA, int lda);

/* Single task-class launches (one step k of the factorization), for tests:
 *   POTRF(k)            src/zpotrf_L.jdf:91-208
 *   TRSM(k+1..MT-1, k)  src/zpotrf_L.jdf:214-343
 *   SYRK/GEMM(m,n,k) for tile columns n in [n_begin, n_end)
 *                       src/zpotrf_L.jdf:349-633 */
int mi_dpotrf_step_potrf(mi_ctx_t *ctx, int N, int NB, double *dT, int k, int *info);
/* Engine of the SYRK/GEMM trailing update: 0 = FP64 VALU FMA, 1 = FP64
 * MFMA 16x16x4 (default; two resident waves per SIMD keep it at full rate).
 * Process-wide. */
int mi_set_update_engine(int which);
/* Engine of POTRF(k) and the panel TRSM: 0 = VALU kernels, 1 = MFMA panel
 * updates with register/LDS diagonal blocks (default).  Process-wide. */
int mi_set_panel_engine(int which);
int mi_dpotrf_step_trsm(mi_ctx_t *ctx, int N, int NB, double *dT, int k);
int mi_dpotrf_step_update(mi_ctx_t *ctx, int N, int NB, double *dT, int k, int n_begin, int n_end);
/* The same three launches on FP32 tiles (the engines do not apply: FP32 is
 * always MFMA). */
int mi_spotrf_step_potrf(mi_ctx_t *ctx, int N, int NB, float *dT, int k, int *info);
int mi_spotrf_step_trsm(mi_ctx_t *ctx, int N, int NB, float *dT, int k);
int mi_spotrf_step_update(mi_ctx_t *ctx, int N, int NB, float *dT, int k, int n_begin, int n_end);
/* Engine of the FP32 trailing update: 0 = FP32 MFMA (default), 1 = BF16
 * MFMA with FP32 accumulation on a two-term bf16 split of both operands
 * (three products, ~16 significant bits).  Process-wide. */
int mi_set_spotrf_update(int which);
/* Where the BF16-split updates of an FP32 factor split their operands:
 * 1 = each panel column once per step, into hi/lo BF16 images the main
 * stream's updates load directly (default), 0 = every update workgroup
 * splits the FP32 panel tiles it reads.  Both give the same factor bit for
 * bit.  Process-wide. */
int mi_set_bf16_panel_split(int which);
/* Kernel of the FP64 residual R = B - A X (mi_dsymm_resid and the
 * refinement of mi_dsposv_tiles) for 64 <= NB <= 512: 1 = whole 16-column
 * chunks of each tile loaded straight into LDS, for NB a multiple of 128
 * (default), 0 = operands loaded into registers.  Both read every stored
 * element once; they sum in different orders.  Process-wide. */
int mi_set_symm_kernel(int which);

/* ------------------------------------------------------------ task insertion */
/* Task-insertion engine behind the DTD front-end of dplasma.h (the tile
 * tasks of src/dtd_wrappers/z{potrf,trsm,herk,gemm}.c as
 * tests/testing_zpotrf_dtd.c:91-302 inserts them).  Tasks work on a packed
 * lower FP64 tile image dT (N, NB) in device memory; each is enqueued when
 * inserted and waits only for the earlier tasks whose tiles it reads or
 * writes (read-after-write, write-after-write, write-after-read), so any
 * insertion order that is valid sequentially gives the sequential result.
 *   POTRF(k)      A(k,k) = chol(A(k,k)); info as mi_dpotrf_tiles
 *   TRSM(m,k)     A(m,k) <- A(m,k) A(k,k)^-T                 (m > k)
 *   SYRK(k,m)     A(m,m) <- A(m,m) - A(m,k) A(m,k)^T          (m > k)
 *   GEMM(m,n,k)   A(m,n) <- A(m,n) - A(m,k) A(n,k)^T          (m > n > k)
 * mi_dtd_wait blocks until every inserted task has run and returns the
 * info word and the number of tasks; the engine can take further tasks. */
typedef struct mi_dtd_s mi_dtd_t;
int  mi_dtd_create(mi_ctx_t *ctx, int N, int NB, double *dT, mi_dtd_t **dtd);
int  mi_dtd_potrf(mi_dtd_t *dtd, int k);
int  mi_dtd_trsm(mi_dtd_t *dtd, int m, int k);
int  mi_dtd_syrk(mi_dtd_t *dtd, int k, int m);
int  mi_dtd_gemm(mi_dtd_t *dtd, int m, int n, int k);
int  mi_dtd_wait(mi_dtd_t *dtd, int *info, long *ntasks);
void mi_dtd_destroy(mi_dtd_t *dtd);

/* ------------------------------------------------------------ mixed precision */
/* Mixed-precision solve of A X = B (no counterpart in the reference: the
 * precision growth path of its dposv, src/zposv_wrapper.c:63-113).  A: FP64
 * lower tiles of an SPD matrix, overwritten by its FP64 factor only when
 * *iter < 0; B, X: N x NRHS column-major.  The factor is computed in FP32
 * (prec 0: FP32 MFMA; prec 1: trailing updates on BF16 MFMA with the
 * two-term split) and the solution refined in FP64 as LAPACK dsposv does:
 * *iter >= 0 refinement steps on success; -2 overflow when rounding A,
 * -3 FP32 factorization failed, -4 the residual stalled, -31 thirty steps
 * without convergence -- each negative value followed by the FP64
 * factorization and solve (then *info as mi_dpotrf_tiles).  Blocking. */
int mi_dsposv_tiles(mi_ctx_t *ctx, int N, int NB, double *dA, int NRHS, const double *dB, int ldb,
                    double *dX, int ldx, int prec, int *iter, int *info);
/* R = B - A X in FP64 for the symmetric matrix A held in lower tiles (the
 * residual of src/dplasma_zcheck.c:179-224 and of the refinement). */
int mi_dsymm_resid(mi_ctx_t *ctx, int N, int NB, const double *dA, int NRHS, const double *dX, int ldx,
                   const double *dB, int ldb, double *dR, int ldr);
/* Round n FP64 values to FP32; *overflow = 1 when a finite value exceeds
 * FLT_MAX (LAPACK dlag2s).  Blocking. */
int mi_dlag2s_tiles(mi_ctx_t *ctx, int64_t n, const double *dS, float *dD, int *overflow);

/* ------------------------------------------------------------ differentiation */
/* Reverse-mode derivative of the tile Cholesky (no counterpart in the
 * reference: the reverse of its right-looking factorization,
 * src/zpotrf_L.jdf:91-633).  dL: the factor L of A = L L^T as returned by
 * mi_dpotrf_tiles (lower tiles; strict upper parts of the diagonal tiles and
 * padding rows are not read).  dLbar: on entry the gradient of a scalar f
 * with respect to L (lower tiles, same layout); on exit the symmetric
 * gradient Abar with respect to A (df = sum_ij Abar_ij dA_ij, lower tiles;
 * the strict upper parts of diagonal tiles and padding are not written).
 * Blocking; returns 0 or a negative error. */
int mi_dpotrf_adjoint_tiles(mi_ctx_t *ctx, int N, int NB, const double *dL, double *dLbar);
int mi_spotrf_adjoint_tiles(mi_ctx_t *ctx, int N, int NB, const float *dL, float *dLbar);
/* A^-1 in place of its Cholesky factor (lower tiles), as the adjoint of
 * logdet A: the inverse step of dplasma_zpotri (src/zpotri_wrapper.c:54-82,
 * trtri + lauum there).  *info = i > 0 if L(i,i) == 0 (dT unchanged).
 * Blocking. */
int mi_dpotri_tiles(mi_ctx_t *ctx, int N, int NB, double *dT, int *info);

/* ------------------------------------------------------------ batched */
/* Many independent SPD systems of uneven orders n_i in [1, 1024], each
 * factored and solved (NRHS = 1) as LAPACK dpotrf + dpotrs would (no
 * counterpart in the reference: its batching growth path, each matrix being
 * the one-tile case of dplasma_dposv, src/zposv_wrapper.c:63-113 with
 * src/cores/core_zpotrf.c:68-74).  Matrix i is column-major, lda = n_i, at
 * element offset sum_{j<i} n_j^2 of dA; its right-hand side at offset
 * sum_{j<i} n_j of dB (dB may be NULL: factor only).  One persistent launch
 * works through the batch; with use_graph the launch is captured in a
 * hipGraph on the first call for given (dA, dB) and replayed after. */
typedef struct mi_batch_s mi_batch_t;
/* n[i] = lo + (splitmix64 word i of seed) mod (hi - lo + 1) */
int mi_batch_sizes(uint64_t seed, int count, int lo, int hi, int *n);
/* Test batch: A_i = G_i^T G_i / n_i + I (lower triangles), G_i Gaussian
 * (consecutive elements of the mi_dgauss stream of seed, G_0 first);
 * rocBLAS dsyrk loaded at run time.  Blocking. */
int mi_dbatch_spd(mi_ctx_t *ctx, int count, const int *n, uint64_t seed, double *dA);
int mi_batch_create(mi_ctx_t *ctx, int count, const int *n, mi_batch_t **plan);
int mi_batch_destroy(mi_batch_t *plan);
int mi_batch_extent(const mi_batch_t *plan, int64_t *a_elems, int64_t *b_elems);
/* tile shape: matrices with n <= nsplit use 64-row strips, the others
 * 128-row strips; wg_per_cu = 1 or 2 persistent workgroups per CU */
int mi_batch_set_shape(mi_batch_t *plan, int nsplit, int wg_per_cu);
int mi_batch_get_shape(const mi_batch_t *plan, int *nsplit, int *wg_per_cu);
/* Times each candidate shape on a copy of (dA0, dB0) in (dA, dB) and keeps
 * the fastest; *choice = its index.  dA/dB are overwritten. */
int mi_dbatch_autotune(mi_batch_t *plan, const double *dA0, const double *dB0, double *dA, double *dB,
                       int *choice);
int mi_sbatch_autotune(mi_batch_t *plan, const float *dA0, const float *dB0, float *dA, float *dB, int *choice);
/* Factor (and solve when dB != NULL) in place; blocking. */
int mi_dposv_batched(mi_batch_t *plan, double *dA, double *dB, int use_graph);
int mi_sposv_batched(mi_batch_t *plan, float *dA, float *dB, int use_graph);
/* The same work as one launch per matrix (the launch-loop baseline). */
int mi_dposv_batched_loop(mi_batch_t *plan, double *dA, double *dB);
int mi_sposv_batched_loop(mi_batch_t *plan, float *dA, float *dB);
/* info[i] of the last run: 0, or the 1-based column of matrix i whose pivot
 * was not positive (then no solve for it). */
int mi_batch_info(const mi_batch_t *plan, int *info);
/* kernel time of the last run (HIP events around the launch), ms */
int mi_batch_last_ms(const mi_batch_t *plan, double *ms);

/* ------------------------------------------------------------ checks */
/* check_dpoinv norms (src/dplasma_zcheck.c:264-314): ||A0||_1, ||Ainv||_1
 * of symmetric lower tiles and ||I - Ainv A0||_1.  Blocking. */
int mi_dcheck_poinv(mi_ctx_t *ctx, int N, int NB, const double *dA0, const double *dAinv,
                    double *Anorm, double *Ainvnorm, double *Rnorm);
/* Norms and residuals behind check_dpotrf / check_daxmb
 * (src/dplasma_zcheck.c:63-133,179-224) computed on the GPU:
 *   *Anorm = ||A0||_oo (symmetric, lower tiles), *Rnorm = ||L L^T - A0||_oo. */
int mi_dcheck_potrf(mi_ctx_t *ctx, int N, int NB, const double *dL, const double *dA0,
                    double *Anorm, double *Rnorm);
/* ||A0||_oo, ||b||_oo, ||x||_oo and ||b - A0 x||_oo for column-major b, x. */
int mi_dcheck_axmb(mi_ctx_t *ctx, int N, int NB, const double *dA0, int NRHS,
                   const double *db, int ldb, const double *dx, int ldx,
                   double *Anorm, double *Bnorm, double *Xnorm, double *Rnorm);

/* ------------------------------------------------------------ profiler */
/* When enabled, every kernel launch is bracketed by HIP events recorded on
 * the stream it runs on; mi_prof_get aggregates per kernel name after the
 * next synchronization. */
int mi_prof_enable(mi_ctx_t *ctx, int on);
int mi_prof_reset(mi_ctx_t *ctx);
int mi_prof_get(mi_ctx_t *ctx, const char *kernel, int *launches, double *total_ms,
                double *flops, double *bytes);
/* Time during which at least one launch of `kernel` (of any recorded
 * kernel when NULL) was running, ms: the union of the launch intervals
 * within each blocking call, summed over the calls -- launches overlapping
 * on different streams count once.  A name ending in '*' selects every
 * kernel name with that prefix. */
int mi_prof_busy_ms(mi_ctx_t *ctx, const char *kernel, double *busy_ms);
/* Number of distinct kernel names recorded, and the i-th name. */
int mi_prof_count(mi_ctx_t *ctx);
const char *mi_prof_name(mi_ctx_t *ctx, int i);
/* Span of the last blocking hot-path call (first launch to last), ms. */
int mi_last_span_ms(mi_ctx_t *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif
