/*
 * Driver of the sanitizer build (ASan + UBSan) of the C host library: the
 * host-only logic that needs no GPU.  Descriptor geometry and tile
 * addressing for every storage and uplo, host tile memory, argument errors
 * of the _New constructors, task-class creation and release, the error path
 * of parsec_init without a usable GPU, and (when a GPU is visible) one small
 * dposv with its residual check.  Exit 1 on a wrong result.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "dplasma.h"

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); fails++; } } while (0)

int main(int argc, char **argv)
{
    /* symmetric descriptors, both uplo, ragged last tile */
    for (int u = 0; u < 2; u++) {
        const int uplo = u ? dplasmaUpper : dplasmaLower;
        parsec_matrix_sym_block_cyclic_t d;
        parsec_matrix_sym_block_cyclic_init(&d, PARSEC_MATRIX_DOUBLE, 0, 93, 93, 378, 378, 0, 0, 378, 378, 1, 1, uplo);
        EXPECT(d.super.mt == 5 && d.super.nb_local_tiles == 15, "sym geometry");
        d.mat = parsec_data_allocate((size_t)d.super.nb_local_tiles * d.super.bsiz * sizeof(double));
        EXPECT(d.mat != NULL, "parsec_data_allocate");
        double *base = (double *)d.mat;
        for (int n = 0; n < 5; n++)
            for (int m = 0; m < 5; m++) {
                double *t = parsec_shim_tile_addr((parsec_tiled_matrix_t *)&d, m, n);
                const int stored = uplo == dplasmaLower ? m >= n : m <= n;
                EXPECT(!t == !stored, "tile (%d,%d) presence", m, n);
                if (t) {
                    EXPECT(t >= base && t + 93 * 93 <= base + d.super.nb_local_tiles * 93 * 93, "tile in range");
                    t[0] = m + 10.0 * n;   /* every stored tile writable */
                }
            }
        parsec_data_free(d.mat);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&d);
    }
    /* general descriptors, tile and LAPACK storage */
    for (int st = 0; st < 2; st++) {
        parsec_matrix_block_cyclic_t g;
        parsec_matrix_block_cyclic_init(&g, PARSEC_MATRIX_DOUBLE, st ? PARSEC_MATRIX_LAPACK : PARSEC_MATRIX_TILE, 0,
                                        64, 64, 200, 3, 0, 0, 200, 3, 1, 1, 1, 1, 0, 0);
        EXPECT(g.super.mt == 4 && g.super.nt == 1, "general geometry");
        g.mat = parsec_data_allocate((size_t)g.super.nb_local_tiles * g.super.bsiz * sizeof(double));
        for (int m = 0; m < 4; m++) {
            double *t = parsec_shim_tile_addr((parsec_tiled_matrix_t *)&g, m, 0);
            EXPECT(t != NULL, "general tile %d", m);
            if (t) t[0] = m;
        }
        parsec_data_free(g.mat);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&g);
    }
    /* constructors: argument errors */
    {
        parsec_matrix_sym_block_cyclic_t d;
        parsec_matrix_sym_block_cyclic_init(&d, PARSEC_MATRIX_DOUBLE, 0, 64, 64, 128, 128, 0, 0, 128, 128, 1, 1,
                                            dplasmaLower);
        int info = -7;
        parsec_taskpool_t *tp = dplasma_dpotrf_New(dplasmaUpperLower, (parsec_tiled_matrix_t *)&d, &info);
        EXPECT(tp == NULL, "dpotrf_New accepted uplo UpperLower");
        tp = dplasma_dpotrf_New(dplasmaLower, (parsec_tiled_matrix_t *)&d, &info);
        EXPECT(tp != NULL && info == 0, "dpotrf_New");
        if (tp) {
            dplasma_dpotrf_setrecursive(tp, 128);
            dplasma_dpotrf_Destruct(tp);
        }
        tp = dplasma_dtrsm_New(dplasmaRight, dplasmaLower, dplasmaNoTrans, dplasmaNonUnit, 1.0,
                               (parsec_tiled_matrix_t *)&d, (parsec_tiled_matrix_t *)&d);
        EXPECT(tp == NULL, "dtrsm_New accepted side Right");
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&d);
    }
    /* DTD task classes without a context */
    {
        parsec_taskpool_t *tp = parsec_dtd_taskpool_new();
        EXPECT(tp != NULL, "dtd taskpool");
        parsec_task_class_t *tc[4] = {
            parsec_dtd_create_dpotrf_task_class(tp, 0, PARSEC_DEV_ALL),
            parsec_dtd_create_dtrsm_task_class(tp, 0, PARSEC_DEV_ALL),
            parsec_dtd_create_dsyrk_task_class(tp, 0, PARSEC_DEV_ALL),
            parsec_dtd_create_dgemm_task_class(tp, 0, PARSEC_DEV_ALL)};
        for (int i = 0; i < 4; i++) {
            EXPECT(tc[i] != NULL, "task class %d", i);
            parsec_dtd_task_class_release(tp, tc[i]);
        }
        EXPECT(parsec_shim_dtd_ntasks(tp) == 0, "ntasks");
        parsec_taskpool_free(tp);
    }
    /* the context: NULL without a usable GPU, else one small dposv */
    parsec_context_t *ctx = parsec_init(1, &argc, &argv);
    const int gpu = ctx != NULL;
    if (ctx) {
        const int N = 378, NB = 93;
        parsec_matrix_sym_block_cyclic_t A, A0;
        parsec_matrix_block_cyclic_t B, X;
        parsec_matrix_sym_block_cyclic_init(&A, PARSEC_MATRIX_DOUBLE, 0, NB, NB, N, N, 0, 0, N, N, 1, 1, dplasmaLower);
        parsec_matrix_sym_block_cyclic_init(&A0, PARSEC_MATRIX_DOUBLE, 0, NB, NB, N, N, 0, 0, N, N, 1, 1, dplasmaLower);
        parsec_matrix_block_cyclic_init(&B, PARSEC_MATRIX_DOUBLE, PARSEC_MATRIX_TILE, 0, NB, NB, N, 2, 0, 0, N, 2,
                                        1, 1, 1, 1, 0, 0);
        parsec_matrix_block_cyclic_init(&X, PARSEC_MATRIX_DOUBLE, PARSEC_MATRIX_TILE, 0, NB, NB, N, 2, 0, 0, N, 2,
                                        1, 1, 1, 1, 0, 0);
        A.mat = parsec_data_allocate((size_t)A.super.nb_local_tiles * A.super.bsiz * sizeof(double));
        A0.mat = parsec_data_allocate((size_t)A0.super.nb_local_tiles * A0.super.bsiz * sizeof(double));
        B.mat = parsec_data_allocate((size_t)B.super.nb_local_tiles * B.super.bsiz * sizeof(double));
        X.mat = parsec_data_allocate((size_t)X.super.nb_local_tiles * X.super.bsiz * sizeof(double));
        dplasma_dplgsy(ctx, (double)N, dplasmaLower, (parsec_tiled_matrix_t *)&A, 3872);
        dplasma_dplgsy(ctx, (double)N, dplasmaLower, (parsec_tiled_matrix_t *)&A0, 3872);
        dplasma_dplrnt(ctx, 0, (parsec_tiled_matrix_t *)&B, 3873);
        dplasma_dlacpy(ctx, dplasmaUpperLower, (parsec_tiled_matrix_t *)&B, (parsec_tiled_matrix_t *)&X);
        EXPECT(dplasma_dposv(ctx, dplasmaLower, (parsec_tiled_matrix_t *)&A, (parsec_tiled_matrix_t *)&X) == 0,
               "dposv");
        EXPECT(check_daxmb(ctx, 0, dplasmaLower, (parsec_tiled_matrix_t *)&A0, (parsec_tiled_matrix_t *)&B,
                           (parsec_tiled_matrix_t *)&X) == 0, "check_daxmb");
        parsec_data_free(A.mat); parsec_data_free(A0.mat); parsec_data_free(B.mat); parsec_data_free(X.mat);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&A);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&A0);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&B);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&X);
        parsec_fini(&ctx);
    }
    if (fails) { fprintf(stderr, "%d failures\n", fails); return 1; }
    printf("host driver: ok (%s)\n", gpu ? "with GPU" : "no GPU: host logic only");
    return 0;
}
