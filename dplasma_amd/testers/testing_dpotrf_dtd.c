/*
 * testing_dpotrf_dtd.c -- the DTD Cholesky tester (tests/testing_zpotrf_dtd.c,
 * d precision) on this library: the tile Cholesky written as inserted tasks
 * (POTRF, TRSM, SYRK, GEMM per tile, the reference's priorities) through the
 * PaRSEC DTD calls of parsec_shim.h and the task classes of dplasma.h, then
 * check_dpotrf, dpotrs with B = dplrnt(seed+1) and check_daxmb.
 * Options: -N, -t/--MB, -K/--NRHS, -u/--uplo L|U, -x/--check, -v, --seed.
 * The reference's ctest case is "potrf_dtd -N 874 -K 367 -t 76"
 * (tests/Testings.cmake:148).  Exit status 0 when every check passes.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>
#include "dplasma.h"

static double now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

static double flops_dpotrf(double n)
{
    return n * (((1. / 6.) * n + 0.5) * n + (1. / 3.)) + n * (((1. / 6.) * n) * n - (1. / 6.));
}

int main(int argc, char **argv)
{
    int N = 0, NB = 180, NRHS = 1, uplo = dplasmaLower, check = 0, loud = 1;
    unsigned long long seed = 3872;
    static struct option opts[] = {
        {"N", required_argument, 0, 'N'},    {"t", required_argument, 0, 't'},
        {"MB", required_argument, 0, 't'},   {"K", required_argument, 0, 'K'},
        {"NRHS", required_argument, 0, 'K'}, {"u", required_argument, 0, 'u'},
        {"uplo", required_argument, 0, 'u'}, {"x", no_argument, 0, 'x'},
        {"check", no_argument, 0, 'x'},      {"v", optional_argument, 0, 'v'},
        {"seed", required_argument, 0, 'R'}, {0, 0, 0, 0}};
    int c;
    while ((c = getopt_long_only(argc, argv, "", opts, NULL)) != -1) {
        switch (c) {
        case 'N': N = atoi(optarg); break;
        case 't': NB = atoi(optarg); break;
        case 'K': NRHS = atoi(optarg); break;
        case 'u': uplo = (optarg[0] == 'u' || optarg[0] == 'U') ? dplasmaUpper : dplasmaLower; break;
        case 'x': check = 1; break;
        case 'v': loud = optarg ? atoi(optarg) : 2; break;
        case 'R': seed = strtoull(optarg, NULL, 10); break;
        default: break;
        }
    }
    if (N <= 0 || NB <= 0) { fprintf(stderr, "usage: %s -N n -t nb [-K nrhs] [-u L|U] [-x]\n", argv[0]); return 2; }
    parsec_context_t *parsec = parsec_init(1, &argc, &argv);
    if (!parsec) return 3;
    int ret = 0, info = 0;

    parsec_matrix_sym_block_cyclic_t dcA;
    parsec_matrix_sym_block_cyclic_init(&dcA, PARSEC_MATRIX_DOUBLE, 0, NB, NB, N, N, 0, 0, N, N, 1, 1, uplo);
    dcA.mat = parsec_data_allocate((size_t)dcA.super.nb_local_tiles * dcA.super.bsiz * sizeof(double));
    parsec_matrix_sym_block_cyclic_t *__dcA = &dcA;
    parsec_dtd_data_collection_init((parsec_data_collection_t *)&dcA);
    dplasma_dplgsy(parsec, (double)N, uplo, (parsec_tiled_matrix_t *)&dcA, seed);

    parsec_taskpool_t *dtd_tp = parsec_dtd_taskpool_new();
    parsec_context_add_taskpool(parsec, dtd_tp);
    double t0 = now();
    parsec_context_start(parsec);
    parsec_task_class_t *potrf_tc = parsec_dtd_create_dpotrf_task_class(dtd_tp, 0, PARSEC_DEV_ALL);
    parsec_task_class_t *trsm_tc = parsec_dtd_create_dtrsm_task_class(dtd_tp, 0, PARSEC_DEV_ALL);
    parsec_task_class_t *syrk_tc = parsec_dtd_create_dsyrk_task_class(dtd_tp, 0, PARSEC_DEV_ALL);
    parsec_task_class_t *gemm_tc = parsec_dtd_create_dgemm_task_class(dtd_tp, 0, PARSEC_DEV_ALL);

    const int total = dcA.super.mt, mb = dcA.super.mb;
    const double one = 1.0, mone = -1.0;
    int side, transA_p, transA_g, transB, trans, diag = dplasmaNonUnit;
    if (uplo == dplasmaLower) {
        side = dplasmaRight; transA_p = dplasmaTrans; trans = dplasmaNoTrans;
        transA_g = dplasmaNoTrans; transB = dplasmaTrans;
    } else {
        side = dplasmaLeft; transA_p = dplasmaTrans; trans = dplasmaTrans;
        transA_g = dplasmaTrans; transB = dplasmaNoTrans;
    }
    int rc = 0;
    /* tile rows / leading dimensions of the task arguments: one array each,
     * so that every argument address stays valid while tasks are queued */
    int *tdim = (int *)malloc(sizeof(int) * total), ld = mb;
    for (int k = 0; k < total; k++) tdim[k] = (k == total - 1) ? N - k * mb : mb;
    for (int k = 0; k < total && !rc; k++) {
        rc |= parsec_dtd_insert_task_with_task_class(
            dtd_tp, potrf_tc, (total - k) * (total - k) * (total - k), PARSEC_DEV_ALL,
            PARSEC_DTD_EMPTY_FLAG, &uplo, PARSEC_DTD_EMPTY_FLAG, &tdim[k],
            PARSEC_PUSHOUT, PARSEC_DTD_TILE_OF(A, k, k), PARSEC_DTD_EMPTY_FLAG, &ld,
            PARSEC_DTD_EMPTY_FLAG, &info, PARSEC_DTD_ARG_END);
        for (int m = k + 1; m < total; m++) {
            const int lo = uplo == dplasmaLower;
            rc |= parsec_dtd_insert_task_with_task_class(
                dtd_tp, trsm_tc, (total - m) * (total - m) * (total - m) + 3 * ((2 * total) - k - m - 1) * (m - k),
                PARSEC_DEV_ALL, PARSEC_DTD_EMPTY_FLAG, &side, PARSEC_DTD_EMPTY_FLAG, &uplo,
                PARSEC_DTD_EMPTY_FLAG, &transA_p, PARSEC_DTD_EMPTY_FLAG, &diag,
                PARSEC_DTD_EMPTY_FLAG, lo ? &tdim[m] : &dcA.super.nb, PARSEC_DTD_EMPTY_FLAG,
                lo ? &dcA.super.nb : &tdim[m], PARSEC_DTD_EMPTY_FLAG, &one,
                PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, k, k), PARSEC_DTD_EMPTY_FLAG, &ld,
                PARSEC_PUSHOUT, lo ? PARSEC_DTD_TILE_OF(A, m, k) : PARSEC_DTD_TILE_OF(A, k, m),
                PARSEC_DTD_EMPTY_FLAG, &ld, PARSEC_DTD_ARG_END);
        }
        parsec_dtd_data_flush(dtd_tp, PARSEC_DTD_TILE_OF(A, k, k));
        for (int m = k + 1; m < total; m++) {
            const int lo = uplo == dplasmaLower;
            rc |= parsec_dtd_insert_task_with_task_class(
                dtd_tp, syrk_tc, (total - m) * (total - m) * (total - m) + 3 * (m - k), PARSEC_DEV_ALL,
                PARSEC_DTD_EMPTY_FLAG, &uplo, PARSEC_DTD_EMPTY_FLAG, &trans,
                PARSEC_DTD_EMPTY_FLAG, &tdim[m], PARSEC_DTD_EMPTY_FLAG, &dcA.super.mb,
                PARSEC_DTD_EMPTY_FLAG, &mone,
                PARSEC_DTD_EMPTY_FLAG, lo ? PARSEC_DTD_TILE_OF(A, m, k) : PARSEC_DTD_TILE_OF(A, k, m),
                PARSEC_DTD_EMPTY_FLAG, &ld, PARSEC_DTD_EMPTY_FLAG, &one,
                PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, m, m), PARSEC_DTD_EMPTY_FLAG, &ld,
                PARSEC_DTD_ARG_END);
            for (int n = m + 1; n < total; n++) {
                const int pri = (total - m) * (total - m) * (total - m) + 3 * ((2 * total) - m - n - 3) * (m - n) +
                                6 * (m - k);
                if (lo)
                    rc |= parsec_dtd_insert_task_with_task_class(
                        dtd_tp, gemm_tc, pri, PARSEC_DEV_ALL, PARSEC_DTD_EMPTY_FLAG, &transA_g,
                        PARSEC_DTD_EMPTY_FLAG, &transB, PARSEC_DTD_EMPTY_FLAG, &tdim[n],
                        PARSEC_DTD_EMPTY_FLAG, &dcA.super.mb, PARSEC_DTD_EMPTY_FLAG, &dcA.super.mb,
                        PARSEC_DTD_EMPTY_FLAG, &mone, PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, n, k),
                        PARSEC_DTD_EMPTY_FLAG, &ld, PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, m, k),
                        PARSEC_DTD_EMPTY_FLAG, &ld, PARSEC_DTD_EMPTY_FLAG, &one,
                        PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, n, m), PARSEC_DTD_EMPTY_FLAG, &ld,
                        PARSEC_DTD_ARG_END);
                else
                    rc |= parsec_dtd_insert_task_with_task_class(
                        dtd_tp, gemm_tc, pri, PARSEC_DEV_ALL, PARSEC_DTD_EMPTY_FLAG, &transA_g,
                        PARSEC_DTD_EMPTY_FLAG, &transB, PARSEC_DTD_EMPTY_FLAG, &dcA.super.mb,
                        PARSEC_DTD_EMPTY_FLAG, &tdim[n], PARSEC_DTD_EMPTY_FLAG, &dcA.super.mb,
                        PARSEC_DTD_EMPTY_FLAG, &mone, PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, k, m),
                        PARSEC_DTD_EMPTY_FLAG, &ld, PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, k, n),
                        PARSEC_DTD_EMPTY_FLAG, &ld, PARSEC_DTD_EMPTY_FLAG, &one,
                        PARSEC_DTD_EMPTY_FLAG, PARSEC_DTD_TILE_OF(A, m, n), PARSEC_DTD_EMPTY_FLAG, &ld,
                        PARSEC_DTD_ARG_END);
            }
            parsec_dtd_data_flush(dtd_tp, lo ? PARSEC_DTD_TILE_OF(A, m, k) : PARSEC_DTD_TILE_OF(A, k, m));
        }
    }
    parsec_dtd_data_flush_all(dtd_tp, (parsec_data_collection_t *)&dcA);
    rc |= parsec_taskpool_wait(dtd_tp);
    rc |= parsec_context_wait(parsec);
    double t1 = now();
    const long ntasks = parsec_shim_dtd_ntasks(dtd_tp);
    printf("[****] TIME(s) %12.5f : dpotrf_dtd\tPxQ= %3d %-3d NB= %4d N= %7d : %14f gflops (%ld tasks)\n",
           t1 - t0, 1, 1, NB, N, flops_dpotrf(N) / 1e9 / (t1 - t0), ntasks);
    parsec_dtd_task_class_release(dtd_tp, potrf_tc);
    parsec_dtd_task_class_release(dtd_tp, trsm_tc);
    parsec_dtd_task_class_release(dtd_tp, syrk_tc);
    parsec_dtd_task_class_release(dtd_tp, gemm_tc);
    parsec_taskpool_free(dtd_tp);
    free(tdim);
    if (rc) {
        printf("-- task insertion or execution failed (%d)\n", rc);
        ret |= 1;
    }
    if (info != 0) {
        printf("-- Factorization is suspicious (info = %d) ! \n", info);
        ret |= 1;
    }
    if (!info && !rc && check) {
        parsec_matrix_sym_block_cyclic_t dcA0;
        parsec_matrix_sym_block_cyclic_init(&dcA0, PARSEC_MATRIX_DOUBLE, 0, NB, NB, N, N, 0, 0, N, N, 1, 1, uplo);
        dcA0.mat = parsec_data_allocate((size_t)dcA0.super.nb_local_tiles * dcA0.super.bsiz * sizeof(double));
        dplasma_dplgsy(parsec, (double)N, uplo, (parsec_tiled_matrix_t *)&dcA0, seed);
        ret |= check_dpotrf(parsec, loud > 2 ? loud : 3, uplo, (parsec_tiled_matrix_t *)&dcA,
                            (parsec_tiled_matrix_t *)&dcA0);
        parsec_matrix_block_cyclic_t dcB, dcX;
        parsec_matrix_block_cyclic_init(&dcB, PARSEC_MATRIX_DOUBLE, PARSEC_MATRIX_TILE, 0, NB, NB, N, NRHS, 0, 0,
                                        N, NRHS, 1, 1, 1, 1, 0, 0);
        dcB.mat = parsec_data_allocate((size_t)dcB.super.nb_local_tiles * dcB.super.bsiz * sizeof(double));
        parsec_matrix_block_cyclic_init(&dcX, PARSEC_MATRIX_DOUBLE, PARSEC_MATRIX_TILE, 0, NB, NB, N, NRHS, 0, 0,
                                        N, NRHS, 1, 1, 1, 1, 0, 0);
        dcX.mat = parsec_data_allocate((size_t)dcX.super.nb_local_tiles * dcX.super.bsiz * sizeof(double));
        dplasma_dplrnt(parsec, 0, (parsec_tiled_matrix_t *)&dcB, seed + 1);
        dplasma_dlacpy(parsec, dplasmaUpperLower, (parsec_tiled_matrix_t *)&dcB, (parsec_tiled_matrix_t *)&dcX);
        dplasma_dpotrs(parsec, uplo, (parsec_tiled_matrix_t *)&dcA, (parsec_tiled_matrix_t *)&dcX);
        ret |= check_daxmb(parsec, loud > 2 ? loud : 3, uplo, (parsec_tiled_matrix_t *)&dcA0,
                           (parsec_tiled_matrix_t *)&dcB, (parsec_tiled_matrix_t *)&dcX);
        parsec_data_free(dcA0.mat);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&dcA0);
        parsec_data_free(dcB.mat);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&dcB);
        parsec_data_free(dcX.mat);
        parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&dcX);
    }
    parsec_dtd_data_collection_fini((parsec_data_collection_t *)&dcA);
    parsec_data_free(dcA.mat);
    parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&dcA);
    parsec_fini(&parsec);
    return ret;
}
