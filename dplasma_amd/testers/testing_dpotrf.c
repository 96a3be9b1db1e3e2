/*
 * testing_dpotrf.c -- the DPLASMA Cholesky tester (tests/testing_zpotrf.c,
 * d precision) driven through this library's drop-in entry points on one
 * MI355X.  Same options (tests/common.c:174-268: -N, -t/--MB, -T/--NB,
 * -K/--NRHS, -u/--uplo, -x/--check, -v, --seed, --nruns, -z/--HMB; -g and -c
 * are accepted), same flow: generate A = dplgsy(N, seed), release device
 * copies, time the factorization from start to wait (the GFLOP/s print of
 * tests/common.h:260-286), then check_dpotrf, solve with B = dplrnt(seed+1)
 * and check_daxmb.  With --resident the matrix is advised onto the device
 * first and the timed region excludes host<->device staging.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "dplasma.h"

static double now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

/* LAWN-41 count, src/flops.h:86-87,272 */
static double flops_dpotrf(double n)
{
    return n * (((1. / 6.) * n + 0.5) * n + (1. / 3.)) + n * (((1. / 6.) * n) * n - (1. / 6.));
}

int main(int argc, char **argv)
{
    int N = 0, MB = 180, NB = 180, NRHS = 1, uplo = dplasmaLower, check = 0, loud = 1;
    int nruns = 1, hmb = 0, resident = 0;
    unsigned long long seed = 3872;
    static struct option opts[] = {
        {"N", required_argument, 0, 'N'},     {"t", required_argument, 0, 't'},
        {"MB", required_argument, 0, 't'},    {"T", required_argument, 0, 'T'},
        {"NB", required_argument, 0, 'T'},    {"K", required_argument, 0, 'K'},
        {"NRHS", required_argument, 0, 'K'},  {"u", required_argument, 0, 'u'},
        {"uplo", required_argument, 0, 'u'},  {"x", no_argument, 0, 'x'},
        {"check", no_argument, 0, 'x'},       {"v", optional_argument, 0, 'v'},
        {"seed", required_argument, 0, 'R'},  {"nruns", required_argument, 0, '3'},
        {"z", required_argument, 0, 'z'},     {"HMB", required_argument, 0, 'z'},
        {"g", required_argument, 0, 'g'},     {"gpus", required_argument, 0, 'g'},
        {"c", required_argument, 0, 'c'},     {"cores", required_argument, 0, 'c'},
        {"resident", no_argument, 0, 'r'},    {0, 0, 0, 0}};
    int c, nbset = 0;
    while ((c = getopt_long_only(argc, argv, "", opts, NULL)) != -1) {
        switch (c) {
        case 'N': N = atoi(optarg); break;
        case 't': MB = atoi(optarg); if (!nbset) NB = MB; break;
        case 'T': NB = atoi(optarg); nbset = 1; break;
        case 'K': NRHS = atoi(optarg); break;
        case 'u': uplo = (optarg[0] == 'u' || optarg[0] == 'U') ? dplasmaUpper : dplasmaLower; break;
        case 'x': check = 1; break;
        case 'v': loud = optarg ? atoi(optarg) : 2; break;
        case 'R': seed = strtoull(optarg, NULL, 10); break;
        case '3': nruns = atoi(optarg); break;
        case 'z': hmb = atoi(optarg); break;
        case 'r': resident = 1; break;
        default: break;
        }
    }
    if (N <= 0) { fprintf(stderr, "usage: %s -N n -t nb [-x] [-v] [--nruns k]\n", argv[0]); return 2; }
    if (MB != NB) NB = MB;   /* square tiles */

    parsec_context_t *parsec = parsec_init(1, &argc, &argv);
    if (!parsec) return 3;
    int ret = 0, info = 0;
    double flops = flops_dpotrf(N);

    parsec_matrix_sym_block_cyclic_t dcA;
    parsec_matrix_sym_block_cyclic_init(&dcA, PARSEC_MATRIX_DOUBLE, 0, MB, NB, N, N, 0, 0, N, N, 1, 1, uplo);
    dcA.mat = parsec_data_allocate((size_t)dcA.super.nb_local_tiles * dcA.super.bsiz * sizeof(double));

    for (int t = 0; t < nruns; t++) {
        dplasma_dplgsy(parsec, (double)N, uplo, (parsec_tiled_matrix_t *)&dcA, seed);
        parsec_devices_release_memory();
        if (resident) dplasma_advise_data_on_device(parsec, uplo, (parsec_tiled_matrix_t *)&dcA, NULL, NULL);
        double t0 = now();
        parsec_taskpool_t *tp = dplasma_dpotrf_New(uplo, (parsec_tiled_matrix_t *)&dcA, &info);
        if (hmb > 0) dplasma_dpotrf_setrecursive(tp, hmb);
        parsec_context_add_taskpool(parsec, tp);
        double t1 = now();
        parsec_context_start(parsec);
        int rc = parsec_context_wait(parsec);
        double t2 = now();
        dplasma_dpotrf_Destruct(tp);
        double t3 = now();
        if (rc) { fprintf(stderr, "dpotrf failed: %d\n", rc); ret |= 1; break; }
        dplasma_mi_timing_t tm;
        dplasma_mi_last_timing(parsec, &tm);
        printf("[****] TIME(s) %12.5f : dpotrf\tPxQxg= %3d %-3d %d NB= %4d N= %7d : %14f gflops"
               " - ENQ&PROG&DEST %12.5f : %14f gflops - ENQ %12.5f - DEST %12.5f"
               " - GPU %10.3f ms (stage in %8.3f, out %8.3f)\n",
               t2 - t1, 1, 1, 1, NB, N, flops / 1e9 / (t2 - t1), t3 - t0, flops / 1e9 / (t3 - t0), t1 - t0,
               t3 - t2, tm.total_ms, tm.stage_in_ms, tm.stage_out_ms);
        fflush(stdout);
        parsec_devices_release_memory();
    }
    if (info != 0) {
        printf("-- Factorization is suspicious (info = %d) ! \n", info);
        ret |= 1;
    }
    if (!info && check) {
        parsec_matrix_sym_block_cyclic_t dcA0;
        parsec_matrix_sym_block_cyclic_init(&dcA0, PARSEC_MATRIX_DOUBLE, 0, MB, NB, N, N, 0, 0, N, N, 1, 1, uplo);
        dcA0.mat = parsec_data_allocate((size_t)dcA0.super.nb_local_tiles * dcA0.super.bsiz * sizeof(double));
        dplasma_dplgsy(parsec, (double)N, uplo, (parsec_tiled_matrix_t *)&dcA0, seed);
        ret |= check_dpotrf(parsec, loud > 2 ? loud : 3, uplo, (parsec_tiled_matrix_t *)&dcA,
                            (parsec_tiled_matrix_t *)&dcA0);

        parsec_matrix_block_cyclic_t dcB, dcX;
        parsec_matrix_block_cyclic_init(&dcB, PARSEC_MATRIX_DOUBLE, PARSEC_MATRIX_TILE, 0, MB, NB, N, NRHS, 0, 0,
                                        N, NRHS, 1, 1, 1, 1, 0, 0);
        dcB.mat = parsec_data_allocate((size_t)dcB.super.nb_local_tiles * dcB.super.bsiz * sizeof(double));
        parsec_matrix_block_cyclic_init(&dcX, PARSEC_MATRIX_DOUBLE, PARSEC_MATRIX_TILE, 0, MB, NB, N, NRHS, 0, 0,
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
    parsec_data_free(dcA.mat);
    parsec_tiled_matrix_destroy((parsec_tiled_matrix_t *)&dcA);
    parsec_fini(&parsec);
    return ret;
}
