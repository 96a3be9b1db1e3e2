/*
 * Driver of the sanitizer builds of the oracle (ASan + UBSan, and TSan for
 * the pthread tile-DAG executor).  It exercises every oracle routine the
 * tests rely on at the reference's ctest sizes, ragged tiles included, and
 * fails (exit 1) on a wrong result; the sanitizers fail it on a memory,
 * undefined-behaviour or data-race error.
 *   oracle_driver [threads]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); fails++; } } while (0)

static void one(int N, int NB, int threads)
{
    const int MT = (N + NB - 1) / NB;
    const size_t ntile = (size_t)MT * (MT + 1) / 2 * NB * NB;
    double *T = malloc(ntile * sizeof *T), *T2 = malloc(ntile * sizeof *T2);
    double *A0 = malloc((size_t)N * N * sizeof *A0), *A = malloc((size_t)N * N * sizeof *A);
    double *b = malloc((size_t)N * 2 * sizeof *b), *x = malloc((size_t)N * 2 * sizeof *x);
    /* generators: tile image == whole matrix, both uplo */
    for (int u = 0; u < 2; u++) {
        const int uplo = u ? ORC_UPPER : ORC_LOWER;
        memset(T, 0, ntile * sizeof *T);
        orc_dplgsy_symtiles((double)N, uplo, N, NB, T, 3872);
        orc_dplgsy_full((double)N, N, A0, N, 3872);
        orc_symtiles_to_full(uplo, N, NB, T, A, N);
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++)
                if ((uplo == ORC_LOWER) ? i >= j : i <= j)
                    EXPECT(A[(size_t)j * N + i] == A0[(size_t)j * N + i], "dplgsy tile/full differ at %d,%d", i, j);
    }
    /* factor: threaded DAG == sequential loop (bitwise), checks pass */
    orc_dplgsy_symtiles((double)N, ORC_LOWER, N, NB, T, 3872);
    memcpy(T2, T, ntile * sizeof *T);
    int i1 = orc_dag_dpotrf_L(N, NB, T, threads, 0, -1);
    int i2 = orc_seq_dpotrf_L(N, NB, T2, 0);
    EXPECT(i1 == 0 && i2 == 0, "info %d %d", i1, i2);
    EXPECT(memcmp(T, T2, ntile * sizeof *T) == 0, "DAG and sequential factors differ (N=%d NB=%d)", N, NB);
    orc_symtiles_to_full(ORC_LOWER, N, NB, T, A, N);
    double an, rn;
    double r = orc_check_dpotrf(ORC_LOWER, N, A, N, A0, N, &an, &rn);
    EXPECT(r < 60.0, "check_dpotrf %g", r);
    /* solve two right-hand sides (B in NB x NB tiles, one tile column) */
    double *Bt = malloc((size_t)MT * NB * NB * sizeof *Bt);
    orc_dplrnt_full(N, 2, b, N, 3873);
    orc_full_to_tiles(N, 2, NB, NB, b, N, Bt);
    orc_tile_dpotrs_L(N, NB, T, 2, Bt);
    orc_tiles_to_full(N, 2, NB, NB, Bt, x, N);
    free(Bt);
    r = orc_check_daxmb(ORC_LOWER, N, 2, A0, N, b, N, x, N);
    EXPECT(r < 60.0, "check_daxmb %g", r);
    /* info of a non-SPD matrix: k*NB + j of the first bad pivot */
    orc_dplgsy_full((double)N, N, A0, N, 3872);
    const int bad = N / 2;
    A0[(size_t)bad * N + bad] = -2.0 * N;
    orc_full_to_symtiles(ORC_LOWER, N, NB, A0, N, T);
    i1 = orc_dag_dpotrf_L(N, NB, T, threads, 0, -1);
    EXPECT(i1 == bad + 1, "info %d, expected %d", i1, bad + 1);
    free(T); free(T2); free(A0); free(A); free(b); free(x);
}

int main(int argc, char **argv)
{
    const int threads = argc > 1 ? atoi(argv[1]) : 4;
    one(378, 93, threads);
    one(700, 128, threads);
    one(256, 256, threads);
    EXPECT(orc_flops_dpotrf(4096.0) == 22914881536.0, "flops");
    if (fails) { fprintf(stderr, "%d failures\n", fails); return 1; }
    printf("oracle driver: ok (%d threads)\n", threads);
    return 0;
}
