/*
 * generators.c -- test-matrix generators of the oracle (test infrastructure).
 *
 * Restates src/cores/random.h:19-41 (64-bit LCG with jump-ahead),
 * src/cores/core_zplgsy.c:72-150 and src/cores/core_zplrnt.c:68-91 for the
 * real (d/s) precisions (NBELEM = 1): entry (i,j) of the global matrix, i >= j,
 * is 0.5f - (float)X_{i + j*gM} * 2^-64 evaluated in single precision, with
 * X_0 = seed and X_{t+1} = A*X_t + C (mod 2^64).  Upper entries mirror lower
 * ones; the bump is added to the diagonal in the target precision.
 */
#include <string.h>
#include "oracle.h"

#define ORC_RND_A 6364136223846793005ULL
#define ORC_RND_C 1ULL

/* X_n given X_0 = seed, by binary powering of the affine map x -> a x + c.
 * random.h:23-38. */
uint64_t orc_rnd64_jump(uint64_t n, uint64_t seed)
{
    uint64_t a = ORC_RND_A, c = ORC_RND_C, x = seed;
    while (n) {
        if (n & 1ULL) x = a * x + c;
        c = c * (a + 1ULL);
        a = a * a;
        n >>= 1;
    }
    return x;
}

/* 0.5f - ran * RndF_Mul in float arithmetic (core_zplgsy.c:89); RndF_Mul is
 * 5.4210108624275222e-20f == 2^-64 exactly. */
float orc_rnd_to_entry(uint64_t ran)
{
    volatile float r = (float)ran;         /* round-to-nearest u64 -> f32 */
    volatile float s = r * 5.4210108624275222e-20f;
    return 0.5f - s;
}

static inline uint64_t next(uint64_t x) { return ORC_RND_A * x + ORC_RND_C; }

/* core_zplgsy.c:72-150 (real case). */
void orc_dplgsy_tile(double bump, int m, int n, double *A, int lda,
                     int gM, int m0, int n0, uint64_t seed)
{
    if (m0 == n0) {
        for (int j = 0; j < n; j++) {
            uint64_t ran = orc_rnd64_jump((uint64_t)(m0 + j) + (uint64_t)(n0 + j) * (uint64_t)gM, seed);
            for (int i = j; i < m; i++) {
                A[(int64_t)j * lda + i] = (double)orc_rnd_to_entry(ran);
                ran = next(ran);
            }
        }
        for (int j = 0; j < n && j < m; j++) {
            A[(int64_t)j * lda + j] += bump;
            for (int i = 0; i < j; i++)
                A[(int64_t)j * lda + i] = A[(int64_t)i * lda + j];
        }
    } else if (m0 > n0) {
        for (int j = 0; j < n; j++) {
            uint64_t ran = orc_rnd64_jump((uint64_t)m0 + (uint64_t)(n0 + j) * (uint64_t)gM, seed);
            for (int i = 0; i < m; i++) {
                A[(int64_t)j * lda + i] = (double)orc_rnd_to_entry(ran);
                ran = next(ran);
            }
        }
    } else {
        /* upper tile: the entry (m0+i, n0+j) mirrors (n0+j, m0+i) */
        for (int i = 0; i < m; i++) {
            uint64_t ran = orc_rnd64_jump((uint64_t)n0 + (uint64_t)(m0 + i) * (uint64_t)gM, seed);
            for (int j = 0; j < n; j++) {
                A[(int64_t)j * lda + i] = (double)orc_rnd_to_entry(ran);
                ran = next(ran);
            }
        }
    }
}

void orc_splgsy_tile(float bump, int m, int n, float *A, int lda,
                     int gM, int m0, int n0, uint64_t seed)
{
    if (m0 == n0) {
        for (int j = 0; j < n; j++) {
            uint64_t ran = orc_rnd64_jump((uint64_t)(m0 + j) + (uint64_t)(n0 + j) * (uint64_t)gM, seed);
            for (int i = j; i < m; i++) {
                A[(int64_t)j * lda + i] = orc_rnd_to_entry(ran);
                ran = next(ran);
            }
        }
        for (int j = 0; j < n && j < m; j++) {
            A[(int64_t)j * lda + j] += bump;
            for (int i = 0; i < j; i++)
                A[(int64_t)j * lda + i] = A[(int64_t)i * lda + j];
        }
    } else if (m0 > n0) {
        for (int j = 0; j < n; j++) {
            uint64_t ran = orc_rnd64_jump((uint64_t)m0 + (uint64_t)(n0 + j) * (uint64_t)gM, seed);
            for (int i = 0; i < m; i++) {
                A[(int64_t)j * lda + i] = orc_rnd_to_entry(ran);
                ran = next(ran);
            }
        }
    } else {
        for (int i = 0; i < m; i++) {
            uint64_t ran = orc_rnd64_jump((uint64_t)n0 + (uint64_t)(m0 + i) * (uint64_t)gM, seed);
            for (int j = 0; j < n; j++) {
                A[(int64_t)j * lda + i] = orc_rnd_to_entry(ran);
                ran = next(ran);
            }
        }
    }
}

/* core_zplrnt.c:68-91 (real case): entry (m0+i, n0+j) = f(X_{m0+i+(n0+j)*gM}). */
void orc_dplrnt_tile(int m, int n, double *A, int lda,
                     int gM, int m0, int n0, uint64_t seed)
{
    for (int j = 0; j < n; j++) {
        uint64_t ran = orc_rnd64_jump((uint64_t)m0 + (uint64_t)(n0 + j) * (uint64_t)gM, seed);
        for (int i = 0; i < m; i++) {
            A[(int64_t)j * lda + i] = (double)orc_rnd_to_entry(ran);
            ran = next(ran);
        }
    }
}

/* Full symmetric N x N matrix, both triangles, column-major. */
void orc_dplgsy_full(double bump, int N, double *A, int lda, uint64_t seed)
{
    /* one call with a single N x N "tile" at (0,0) follows the diagonal-tile
     * branch, which fills the lower triangle and mirrors it. */
    orc_dplgsy_tile(bump, N, N, A, lda, N, 0, 0, seed);
}

void orc_dplrnt_full(int M, int N, double *A, int lda, uint64_t seed)
{
    orc_dplrnt_tile(M, N, A, lda, M, 0, 0, seed);
}

/* Position of tile (m,n) in the symmetric descriptor's local storage (one
 * rank): lower tiles stored column by column, column n holding tiles n..MT-1;
 * upper tiles column by column, column n holding tiles 0..n. */
int64_t orc_sym_tile_index(int uplo, int MT, int m, int n)
{
    if (uplo == ORC_LOWER)
        return (int64_t)n * MT - ((int64_t)n * (n - 1)) / 2 + (m - n);
    return ((int64_t)n * (n + 1)) / 2 + m;
}

/* dplasma_dplgsy on a symmetric tile descriptor: one CORE_dplgsy call per
 * stored tile (src/zplgsy_wrapper.c:24-46), with tempmm/tempnn for the ragged
 * last row/column of tiles and ld = MB. */
void orc_dplgsy_symtiles(double bump, int uplo, int N, int NB, double *T, uint64_t seed)
{
    int MT = (N + NB - 1) / NB;
    for (int n = 0; n < MT; n++) {
        int m_lo = (uplo == ORC_LOWER) ? n : 0;
        int m_hi = (uplo == ORC_LOWER) ? MT : n + 1;
        for (int m = m_lo; m < m_hi; m++) {
            int tm = (m == MT - 1) ? N - m * NB : NB;
            int tn = (n == MT - 1) ? N - n * NB : NB;
            double *t = T + orc_sym_tile_index(uplo, MT, m, n) * (int64_t)NB * NB;
            orc_dplgsy_tile(bump, tm, tn, t, NB, N, m * NB, n * NB, seed);
        }
    }
}

void orc_dplrnt_tiles(int M, int N, int MB, int NB, double *T, uint64_t seed)
{
    int MT = (M + MB - 1) / MB, NT = (N + NB - 1) / NB;
    for (int n = 0; n < NT; n++)
        for (int m = 0; m < MT; m++) {
            int tm = (m == MT - 1) ? M - m * MB : MB;
            int tn = (n == NT - 1) ? N - n * NB : NB;
            double *t = T + ((int64_t)n * MT + m) * (int64_t)MB * NB;
            orc_dplrnt_tile(tm, tn, t, MB, M, m * MB, n * NB, seed);
        }
}

void orc_symtiles_to_full(int uplo, int N, int NB, const double *T, double *A, int lda)
{
    int MT = (N + NB - 1) / NB;
    for (int n = 0; n < MT; n++) {
        int m_lo = (uplo == ORC_LOWER) ? n : 0;
        int m_hi = (uplo == ORC_LOWER) ? MT : n + 1;
        for (int m = m_lo; m < m_hi; m++) {
            int tm = (m == MT - 1) ? N - m * NB : NB;
            int tn = (n == MT - 1) ? N - n * NB : NB;
            const double *t = T + orc_sym_tile_index(uplo, MT, m, n) * (int64_t)NB * NB;
            for (int j = 0; j < tn; j++)
                memcpy(A + (int64_t)(n * NB + j) * lda + m * NB, t + (int64_t)j * NB,
                       sizeof(double) * tm);
        }
    }
}

void orc_full_to_symtiles(int uplo, int N, int NB, const double *A, int lda, double *T)
{
    int MT = (N + NB - 1) / NB;
    for (int n = 0; n < MT; n++) {
        int m_lo = (uplo == ORC_LOWER) ? n : 0;
        int m_hi = (uplo == ORC_LOWER) ? MT : n + 1;
        for (int m = m_lo; m < m_hi; m++) {
            int tm = (m == MT - 1) ? N - m * NB : NB;
            int tn = (n == MT - 1) ? N - n * NB : NB;
            double *t = T + orc_sym_tile_index(uplo, MT, m, n) * (int64_t)NB * NB;
            for (int j = 0; j < tn; j++)
                memcpy(t + (int64_t)j * NB, A + (int64_t)(n * NB + j) * lda + m * NB,
                       sizeof(double) * tm);
        }
    }
}

void orc_tiles_to_full(int M, int N, int MB, int NB, const double *T, double *A, int lda)
{
    int MT = (M + MB - 1) / MB, NT = (N + NB - 1) / NB;
    for (int n = 0; n < NT; n++)
        for (int m = 0; m < MT; m++) {
            int tm = (m == MT - 1) ? M - m * MB : MB;
            int tn = (n == NT - 1) ? N - n * NB : NB;
            const double *t = T + ((int64_t)n * MT + m) * (int64_t)MB * NB;
            for (int j = 0; j < tn; j++)
                memcpy(A + (int64_t)(n * NB + j) * lda + m * MB, t + (int64_t)j * MB,
                       sizeof(double) * tm);
        }
}

void orc_full_to_tiles(int M, int N, int MB, int NB, const double *A, int lda, double *T)
{
    int MT = (M + MB - 1) / MB, NT = (N + NB - 1) / NB;
    for (int n = 0; n < NT; n++)
        for (int m = 0; m < MT; m++) {
            int tm = (m == MT - 1) ? M - m * MB : MB;
            int tn = (n == NT - 1) ? N - n * NB : NB;
            double *t = T + ((int64_t)n * MT + m) * (int64_t)MB * NB;
            for (int j = 0; j < tn; j++)
                memcpy(t + (int64_t)j * MB, A + (int64_t)(n * NB + j) * lda + m * MB,
                       sizeof(double) * tm);
        }
}
