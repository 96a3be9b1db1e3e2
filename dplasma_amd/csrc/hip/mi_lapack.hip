// mi_lapack.hip -- LAPACK-layout entry points over the tile path.
//
// The reference reaches LAPACK-layout memory through a block-cyclic
// descriptor with PARSEC_MATRIX_LAPACK storage over the caller's
// column-major array (lda = LLD), built by the ScaLAPACK-style wrapper
// pdpotrf_w (src/scalapack_wrappers/dplasma_wrapper_pdpotrf.c:139-232)
// and staged tile by tile with 2-D pitched copies
// (src/cuda/lapack_cuda_stage_in.c:18-162).  Here the whole uplo triangle of
// a column-major array already in HBM is staged at once into the packed
// lower tile image every kernel works on, and the factor is written back
// into the same triangle:
//   - Lower: image element (i, j), i >= j, is A(i, j);
//   - Upper: image element (i, j) is A(j, i) (the image holds L = U^T), so
//     the copy transposes, through a 32 x 32 LDS block that keeps both the
//     reads and the writes in whole columns.
// The other triangle of A is neither read nor written (LAPACK dpotrf
// semantics).  Image entries outside the triangle or beyond N are zero.
#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

constexpr int UPLO_UPPER = 121;   // dplasmaUpper (src/include/dplasma/constants.h)
constexpr int UPLO_LOWER = 122;   // dplasmaLower

// tile index t of the packed lower layout -> (m, n): column n starts at
// B(n) = n (2 MT - n + 1) / 2, the root of the quadratic, corrected by one
// step either way for rounding
__device__ inline void tile_of(int64_t t, int MT, int &m, int &n)
{
    const double b = 2.0 * MT + 1.0;
    int c = (int)((b - sqrt(b * b - 8.0 * (double)t)) * 0.5);
    auto base = [&](int64_t x) { return x * (2 * (int64_t)MT - x + 1) / 2; };
    while (c > 0 && base(c) > t) c--;
    while (c + 1 < MT && base(c + 1) <= t) c++;
    n = c;
    m = c + (int)(t - base(c));
}

// One 32 x 32 block of tile t = blockIdx.x (block blockIdx.y of the tile's
// ceil(NB/32)^2).  TO_IMAGE: column-major A -> image T; otherwise T's lower
// triangle (the factor) -> A's uplo triangle.
template <bool TO_IMAGE, bool UPPER>
__global__ __launch_bounds__(256) void k_sym_stage(double *__restrict__ T, double *__restrict__ A, int N, int NB,
                                                   int MT, int64_t lda)
{
    __shared__ double buf[32][33];
    const int sb = (NB + 31) / 32;
    int m, n;
    tile_of(blockIdx.x, MT, m, n);
    const int r0 = (blockIdx.y % sb) * 32, c0 = (blockIdx.y / sb) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    double *Tt = T + (int64_t)blockIdx.x * NB * NB;
    // image element (r, c) of this tile: global (m NB + r, n NB + c)
    auto in_tile = [&](int r, int c) { return r < NB && c < NB; };
    auto in_tri = [&](int r, int c) {
        const int i = m * NB + r, j = n * NB + c;
        return in_tile(r, c) && i < N && j < N && i >= j;
    };
    if (!UPPER) {
        for (int y = ty; y < 32; y += 8) {
            const int r = r0 + tx, c = c0 + y;
            const int64_t a = (int64_t)(n * NB + c) * lda + (m * NB + r);
            if (TO_IMAGE) {
                if (in_tile(r, c)) Tt[(int64_t)c * NB + r] = in_tri(r, c) ? A[a] : 0.0;
            } else if (in_tri(r, c)) {
                A[a] = Tt[(int64_t)c * NB + r];
            }
        }
        return;
    }
    // Upper: image (r, c) <-> A(n NB + c, m NB + r), i.e. A[(m NB + r) lda + n NB + c]
    if (TO_IMAGE) {
        for (int y = ty; y < 32; y += 8) {                // A column m NB + r0 + y, rows n NB + c0 + tx
            const int r = r0 + y, c = c0 + tx;
            buf[y][tx] = in_tri(r, c) ? A[(int64_t)(m * NB + r) * lda + (n * NB + c)] : 0.0;
        }
        __syncthreads();
        for (int y = ty; y < 32; y += 8) {                // image column c0 + y, rows r0 + tx
            const int r = r0 + tx, c = c0 + y;
            if (in_tile(r, c)) Tt[(int64_t)c * NB + r] = buf[tx][y];
        }
    } else {
        for (int y = ty; y < 32; y += 8) {
            const int r = r0 + tx, c = c0 + y;
            buf[tx][y] = in_tile(r, c) ? Tt[(int64_t)c * NB + r] : 0.0;
        }
        __syncthreads();
        for (int y = ty; y < 32; y += 8) {
            const int r = r0 + y, c = c0 + tx;
            if (in_tri(r, c)) A[(int64_t)(m * NB + r) * lda + (n * NB + c)] = buf[y][tx];
        }
    }
}

int check_args(const char *f, int uplo, int N, int lda, int NB)
{
    if (uplo != UPLO_UPPER && uplo != UPLO_LOWER) {
        set_error("%s: uplo must be 121 (Upper) or 122 (Lower), got %d", f, uplo);
        return -2;
    }
    if (N < 0) { set_error("%s: N = %d < 0", f, N); return -3; }
    if (lda < (N > 1 ? N : 1)) { set_error("%s: lda = %d < max(1, N = %d)", f, lda, N); return -5; }
    if (NB <= 0) { set_error("%s: tile size NB = %d must be positive", f, NB); return -6; }
    return 0;
}

int stage(mi_ctx_t *c, int to_image, int uplo, int N, int NB, double *dT, double *dA, int lda)
{
    const int MT = (N + NB - 1) / NB;
    const int sb = (NB + 31) / 32;
    const dim3 grid((unsigned)((int64_t)MT * (MT + 1) / 2), (unsigned)(sb * sb));
    if (sb * sb > 65535) { set_error("tile size NB = %d too large for the staging grid", NB); return -6; }
    const double bytes = 2.0 * 8.0 * ((double)N * (N + 1) / 2);
    const bool up = uplo == UPLO_UPPER;
    if (to_image) {
        if (up)
            MI_LAUNCH(c, c->s_main, "dsym_stage_in", 0.0, bytes,
                      k_sym_stage<true, true><<<grid, 256, 0, c->s_main>>>(dT, dA, N, NB, MT, lda));
        else
            MI_LAUNCH(c, c->s_main, "dsym_stage_in", 0.0, bytes,
                      k_sym_stage<true, false><<<grid, 256, 0, c->s_main>>>(dT, dA, N, NB, MT, lda));
    } else {
        if (up)
            MI_LAUNCH(c, c->s_main, "dsym_stage_out", 0.0, bytes,
                      k_sym_stage<false, true><<<grid, 256, 0, c->s_main>>>(dT, dA, N, NB, MT, lda));
        else
            MI_LAUNCH(c, c->s_main, "dsym_stage_out", 0.0, bytes,
                      k_sym_stage<false, false><<<grid, 256, 0, c->s_main>>>(dT, dA, N, NB, MT, lda));
    }
    return mi_synchronize(c);
}

// packed image of the N x N matrix (MT(MT+1)/2 tiles of NB x NB)
int alloc_image(mi_ctx_t *c, int N, int NB, double **dT)
{
    const int64_t MT = (N + NB - 1) / NB;
    return mi_malloc(c, (void **)dT, (size_t)(MT * (MT + 1) / 2) * NB * NB * sizeof(double));
}

}  // namespace

extern "C" int mi_dsym_lapack_to_tiles(mi_ctx_t *c, int uplo, int N, const double *dA, int lda, int NB,
                                       double *dT)
{
    int rc = check_args("mi_dsym_lapack_to_tiles", uplo, N, lda, NB);
    if (rc || N == 0) return rc;
    return stage(c, 1, uplo, N, NB, dT, const_cast<double *>(dA), lda);
}

extern "C" int mi_dsym_tiles_to_lapack(mi_ctx_t *c, int uplo, int N, int NB, const double *dT, double *dA,
                                       int lda)
{
    int rc = check_args("mi_dsym_tiles_to_lapack", uplo, N, lda, NB);
    if (rc || N == 0) return rc;
    return stage(c, 0, uplo, N, NB, const_cast<double *>(dT), dA, lda);
}

extern "C" int mi_dpotrf_lapack(mi_ctx_t *c, int uplo, int N, double *dA, int lda, int NB, int *info)
{
    *info = 0;
    int rc = check_args("mi_dpotrf_lapack", uplo, N, lda, NB);
    if (rc || N == 0) return rc;
    double *dT = nullptr;
    if ((rc = alloc_image(c, N, NB, &dT))) return rc;
    rc = stage(c, 1, uplo, N, NB, dT, dA, lda);
    if (!rc) rc = mi_dpotrf_tiles(c, N, NB, dT, info);
    if (!rc) rc = stage(c, 0, uplo, N, NB, dT, dA, lda);
    mi_free(c, dT);
    return rc;
}

extern "C" int mi_dpotrs_lapack(mi_ctx_t *c, int uplo, int N, int NRHS, const double *dA, int lda, double *dB,
                                int ldb, int NB)
{
    int rc = check_args("mi_dpotrs_lapack", uplo, N, lda, NB);
    if (rc) return rc;
    if (NRHS < 0) { set_error("mi_dpotrs_lapack: NRHS = %d < 0", NRHS); return -4; }
    if (ldb < (N > 1 ? N : 1)) { set_error("mi_dpotrs_lapack: ldb = %d < max(1, N = %d)", ldb, N); return -8; }
    if (N == 0 || NRHS == 0) return 0;
    double *dT = nullptr;
    if ((rc = alloc_image(c, N, NB, &dT))) return rc;
    rc = stage(c, 1, uplo, N, NB, dT, const_cast<double *>(dA), lda);
    if (!rc) rc = mi_dpotrs_tiles(c, N, NB, dT, NRHS, dB, ldb);
    mi_free(c, dT);
    return rc;
}

extern "C" int mi_dposv_lapack(mi_ctx_t *c, int uplo, int N, int NRHS, double *dA, int lda, double *dB, int ldb,
                               int NB, int *info)
{
    *info = 0;
    int rc = check_args("mi_dposv_lapack", uplo, N, lda, NB);
    if (rc) return rc;
    if (NRHS < 0) { set_error("mi_dposv_lapack: NRHS = %d < 0", NRHS); return -4; }
    if (ldb < (N > 1 ? N : 1)) { set_error("mi_dposv_lapack: ldb = %d < max(1, N = %d)", ldb, N); return -8; }
    if (N == 0) return 0;
    double *dT = nullptr;
    if ((rc = alloc_image(c, N, NB, &dT))) return rc;
    rc = stage(c, 1, uplo, N, NB, dT, dA, lda);
    if (!rc) rc = mi_dpotrf_tiles(c, N, NB, dT, info);
    if (!rc) rc = stage(c, 0, uplo, N, NB, dT, dA, lda);
    if (!rc && *info == 0 && NRHS > 0) rc = mi_dpotrs_tiles(c, N, NB, dT, NRHS, dB, ldb);
    mi_free(c, dT);
    return rc;
}
