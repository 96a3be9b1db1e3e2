// mi_solve.hip -- tile triangular solves on gfx950 behind dplasma_dtrsm /
// dplasma_dpotrs (Left, Lower; NoTrans = src/ztrsm_LLN.jdf, Trans =
// src/ztrsm_LLT.jdf).  The right-hand sides are column-major N x NRHS in HBM
// and are processed in chunks of up to 16 columns.
//
// Per tile step k (forward; backward mirrors it):
//   k_trsv_fwd   X_k = L_kk^{-1} B_k            (ztrsm task, LLN.jdf:29-60)
//   k_gemv_fwd   B_m -= L_mk X_k, m > k         (zgemm task, LLN.jdf:79-174)
// The diagonal solve walks 32-row blocks and applies the inverse of each
// 32x32 diagonal block (computed for all k by k_inv_diag_blocks); the update
// streams each L tile from HBM exactly once per sweep, which is the bound of
// this HBM-bound phase.
#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

constexpr int RHSB = 16;   // right-hand-side chunk

// Inverse of every 32x32 diagonal block of every diagonal tile.
// grid = MT * nblk, 32 threads (thread j solves L w = e_j).
__global__ __launch_bounds__(32) void k_inv_diag_blocks(const double *__restrict__ T, int N, int NB,
                                                       int MT, int unit, double *__restrict__ W)
{
    const int nblk = (NB + 31) / 32;
    const int k = blockIdx.x / nblk, cb = blockIdx.x % nblk;
    const int n = tile_rows(N, NB, MT, k);
    const int c0 = cb * 32;
    double *Wb = W + ((int64_t)k * nblk + cb) * 1024;
    const int j = threadIdx.x;
    if (c0 >= n) {                        // padding block: identity
        for (int i = 0; i < 32; i++) Wb[j * 32 + i] = (i == j) ? 1.0 : 0.0;
        return;
    }
    const int jw = min(32, n - c0);
    const double *L = T + tile_index(MT, k, k) * (int64_t)NB * NB + (int64_t)c0 * NB + c0;
    __shared__ double Ls[32][33];
    for (int c = 0; c < 32; c++) Ls[c][j] = (j < jw && c < jw) ? L[(int64_t)c * NB + j] : 0.0;
    __syncthreads();
    double w[32];
#pragma unroll
    for (int i = 0; i < 32; i++) {
        if (i < jw) {
            double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int p = 0; p < 32; p++)
                if (p < i) s -= Ls[p][i] * w[p];
            w[i] = (i >= j) ? (unit ? s : s / Ls[i][i]) : 0.0;
        } else {
            w[i] = (i == j) ? 1.0 : 0.0;
        }
    }
#pragma unroll
    for (int i = 0; i < 32; i++) Wb[j * 32 + i] = w[i];
}

// Forward diagonal solve of tile k: X_k = L_kk^{-1} B_k for nr <= 16 columns.
// 256 threads: r = t % 32 (row in block), s = t / 32 (k-slice).
__global__ __launch_bounds__(256) void k_trsv_fwd(const double *__restrict__ T, int N, int NB, int MT,
                                                 int k, const double *__restrict__ W,
                                                 double *__restrict__ B, int ldb, int nr)
{
    extern __shared__ double sm[];
    double *Xs = sm;                        // [NB][RHSB]
    double *red = sm + NB * RHSB;           // [8][32][RHSB]
    double *Ys = red + 8 * 32 * RHSB;       // [32][RHSB]
    const int tid = threadIdx.x, r = tid & 31, s = tid >> 5;
    const int n = tile_rows(N, NB, MT, k);
    const int nblk = (NB + 31) / 32;
    const double *L = T + tile_index(MT, k, k) * (int64_t)NB * NB;
    double *Bk = B + (int64_t)k * NB;
    for (int idx = tid; idx < NB * RHSB; idx += 256) {
        int i = idx / RHSB, j = idx % RHSB;
        Xs[idx] = (i < n && j < nr) ? Bk[(int64_t)j * ldb + i] : 0.0;
    }
    __syncthreads();
    for (int cb = 0; cb * 32 < n; cb++) {
        const int c0 = cb * 32;
        double part[RHSB];
#pragma unroll
        for (int j = 0; j < RHSB; j++) part[j] = 0.0;
        if (c0 + r < n)
            for (int p = s; p < c0; p += 8) {
                const double l = L[(int64_t)p * NB + c0 + r];
#pragma unroll
                for (int j = 0; j < RHSB; j++) part[j] += l * Xs[p * RHSB + j];
            }
#pragma unroll
        for (int j = 0; j < RHSB; j++) red[(s * 32 + r) * RHSB + j] = part[j];
        __syncthreads();
        for (int idx = tid; idx < 32 * RHSB; idx += 256) {
            const int rr = idx / RHSB, j = idx % RHSB;
            double y = Xs[(c0 + rr) * RHSB + j];
            for (int q = 0; q < 8; q++) y -= red[(q * 32 + rr) * RHSB + j];
            Ys[rr * RHSB + j] = (c0 + rr < n) ? y : 0.0;
        }
        __syncthreads();
        const double *Wb = W + ((int64_t)k * nblk + cb) * 1024;
        for (int idx = tid; idx < 32 * RHSB; idx += 256) {
            const int rr = idx / RHSB, j = idx % RHSB;
            double x = 0.0;
            for (int q = 0; q <= rr; q++) x += Wb[q * 32 + rr] * Ys[q * RHSB + j];
            if (c0 + rr < n) Xs[(c0 + rr) * RHSB + j] = x;
        }
        __syncthreads();
    }
    for (int idx = tid; idx < NB * RHSB; idx += 256) {
        int i = idx / RHSB, j = idx % RHSB;
        if (i < n && j < nr) Bk[(int64_t)j * ldb + i] = Xs[idx];
    }
}

// B_m -= L_mk X_k for m = k+1 .. MT-1.  Grid = (MT-k-1) * strips of 256 rows;
// thread = one row; X_k staged in LDS.
__global__ __launch_bounds__(256) void k_gemv_fwd(const double *__restrict__ T, int N, int NB, int MT,
                                                 int k, double *__restrict__ B, int ldb, int nr)
{
    extern __shared__ double Xs[];          // [NB][RHSB]
    const int strips = (NB + 255) / 256;
    const int m = k + 1 + blockIdx.x / strips;
    const int r = (blockIdx.x % strips) * 256 + threadIdx.x;
    const int rows_m = tile_rows(N, NB, MT, m);
    const double *Bk = B + (int64_t)k * NB;
    for (int idx = threadIdx.x; idx < NB * RHSB; idx += 256) {
        int i = idx / RHSB, j = idx % RHSB;
        Xs[idx] = (j < nr) ? Bk[(int64_t)j * ldb + i] : 0.0;
    }
    __syncthreads();
    if (r >= rows_m) return;
    const double *L = T + tile_index(MT, m, k) * (int64_t)NB * NB + r;
    double acc[RHSB];
#pragma unroll
    for (int j = 0; j < RHSB; j++) acc[j] = 0.0;
    if (nr == 1) {
        for (int p = 0; p < NB; p++) acc[0] += L[(int64_t)p * NB] * Xs[p * RHSB];
    } else {
        for (int p = 0; p < NB; p++) {
            const double l = L[(int64_t)p * NB];
#pragma unroll
            for (int j = 0; j < RHSB; j++) acc[j] += l * Xs[p * RHSB + j];
        }
    }
    double *Bm = B + (int64_t)m * NB + r;
    for (int j = 0; j < nr; j++) Bm[(int64_t)j * ldb] -= acc[j];
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Backward diagonal solve of tile k: X_k = L_kk^{-T} B_k.
// 256 threads = 4 waves; block rows are walked from the bottom.
__global__ __launch_bounds__(256) void k_trsv_bwd(const double *__restrict__ T, int N, int NB, int MT,
                                                 int k, const double *__restrict__ W,
                                                 double *__restrict__ B, int ldb, int nr)
{
    extern __shared__ double sm[];
    double *Xs = sm;                        // [NB][RHSB]
    double *Ys = sm + NB * RHSB;            // [32][RHSB]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = tile_rows(N, NB, MT, k);
    const int nblk = (NB + 31) / 32;
    const double *L = T + tile_index(MT, k, k) * (int64_t)NB * NB;
    double *Bk = B + (int64_t)k * NB;
    for (int idx = tid; idx < NB * RHSB; idx += 256) {
        int i = idx / RHSB, j = idx % RHSB;
        Xs[idx] = (i < n && j < nr) ? Bk[(int64_t)j * ldb + i] : 0.0;
    }
    __syncthreads();
    const int ncb = (n + 31) / 32;
    for (int cb = ncb - 1; cb >= 0; cb--) {
        const int c0 = cb * 32, c1 = min(n, c0 + 32);
        // y_c = b_c - sum_{p >= c1} L(p, c) x_p, one wave per column c
        for (int cc = wave; cc < 32; cc += 4) {
            const int c = c0 + cc;
            double part[RHSB];
#pragma unroll
            for (int j = 0; j < RHSB; j++) part[j] = 0.0;
            if (c < n)
                for (int p = c1 + lane; p < n; p += 64) {
                    const double l = L[(int64_t)c * NB + p];
#pragma unroll
                    for (int j = 0; j < RHSB; j++) part[j] += l * Xs[p * RHSB + j];
                }
            for (int j = 0; j < nr; j++) {
                const double t = wave_sum(part[j]);
                if (lane == 0) Ys[cc * RHSB + j] = (c < n) ? Xs[c * RHSB + j] - t : 0.0;
            }
        }
        __syncthreads();
        // x = W^T y with W = inv(L_cb,cb): x_r = sum_{q >= r} W(q, r) y_q
        const double *Wb = W + ((int64_t)k * nblk + cb) * 1024;
        for (int idx = tid; idx < 32 * RHSB; idx += 256) {
            const int rr = idx / RHSB, j = idx % RHSB;
            double x = 0.0;
            for (int q = rr; q < 32; q++) x += Wb[rr * 32 + q] * Ys[q * RHSB + j];
            if (c0 + rr < n && j < nr) Xs[(c0 + rr) * RHSB + j] = x;
        }
        __syncthreads();
    }
    for (int idx = tid; idx < NB * RHSB; idx += 256) {
        int i = idx / RHSB, j = idx % RHSB;
        if (i < n && j < nr) Bk[(int64_t)j * ldb + i] = Xs[idx];
    }
}

// B_m -= L_km^T X_k for m = 0 .. k-1.  Grid = k tiles * (NB/16) column groups;
// each wave reduces 4 columns of L_km (contiguous in HBM) against X_k.
__global__ __launch_bounds__(256) void k_gemv_bwd(const double *__restrict__ T, int N, int NB, int MT,
                                                 int k, double *__restrict__ B, int ldb, int nr)
{
    extern __shared__ double Xs[];          // [NB][RHSB]
    const int groups = (NB + 15) / 16;
    const int m = blockIdx.x / groups;
    const int g = blockIdx.x % groups;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rows_k = tile_rows(N, NB, MT, k);
    const double *Bk = B + (int64_t)k * NB;
    for (int idx = threadIdx.x; idx < NB * RHSB; idx += 256) {
        int i = idx / RHSB, j = idx % RHSB;
        Xs[idx] = (i < rows_k && j < nr) ? Bk[(int64_t)j * ldb + i] : 0.0;
    }
    __syncthreads();
    const double *L = T + tile_index(MT, k, m) * (int64_t)NB * NB;
    for (int cc = 0; cc < 4; cc++) {
        const int i = g * 16 + wave * 4 + cc;            // column of L_km = row of B_m
        if (i >= NB) break;
        double part[RHSB];
#pragma unroll
        for (int j = 0; j < RHSB; j++) part[j] = 0.0;
        for (int p = lane; p < rows_k; p += 64) {
            const double l = L[(int64_t)i * NB + p];
#pragma unroll
            for (int j = 0; j < RHSB; j++) part[j] += l * Xs[p * RHSB + j];
        }
        for (int j = 0; j < nr; j++) {
            const double t = wave_sum(part[j]);
            if (lane == 0) B[(int64_t)m * NB + i + (int64_t)j * ldb] -= t;
        }
    }
}

__global__ void k_scale(double *__restrict__ B, int N, int nrhs, int ldb, double alpha)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < (int64_t)N * nrhs) {
        const int64_t c = i / N, r = i % N;
        B[c * ldb + r] *= alpha;
    }
}

struct SolveWs {
    double *w = nullptr;
    size_t bytes = 0;
};
SolveWs g_sws;

int ensure_sws(size_t bytes)
{
    if (g_sws.bytes >= bytes) return 0;
    if (g_sws.w) (void)hipFree(g_sws.w);
    g_sws.w = nullptr;
    g_sws.bytes = 0;
    MI_CHECK(hipMalloc(&g_sws.w, bytes));
    g_sws.bytes = bytes;
    return 0;
}

int set_lds(const void *f, size_t bytes)
{
    if (bytes > 65536) MI_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return 0;
}

}  // namespace

extern "C" int mi_dtrsm_LL(mi_ctx_t *c, int trans, int diag, double alpha, int N, int NB,
                           const double *dT, int NRHS, double *dB, int ldb)
{
    if (N <= 0 || NRHS <= 0) return 0;
    if (NB <= 0 || ldb < N) { set_error("bad sizes"); return -2; }
    const int MT = (N + NB - 1) / NB;
    const int nblk = (NB + 31) / 32;
    int rc;
    if ((rc = ensure_sws((size_t)MT * nblk * 1024 * sizeof(double)))) return rc;
    hipStream_t s = c->s_main;
    MI_CHECK(hipEventRecord(c->t0, s));
    if (alpha != 1.0) {
        const int64_t tot = (int64_t)N * NRHS;
        MI_LAUNCH(c, s, "dtrsm_scale", (double)tot, 16.0 * tot,
                  k_scale<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(dB, N, NRHS, ldb, alpha));
    }
    MI_LAUNCH(c, s, "dtrsm_inv_diag", (double)MT * NB * 32 * 32 / 3.0, (double)MT * NB * 32 * 16.0,
              k_inv_diag_blocks<<<MT * nblk, 32, 0, s>>>(dT, N, NB, MT, diag == 132, g_sws.w));
    const size_t lds_x = (size_t)NB * RHSB * sizeof(double);
    const size_t lds_f = lds_x + (8 * 32 * RHSB + 32 * RHSB) * sizeof(double);
    const size_t lds_b = lds_x + 32 * RHSB * sizeof(double);
    if ((rc = set_lds((const void *)k_trsv_fwd, lds_f))) return rc;
    if ((rc = set_lds((const void *)k_trsv_bwd, lds_b))) return rc;
    if ((rc = set_lds((const void *)k_gemv_fwd, lds_x))) return rc;
    if ((rc = set_lds((const void *)k_gemv_bwd, lds_x))) return rc;

    for (int j0 = 0; j0 < NRHS; j0 += RHSB) {
        const int nr = (NRHS - j0 < RHSB) ? NRHS - j0 : RHSB;
        double *Bj = dB + (int64_t)j0 * ldb;
        if (trans == 111) {
            for (int k = 0; k < MT; k++) {
                const int n = tile_rows(N, NB, MT, k);
                MI_LAUNCH(c, s, "dtrsm_diag", (double)n * n * nr, (double)n * n * 4.0,
                          k_trsv_fwd<<<1, 256, lds_f, s>>>(dT, N, NB, MT, k, g_sws.w, Bj, ldb, nr));
                if (k + 1 < MT) {
                    const int strips = (NB + 255) / 256;
                    const double rows = (double)(N - (k + 1) * NB);
                    MI_LAUNCH(c, s, "dtrsm_gemv", 2.0 * rows * NB * nr, rows * NB * 8.0,
                              k_gemv_fwd<<<(MT - k - 1) * strips, 256, lds_x, s>>>(dT, N, NB, MT, k,
                                                                                  Bj, ldb, nr));
                }
            }
        } else {
            for (int k = MT - 1; k >= 0; k--) {
                const int n = tile_rows(N, NB, MT, k);
                MI_LAUNCH(c, s, "dtrsm_diag", (double)n * n * nr, (double)n * n * 4.0,
                          k_trsv_bwd<<<1, 256, lds_b, s>>>(dT, N, NB, MT, k, g_sws.w, Bj, ldb, nr));
                if (k > 0) {
                    const int groups = (NB + 15) / 16;
                    const double cols = (double)k * NB;
                    MI_LAUNCH(c, s, "dtrsm_gemv", 2.0 * cols * n * nr, cols * n * 8.0,
                              k_gemv_bwd<<<k * groups, 256, lds_x, s>>>(dT, N, NB, MT, k, Bj, ldb, nr));
                }
            }
        }
    }
    MI_CHECK(hipEventRecord(c->t1, s));
    if ((rc = mi_synchronize(c))) return rc;
    float ms = 0.f;
    MI_CHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_ms = ms;
    return 0;
}

extern "C" int mi_dpotrs_tiles(mi_ctx_t *c, int N, int NB, const double *dT, int NRHS, double *dB, int ldb)
{
    double ms = 0.0;
    int rc = mi_dtrsm_LL(c, 111, 131, 1.0, N, NB, dT, NRHS, dB, ldb);
    if (rc) return rc;
    ms = c->last_ms;
    rc = mi_dtrsm_LL(c, 112, 131, 1.0, N, NB, dT, NRHS, dB, ldb);
    c->last_ms += ms;
    return rc;
}
