// mi_mixed.hip -- mixed-precision solve of A X = B on gfx950: factor in
// FP32 (FP32 MFMA, or BF16 MFMA with a two-term split for the trailing
// updates), refine in FP64.
//
// The refinement follows LAPACK's dsposv (the same ITER codes): the
// FP64 matrix is rounded to FP32 tiles (overflow -> ITER = -2), factored by
// the FP32 tile Cholesky (failure -> ITER = -3), and then
//     X = L32^-T L32^-1 B                     (FP32 solves, widened)
//     repeat: R = B - A X                     (FP64, symmetric tiles of A)
//             stop when every column has ||r_j||inf <= ||x_j||inf ||A||inf eps sqrt(N)
//             X += L32^-T L32^-1 R             (FP32 solves, FP64 update)
// for at most 30 iterations (ITER = -31), with a stall test on top: when
// the largest column residual fails to halve for two iterations in a row,
// the refinement stops early (ITER = -4).  Every negative ITER ends in the
// FP64 factorization of A in place and an FP64 solve, as dsposv does.
//
// Kernels here: FP64 <-> FP32 rounding (tiles and column-major blocks),
// the FP64 symmetric residual R = B - A X over the lower tiles, and column
// infinity norms.  The residual is deterministic: one workgroup owns a
// 64-row strip of R and walks tile row m (tiles (m, n), n <= m) and tile
// column m transposed (tiles (n, m), n > m); every A element is read twice
// per residual (once per side of the diagonal), 64-row x 64 chunks staged
// through LDS so that both sides are read coalesced.
#include <cfloat>
#include <chrono>
#include <cmath>
#include <vector>

#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

inline unsigned grid_of(int64_t work, int per = 256)
{
    int64_t g = (work + per - 1) / per;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}

// out = float(in) over n elements; *ovf = 1 when a finite |in| exceeds
// FLT_MAX (dlag2s semantics).  Vector stores only.
__global__ __launch_bounds__(256) void k_d2s(const double *__restrict__ in, float *__restrict__ out, int64_t n,
                                            int *__restrict__ ovf)
{
    bool bad = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = in[i];
        bad |= fabs(v) > (double)FLT_MAX;
        out[i] = (float)v;
    }
    if (bad) *ovf = 1;
}

// column-major blocks: D (ldd) = float(S (lds)) for M x Nc
__global__ __launch_bounds__(256) void k_d2s_cm(const double *__restrict__ S, int lds, float *__restrict__ D,
                                               int ldd, int M, int Nc, int *__restrict__ ovf)
{
    bool bad = false;
    const int64_t n = (int64_t)M * Nc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / M, r = i - c * M;
        const double v = S[c * lds + r];
        bad |= fabs(v) > (double)FLT_MAX;
        D[c * ldd + r] = (float)v;
    }
    if (bad) *ovf = 1;
}

// X (ldx) = double(S) (add = 0) or X += double(S) (add = 1), M x Nc
__global__ __launch_bounds__(256) void k_s2d_cm(const float *__restrict__ S, int lds, double *__restrict__ X,
                                               int ldx, int M, int Nc, int add)
{
    const int64_t n = (int64_t)M * Nc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / M, r = i - c * M;
        const double v = (double)S[c * lds + r];
        X[c * ldx + r] = add ? X[c * ldx + r] + v : v;
    }
}

// R (ldr) = B (ldb) - A X (ldx) for nr <= 16 right-hand sides; A symmetric,
// lower packed tiles (N, NB).  Grid = MT * ceil(NB/64) strips; 256 threads:
// lane = row of the strip, wave w = right-hand sides 4w .. 4w+3.  Per 64-wide
// chunk of a tile the strip's 64 x 64 block of A (rows i of the strip,
// columns kk) is loaded coalesced -- tile (m, n) by columns when n <= m, tile
// (n, m) by its columns (= our rows) when n > m, the diagonal tile by
// whichever side holds the lower triangle -- and staged in LDS as Ls[i][kk]
// together with the chunk of X; the next chunk is loaded while this one is
// used.
constexpr int RS_ROWS = 64;
__global__ __launch_bounds__(256) void k_dsymm_resid(const double *__restrict__ A, int N, int NB, int MT,
                                                    const double *__restrict__ X, int ldx,
                                                    const double *__restrict__ B, int ldb,
                                                    double *__restrict__ R, int ldr, int nr,
                                                    double *__restrict__ rowabs)
{
    __shared__ double Ls[RS_ROWS][RS_ROWS + 1];
    __shared__ double Xs[RS_ROWS][16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int strips = (NB + RS_ROWS - 1) / RS_ROWS;
    const int m = blockIdx.x / strips;
    const int i0 = (blockIdx.x % strips) * RS_ROWS;        // strip start inside tile m
    const int rows_m = tile_rows(N, NB, MT, m);
    if (i0 >= rows_m) return;
    const int chunks = (NB + RS_ROWS - 1) / RS_ROWS;
    const int nchunk = MT * chunks;

    // loader state: 16 A values + 4 X values per thread for one chunk
    double la[16], lx[4];
    auto load = [&](int c) {
        const int n = c / chunks, k0 = (c % chunks) * RS_ROWS;
        const int cols_n = tile_rows(N, NB, MT, n);
        // A block: element (i, kk), i = strip row, kk = chunk column;
        // thread covers kk = (tid >> 6) + 4s (s < 16) at i = lane (n <= m), or
        // i = (tid >> 6) + 4s at kk = lane (n > m), so that lanes walk memory
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const int a = wave + 4 * s;
            double v = 0.0;
            if (n < m) {                 // tile (m, n): column k0+a, row i0+lane
                const int i = i0 + lane, k = k0 + a;
                if (i < rows_m && k < cols_n) v = A[tile_index(MT, m, n) * (int64_t)NB * NB + (int64_t)k * NB + i];
            } else if (n > m) {          // tile (n, m): column i0+a (our row), row k0+lane (our column)
                const int i = i0 + a, k = k0 + lane;
                if (i < rows_m && k < cols_n) v = A[tile_index(MT, n, m) * (int64_t)NB * NB + (int64_t)i * NB + k];
            } else {                     // diagonal tile, lower triangle only: (max, min)
                const int i = i0 + lane, k = k0 + a;
                if (i < rows_m && k < cols_n) {
                    const int64_t base = tile_index(MT, m, m) * (int64_t)NB * NB;
                    v = (i >= k) ? A[base + (int64_t)k * NB + i] : A[base + (int64_t)i * NB + k];
                }
            }
            la[s] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;             // 1024 = 64 rows x 16 rhs, rows across lanes
            const int kk = e & 63, j = e >> 6;
            const int k = k0 + kk;
            lx[q] = (k < cols_n && j < nr) ? X[(int64_t)j * ldx + (int64_t)n * NB + k] : 0.0;
        }
    };
    auto stage = [&](int c) {
        const int n = c / chunks;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const int a = wave + 4 * s;
            if (n > m) Ls[a][lane] = la[s];          // (i = a, kk = lane)
            else Ls[lane][a] = la[s];                // (i = lane, kk = a)
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 256 * q;
            Xs[e & 63][e >> 6] = lx[q];
        }
    };

    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double ra = 0.0;                            // sum |A(i, :)| (wave 0, when asked)
    const bool want_abs = rowabs != nullptr && wave == 0;
    load(0);
    for (int c = 0; c < nchunk; c++) {
        stage(c);
        __syncthreads();
        if (c + 1 < nchunk) load(c + 1);
        if (want_abs) {
#pragma unroll 8
            for (int kk = 0; kk < RS_ROWS; kk++) {
                const double a = Ls[lane][kk];
                ra += fabs(a);
#pragma unroll
                for (int jj = 0; jj < 4; jj++) acc[jj] += a * Xs[kk][4 * wave + jj];
            }
        } else {
#pragma unroll 8
            for (int kk = 0; kk < RS_ROWS; kk++) {
                const double a = Ls[lane][kk];
#pragma unroll
                for (int jj = 0; jj < 4; jj++) acc[jj] += a * Xs[kk][4 * wave + jj];
            }
        }
        __syncthreads();
    }
    const int i = i0 + lane;
    if (i < rows_m) {
        const int64_t row = (int64_t)m * NB + i;
        if (want_abs) rowabs[row] = ra;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const int j = 4 * wave + jj;
            if (j < nr) R[(int64_t)j * ldr + row] = B[(int64_t)j * ldb + row] - acc[jj];
        }
    }
}

// out[j] = max_i |V(i, j)| for the nc columns of V (M x nc, ld); NaN wins.
// One workgroup per column.
__global__ __launch_bounds__(256) void k_colamax(const double *__restrict__ V, int ld, int M,
                                                double *__restrict__ out)
{
    __shared__ double sm[256];
    const double *v = V + (int64_t)blockIdx.x * ld;
    double mx = 0.0;
    for (int i = threadIdx.x; i < M; i += 256) {
        const double a = fabs(v[i]);
        mx = (a != a || mx != mx) ? NAN : fmax(mx, a);
    }
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double a = sm[threadIdx.x], b = sm[threadIdx.x + w];
            sm[threadIdx.x] = (a != a || b != b) ? NAN : fmax(a, b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

__global__ void k_flag_clear(int *p) { *p = 0; }

// HIP check that records the error and jumps to the cleanup label
#define MX_CHECK(call, label)                                                     \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            mi::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                          hipGetErrorString(e_));                                 \
            rc = -1000 - (int)e_;                                                 \
            goto label;                                                           \
        }                                                                         \
    } while (0)

// Workspaces of the mixed solve, kept across calls (grown, never shrunk) so
// that a repeated solve allocates nothing: FP32 tiles, FP32 right-hand sides,
// FP64 residual + norms + row sums, FP32 diagonal-block inverses.
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t need)
    {
        if (bytes >= need) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, need) != hipSuccess) { p = nullptr; return nullptr; }
        bytes = need;
        return p;
    }
};
Buf g_l32, g_r32, g_r64, g_w32;

}  // namespace

namespace mi {

// R = B - A X for NRHS columns (chunks of 16), enqueued on s.
int dsymm_resid(Ctx *c, hipStream_t s, int N, int NB, const double *dA, int NRHS, const double *dX, int ldx,
                const double *dB, int ldb, double *dR, int ldr, double *rowabs)
{
    const int MT = (N + NB - 1) / NB;
    const int strips = (NB + RS_ROWS - 1) / RS_ROWS;
    for (int j0 = 0; j0 < NRHS; j0 += 16) {
        const int nr = NRHS - j0 < 16 ? NRHS - j0 : 16;
        MI_LAUNCH(c, s, "dsymm_resid", 2.0 * N * (double)N * nr, 2.0 * 8.0 * N * (double)N / 2.0,
                  k_dsymm_resid<<<MT * strips, 256, 0, s>>>(dA, N, NB, MT, dX + (int64_t)j0 * ldx, ldx,
                                                           dB + (int64_t)j0 * ldb, ldb, dR + (int64_t)j0 * ldr,
                                                           ldr, nr, j0 == 0 ? rowabs : nullptr));
    }
    return 0;
}

}  // namespace mi

extern "C" {

int mi_dsymm_resid(mi_ctx_t *c, int N, int NB, const double *dA, int NRHS, const double *dX, int ldx,
                   const double *dB, int ldb, double *dR, int ldr)
{
    if (N <= 0 || NRHS <= 0) return 0;
    if (NB <= 0 || ldx < N || ldb < N || ldr < N) { set_error("bad sizes"); return -2; }
    int rc = dsymm_resid(c, c->s_main, N, NB, dA, NRHS, dX, ldx, dB, ldb, dR, ldr, nullptr);
    if (rc) return rc;
    return mi_synchronize(c);
}

int mi_dlag2s_tiles(mi_ctx_t *c, int64_t n, const double *dS, float *dD, int *overflow)
{
    *overflow = 0;
    if (n <= 0) return 0;
    int *flag = c->d_info;   // reused: the factorization clears it again
    hipStream_t s = c->s_main;
    k_flag_clear<<<1, 1, 0, s>>>(flag);
    MI_LAUNCH(c, s, "dlag2s", 0.0, 12.0 * (double)n, k_d2s<<<grid_of(n), 256, 0, s>>>(dS, dD, n, flag));
    MI_CHECK(hipMemcpyAsync(overflow, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    return mi_synchronize(c);
}

// Mixed-precision solve (dsposv semantics).  dA: FP64 lower tiles of the SPD
// matrix (overwritten by its FP64 factor only when *iter < 0); dB, dX: N x
// NRHS column-major.  prec: 0 = FP32 factor, 1 = FP32 factor whose trailing
// updates run on BF16 MFMA with the two-term split.
int mi_dsposv_tiles(mi_ctx_t *c, int N, int NB, double *dA, int NRHS, const double *dB, int ldb, double *dX,
                    int ldx, int prec, int *iter, int *info)
{
    *iter = 0;
    *info = 0;
    if (N <= 0 || NRHS <= 0) return 0;
    if (NB <= 0 || ldb < N || ldx < N) { set_error("bad sizes"); return -2; }
    if (prec != 0 && prec != 1) { set_error("mixed precision must be 0 (FP32) or 1 (BF16 split)"); return -2; }
    const int MT = (N + NB - 1) / NB;
    const int64_t ntile_el = (int64_t)MT * (MT + 1) / 2 * NB * NB;
    hipStream_t s = c->s_main;
    const int ITERMAX = 30;
    const double eps = 0x1p-53;
    int rc = 0;

    float *L32 = (float *)g_l32.get((size_t)ntile_el * sizeof(float));
    float *R32 = (float *)g_r32.get((size_t)N * NRHS * sizeof(float));
    double *R = (double *)g_r64.get(((size_t)N * NRHS + 2 * NRHS + 1 + (size_t)N) * sizeof(double));
    float *W32 = (float *)g_w32.get(trsm_ws_elems(N, NB) * sizeof(float));
    if (!L32 || !R32 || !R || !W32) {
        set_error("out of device memory for the mixed-precision solve");
        return -3;
    }
    double *nrm = R + (size_t)N * NRHS;           // [0, NRHS): ||r_j||, [NRHS, 2 NRHS): ||x_j||, ||A||
    double *rs = nrm + 2 * NRHS + 1;              // row sums of |A|
    std::vector<double> h(2 * NRHS + 1);
    double anrm = 0.0, best = INFINITY;
    int stall = 0, ovf = 0;
    auto fall_back = [&](int code) -> int {
        *iter = code;
        int r = mi_dpotrf_tiles(c, N, NB, dA, info);
        if (r) return r;
        if (*info) return 0;
        MI_CHECK(hipMemcpy2DAsync(dX, (size_t)ldx * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                  (size_t)N * sizeof(double), NRHS, hipMemcpyDeviceToDevice, s));
        return mi_dpotrs_tiles(c, N, NB, dA, NRHS, dX, ldx);
    };

    // span on the host clock: the nested blocking calls use the context's events
    const auto t_start = std::chrono::steady_clock::now();
    // A -> FP32 tiles
    if ((rc = mi_dlag2s_tiles(c, ntile_el, dA, L32, &ovf))) goto done;
    if (ovf) { rc = fall_back(-2); goto done; }
    {
        int sinfo = 0;
        mi_set_spotrf_update(prec);
        rc = mi_spotrf_tiles(c, N, NB, L32, &sinfo);
        mi_set_spotrf_update(0);
        if (rc) goto done;
        if (sinfo) { rc = fall_back(-3); goto done; }
    }
    {
        auto solve32 = [&]() -> int {
            int r;
            if ((r = trsm_enqueue<float>(c, s, 111, N, NB, L32, W32, NRHS, R32, N))) return r;
            return trsm_enqueue<float>(c, s, 112, N, NB, L32, W32, NRHS, R32, N);
        };
        double cte = 0.0;                         // ||A||inf eps sqrt(N), from the first residual
        int it = 0;
        bool conv = false;
        if ((rc = trsm_prepare<float>(c, s, 0, N, NB, L32, W32))) goto done;
        // X0 = L32^-T L32^-1 B
        k_flag_clear<<<1, 1, 0, s>>>(c->d_info);
        k_d2s_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(dB, ldb, R32, N, N, NRHS, c->d_info);
        if ((rc = solve32())) goto done;
        k_s2d_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(R32, N, dX, ldx, N, NRHS, 0);
        for (it = 0; it <= ITERMAX; it++) {
            // the first residual also yields the row sums of |A|, hence ||A||inf
            if ((rc = dsymm_resid(c, s, N, NB, dA, NRHS, dX, ldx, dB, ldb, R, N, it == 0 ? rs : nullptr)))
                goto done;
            k_colamax<<<NRHS, 256, 0, s>>>(R, N, N, nrm);
            k_colamax<<<NRHS, 256, 0, s>>>(dX, ldx, N, nrm + NRHS);
            if (it == 0) k_colamax<<<1, 256, 0, s>>>(rs, N, N, nrm + 2 * NRHS);
            MX_CHECK(hipMemcpyAsync(h.data(), nrm, (2 * NRHS + 1) * sizeof(double), hipMemcpyDeviceToHost, s),
                     done);
            MX_CHECK(hipStreamSynchronize(s), done);
            if (it == 0) {
                anrm = h[2 * NRHS];
                cte = anrm * eps * std::sqrt((double)N);
            }
            double worst = 0.0;
            conv = true;
            for (int j = 0; j < NRHS; j++) {
                if (!(h[j] <= h[NRHS + j] * cte)) conv = false;
                worst = (h[j] != h[j]) ? INFINITY : std::max(worst, h[j]);
            }
            if (conv || it == ITERMAX) break;
            // stall: the largest residual failed to halve twice in a row
            stall = (worst > 0.5 * best) ? stall + 1 : 0;
            best = std::min(best, worst);
            if (stall >= 2 || !(worst < INFINITY)) break;
            k_d2s_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(R, N, R32, N, N, NRHS, c->d_info);
            if ((rc = solve32())) goto done;
            k_s2d_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(R32, N, dX, ldx, N, NRHS, 1);
        }
        if (conv) {
            *iter = it;
        } else {
            rc = fall_back(it >= ITERMAX ? -31 : -4);
        }
    }
done:
    if (rc) return rc;
    MI_CHECK(hipStreamSynchronize(s));
    c->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return 0;
}

}  // extern "C"
