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
// infinity norms.  The residual is deterministic.  For even 64 <= NB <= 512 every
// stored A element is read once: one workgroup per tile forms A_mn X_n and
// A_mn^T X_m on FP64 MFMA into per-tile partials, and a second pass sums the
// partials of each row in a fixed order.  Other tiles use a strip kernel
// that reads every element twice (once per side of the diagonal).
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

// Packed lower tiles (N, NB): D = float(S) on the lower triangle within N,
// 0 on the strict upper parts of the diagonal tiles and on padding rows, and
// *ovf = 1 when a value of the lower triangle exceeds FLT_MAX in magnitude
// (dlat2s semantics: only the referenced triangle is converted and tested).
// Workgroup b converts part b % per of packed tile b / per (tiles are stored
// column by column, so b / per is the tile's storage index); with NB a
// multiple of 4 each thread moves runs of four elements of one column.
__global__ __launch_bounds__(256) void k_d2s_lower(const double *__restrict__ S, float *__restrict__ D, int N,
                                                  int NB, int MT, int per, int *__restrict__ ovf)
{
    int t = blockIdx.x / per;
    const int part = blockIdx.x - t * per;
    const int64_t base = (int64_t)t * NB * NB;
    int col = 0;                                   // tile column of tile t
    while (t >= MT - col) { t -= MT - col; col++; }
    const int row = col + t;
    const int rows = min(NB, N - row * NB), cols = min(NB, N - col * NB);
    const bool dg = row == col;
    const int T2 = NB * NB;
    auto conv = [&](double x, int r, int cl, bool &bad) {
        const bool ref = r < rows && cl < cols && (!dg || r >= cl);
        bad |= ref && fabs(x) > (double)FLT_MAX;
        return ref ? (float)x : 0.f;
    };
    bool bad = false;
    if ((NB & 3) == 0) {
        for (int q = part * 256 + (int)threadIdx.x; q < T2 / 4; q += per * 256) {
            const int e = 4 * q, cl = e / NB, r = e - cl * NB;
            const double2 x0 = *(const double2 *)(S + base + e), x1 = *(const double2 *)(S + base + e + 2);
            const float4 v = {conv(x0.x, r, cl, bad), conv(x0.y, r + 1, cl, bad), conv(x1.x, r + 2, cl, bad),
                              conv(x1.y, r + 3, cl, bad)};
            *(float4 *)(D + base + e) = v;
        }
    } else {
        for (int e = part * 256 + (int)threadIdx.x; e < T2; e += per * 256) {
            const int cl = e / NB;
            D[base + e] = conv(S[base + e], e - cl * NB, cl, bad);
        }
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

// The same residual with every stored element of A read once (64 <= NB <= 512, NB even),
// in two launches.  k_symm_part: one workgroup per stored tile (m, n),
// m >= n, computes on FP64 MFMA (16x16x4, the 16 columns are the right-hand
// sides)
//     P1(m, n) = A_mn X_n        (NB x 16, rows of tile m)
//     P2(m, n) = A_mn^T X_m      (NB x 16, rows of tile n)
// and, when asked, the row and column sums of |A_mn|.  On the diagonal
// tile only the lower triangle is stored: P1 takes it with the diagonal,
// P2 the strict lower part, so P1 + P2 is the full symmetric product.  Wave
// w owns rows [128 w, 128 w + 128) of the tile for P1.  Per 32-column chunk
// each wave reads its 32 x 32 A blocks as pairs of consecutive rows (MFMA
// row i of product t is row 2 i + t); the a-operands of P2 (rows along k)
// are the same values transposed through a per-wave LDS buffer.  The four waves'
// P2 contributions to a chunk are summed in LDS in a fixed order.
// k_symm_sum then forms R = B - sum_n P1(m, n) - sum_m' P2(m', m) for each
// row, in a fixed order, so the result is deterministic.
// Partials per tile: P1 [NB][16], P2 [NB][16], |A| row sums [NB], column
// sums [NB] -- 34 NB doubles at tile_index(m, n) * 34 NB.
constexpr int SY_PART = 34;    // partial doubles per tile row index

template <bool ABS>
__global__ __launch_bounds__(256, 2) void k_symm_part(const double *__restrict__ A, int N, int NB, int MT,
                                                     const double *__restrict__ X, int ldx, int nr,
                                                     double *__restrict__ P)
{
    constexpr int CW = 32;                       // columns per chunk
    __shared__ double Tb[4][CW * 33];            // per wave: 32 x 32 block of A, [column][row], stride 33
    __shared__ double red[4][CW * 16 + CW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    int t = blockIdx.x, n = 0;
    while (t >= MT - n) { t -= MT - n; n++; }
    const int m = n + t;
    const bool dg = (m == n);
    const int rows_m = tile_rows(N, NB, MT, m), cols_n = tile_rows(N, NB, MT, n);
    const int64_t ti = tile_index(MT, m, n);
    const double *At = A + ti * (int64_t)NB * NB;
    double *P1 = P + ti * (int64_t)SY_PART * NB;
    double *P2 = P1 + 16 * NB;
    double *A1 = P2 + 16 * NB;
    double *A2 = A1 + NB;
    const int jc = min(fr, nr - 1);              // clamped right-hand side of this lane
    const bool jok = fr < nr;
    const double *Xn = X + (int64_t)jc * ldx + (int64_t)n * NB;
    const double *Xm = X + (int64_t)jc * ldx + (int64_t)m * NB;
    double *T = Tb[w];
    // row groups of 32 (two MFMA row groups t, MFMA row i at row 2 i + t, so
    // that a lane loads two consecutive rows at once)
    const int ng = min(4, max(0, (rows_m - 128 * w + 31) / 32));   // uniform per wave

    dx4 p1[4][2];
    double a1[4][2];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
        for (int tt = 0; tt < 2; tt++) {
            p1[g][tt] = dx4{0.0, 0.0, 0.0, 0.0};
            a1[g][tt] = 0.0;
        }
    // a-operands of row group g of chunk c0: A(rb + 2 fr + tt, c0 + 4 u + fk) as
    // pairs (the tile is column-major, NB even); the row base is made opaque
    // per call, so that row-dependent addresses are not precomputed for all
    // row groups
    auto load_a = [&](double2 (&a)[8], int c0, int g) {
        int rb = 128 * w + 32 * g;
        asm volatile("" : "+v"(rb));
        const int r = rb + 2 * fr;
        const unsigned rc = (unsigned)min(r, ((rows_m - 1) & ~1));
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = c0 + 4 * u + fk;
            const double2 v = *(const double2 *)(At + (unsigned)min(c, cols_n - 1) * (unsigned)NB + rc);
            // in the tile, and in the lower triangle on the diagonal tile
            const int ok0 = (rows_m - 1 - r) | (cols_n - 1 - c) | (dg ? r - c : 0);
            const int ok1 = (rows_m - 2 - r) | (cols_n - 1 - c) | (dg ? r + 1 - c : 0);
            a[u].x = ok0 >= 0 ? v.x : 0.0;
            a[u].y = ok1 >= 0 ? v.y : 0.0;
        }
    };
    // b-operands of P2 for row group g: X_m(rb + 4 s + fk, fr)
    auto load_xm = [&](double (&xm)[8], int g) {
        int rb = 128 * w + 32 * g;
        asm volatile("" : "+v"(rb));
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int rr = rb + 4 * s + fk;
            const double v = Xm[(unsigned)min(rr, rows_m - 1)];
            xm[s] = (jok && rr < rows_m) ? v : 0.0;
        }
    };
    const int nch = (NB + CW - 1) / CW;
    for (int ch = 0; ch < nch; ch++) {
        const int c0 = CW * ch;
        // b-operands of P1: X_n(c0 + 4 u + fk, fr)
        double xn[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = c0 + 4 * u + fk;
            const double v = Xn[(unsigned)min(c, cols_n - 1)];
            xn[u] = (jok && c < cols_n) ? v : 0.0;
        }
        dx4 p2[2];
        double a2[2];
#pragma unroll
        for (int cg = 0; cg < 2; cg++) {
            p2[cg] = dx4{0.0, 0.0, 0.0, 0.0};
            a2[cg] = 0.0;
        }
#pragma unroll
        for (int g = 0; g < 4; g++) {
            if (g < ng) {                         // wave-uniform
                // this group's operands (the second workgroup on the CU hides
                // their latency)
                asm volatile("" ::: "memory");
                double2 a[8];
                load_a(a, c0, g);
                double xm[8];
                load_xm(xm, g);
                const int r0 = 128 * w + 32 * g;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    p1[g][0] = MF<double>::mfma(a[u].x, xn[u], p1[g][0]);
                    p1[g][1] = MF<double>::mfma(a[u].y, xn[u], p1[g][1]);
                    if (ABS) {
                        a1[g][0] += fabs(a[u].x);
                        a1[g][1] += fabs(a[u].y);
                    }
                    T[(4 * u + fk) * 33 + 2 * fr] = a[u].x;
                    T[(4 * u + fk) * 33 + 2 * fr + 1] = a[u].y;
                }
#pragma unroll
                for (int cg = 0; cg < 2; cg++)
#pragma unroll
                    for (int s = 0; s < 8; s++) {
                        // A(r0 + 4 s + fk, c0 + 16 cg + fr); strict lower part on the diagonal tile
                        double v = T[(16 * cg + fr) * 33 + 4 * s + fk];
                        if (dg && !(r0 + 4 * s + fk > c0 + 16 * cg + fr)) v = 0.0;
                        p2[cg] = MF<double>::mfma(v, xm[s], p2[cg]);
                        if (ABS) a2[cg] += fabs(v);
                    }
            }
        }
        // element q of p2[cg] in lane (fr, fk): P2(c0 + 16 cg + 4 q + fk, fr)
        double *Rw = red[w];
#pragma unroll
        for (int cg = 0; cg < 2; cg++)
#pragma unroll
            for (int q = 0; q < 4; q++) Rw[(16 * cg + 4 * q + fk) * 16 + fr] = p2[cg][q];
        if (ABS) {
#pragma unroll
            for (int cg = 0; cg < 2; cg++) {
                double v = a2[cg];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (fk == 0) Rw[CW * 16 + 16 * cg + fr] = v;
            }
        }
        __syncthreads();
        for (int e = tid; e < CW * 16 + (ABS ? CW : 0); e += 256) {
            const double v = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
            if (e < CW * 16) {
                const int cl = e >> 4;
                if (c0 + cl < NB) P2[(int64_t)(c0 + cl) * 16 + (e & 15)] = v;
            } else if (c0 + e - CW * 16 < NB) {
                A2[c0 + e - CW * 16] = v;
            }
        }
        __syncthreads();
    }
    // element q of p1[g][tt] in lane (fr, fk): P1(128 w + 32 g + 2 (4 q + fk) + tt, fr)
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int r0 = 128 * w + 32 * g;
#pragma unroll
        for (int tt = 0; tt < 2; tt++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = r0 + 2 * MF<double>::row(q, fk) + tt;
                if (r < NB) P1[(int64_t)r * 16 + fr] = p1[g][tt][q];
            }
            if (ABS) {
                double v = a1[g][tt];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (fk == 0 && r0 + 2 * fr + tt < NB) A1[r0 + 2 * fr + tt] = v;
            }
        }
    }
}

// The same partials for NB = 128 RT (RT = 1..4), the tile crossing HBM once
// as whole 16-column chunks (NB x 16 contiguous doubles) loaded straight
// into LDS by global_load_lds_dwordx4, double-buffered, so that the next
// chunk is in flight while this one is used.  One workgroup of eight waves
// per stored tile (m, n); wave w owns rows [16 RT w, 16 RT (w + 1)):
//   P1 rows of its slice:   A(rows, chunk) X_n(chunk)       (RT x 4 MFMAs)
//   P2 of the chunk's 16 columns over its rows:
//                           A(rows, chunk)^T X_m(rows)      (4 RT MFMAs)
// Both read their a-operands from the same LDS copy (P2's transposed).  The
// eight waves' P2 contributions are summed in LDS in a fixed order, so the
// result is deterministic.  Masks (diagonal tile, ragged tiles) are applied
// to the operands as they are read.
typedef __attribute__((address_space(3))) void *lds_ptr;
// LDS reads at byte address a + immediate offsets, waited for in the same
// block: the compiler would otherwise wait for the chunk loads in flight
// into the other buffer before every LDS read (it cannot tell the two
// buffers apart).  Offsets: B + S2 k (four), B + S1 (k / 4) + S2 (k % 4)
// (eight).
template <int B, int S2>
__device__ inline void lds_rd4(unsigned a, double (&x)[4])
{
    asm volatile("ds_read_b64 %0, %4 offset:%5\n\t"
                 "ds_read_b64 %1, %4 offset:%6\n\t"
                 "ds_read_b64 %2, %4 offset:%7\n\t"
                 "ds_read_b64 %3, %4 offset:%8\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(x[0]), "=&v"(x[1]), "=&v"(x[2]), "=&v"(x[3])
                 : "v"(a), "i"(B), "i"(B + S2), "i"(B + 2 * S2), "i"(B + 3 * S2)
                 : "memory");
}
template <int B, int S1, int S2>
__device__ inline void lds_rd8(unsigned a, double (&x)[8])
{
    asm volatile("ds_read_b64 %0, %8 offset:%9\n\t"
                 "ds_read_b64 %1, %8 offset:%10\n\t"
                 "ds_read_b64 %2, %8 offset:%11\n\t"
                 "ds_read_b64 %3, %8 offset:%12\n\t"
                 "ds_read_b64 %4, %8 offset:%13\n\t"
                 "ds_read_b64 %5, %8 offset:%14\n\t"
                 "ds_read_b64 %6, %8 offset:%15\n\t"
                 "ds_read_b64 %7, %8 offset:%16\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(x[0]), "=&v"(x[1]), "=&v"(x[2]), "=&v"(x[3]), "=&v"(x[4]), "=&v"(x[5]), "=&v"(x[6]),
                   "=&v"(x[7])
                 : "v"(a), "i"(B), "i"(B + S2), "i"(B + 2 * S2), "i"(B + 3 * S2), "i"(B + S1), "i"(B + S1 + S2),
                   "i"(B + S1 + 2 * S2), "i"(B + S1 + 3 * S2)
                 : "memory");
}
template <bool ABS, int RT>
__global__ __launch_bounds__(512, 1) void k_symm_part2(const double *__restrict__ A, int N, int MT,
                                                      const double *__restrict__ X, int ldx, int nr,
                                                      double *__restrict__ P)
{
    constexpr int CW = 16;                       // columns per chunk
    constexpr int TB = RT * 128;                 // tile order
    constexpr int CS = TB + 16;                  // LDS column stride (doubles): 128 bytes of padding
    constexpr int RWV = RT * 16;                 // rows per wave
    // two buffers as two objects, so that reads of one need not wait for
    // the direct loads into the other
    __shared__ __align__(16) double S0[CW * CS];
    __shared__ __align__(16) double S1[CW * CS];
    __shared__ double red[8][CW * 16 + CW];
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, w = tid >> 6, fr = lane & 15, fk = lane >> 4;
    int t = blockIdx.x, n = 0;
    while (t >= MT - n) { t -= MT - n; n++; }
    const int m = n + t;
    const bool dg = m == n;
    const int rows_m = tile_rows(N, TB, MT, m), cols_n = tile_rows(N, TB, MT, n);
    const bool edge = dg || rows_m < TB || cols_n < TB;      // workgroup-uniform
    const int64_t ti = tile_index(MT, m, n);
    const double *At = A + ti * (int64_t)TB * TB;
    double *P1 = P + ti * (int64_t)SY_PART * TB;
    double *P2 = P1 + 16 * TB;
    double *A1 = P2 + 16 * TB;
    double *A2 = A1 + TB;
    const int jc = min(fr, nr - 1);              // clamped right-hand side of this lane
    const double jm = fr < nr ? 1.0 : 0.0;
    const double *Xn = X + (int64_t)jc * ldx + (int64_t)n * TB;
    const double *Xm = X + (int64_t)jc * ldx + (int64_t)m * TB;
    const int rw0 = w * RWV;

    // X values are read at clamped (valid, finite) positions and scaled by
    // 0 or 1, so that no load depends on a condition.  b-operands of P2:
    // X_m(rw0 + 4 s + fk, fr)
    double xm[4 * RT];
#pragma unroll
    for (int s = 0; s < 4 * RT; s++) {
        const int r = rw0 + 4 * s + fk;
        xm[s] = Xm[(unsigned)min(r, rows_m - 1)] * (r < rows_m ? jm : 0.0);
    }
    dx4 p1[RT];
    double a1[RT];
#pragma unroll
    for (int i = 0; i < RT; i++) {
        p1[i] = dx4{0.0, 0.0, 0.0, 0.0};
        a1[i] = 0.0;
    }
    // chunk ch into buffer Sd: 16 RT pieces of 128 doubles (column ch CW +
    // piece / RT, rows 128 (piece % RT) ..), two RT per wave
    auto dload = [&](double *Sd, int ch) {
#pragma unroll
        for (int q = 0; q < 2 * RT; q++) {
            const int pc = w + 8 * q, col = pc / RT, part = pc - col * RT;
            __builtin_amdgcn_global_load_lds(At + (int64_t)(ch * CW + col) * TB + part * 128 + 2 * lane,
                                             (lds_ptr)&Sd[col * CS + part * 128], 16, 0, 0);
        }
    };
    // the products of chunk ch from Sb; masked operands (diagonal tile:
    // P1 the lower triangle, P2 the strict lower; rows and columns past the
    // matrix: zero) when E
    auto compute = [&](auto edge_tag, const double *Sb, int c0, const double (&xn)[4]) {
        constexpr bool E = decltype(edge_tag)::value;
        // a-operands of P1, A(rw0 + 16 i + fr, c0 + 4 u + fk), as ap[4 i + u]
        double ap[4 * RT];
        {
            const unsigned b1 = (unsigned)(size_t)(lds_ptr)&Sb[fk * CS + rw0 + fr];
            if constexpr (RT == 1) {
                double x[4];
                lds_rd4<0, 32 * CS>(b1, x);
#pragma unroll
                for (int k = 0; k < 4; k++) ap[k] = x[k];
            } else {
#pragma unroll
                for (int i0 = 0; i0 + 1 < RT; i0 += 2) {
                    double x[8];
                    if (i0 == 0) lds_rd8<0, 128, 32 * CS>(b1, x);
                    else lds_rd8<256, 128, 32 * CS>(b1, x);
#pragma unroll
                    for (int k = 0; k < 8; k++) ap[4 * i0 + k] = x[k];
                }
                if constexpr (RT == 3) {
                    double x[4];
                    lds_rd4<256, 32 * CS>(b1, x);
#pragma unroll
                    for (int k = 0; k < 4; k++) ap[8 + k] = x[k];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < RT; i++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = rw0 + 16 * i + fr, c = 4 * u + fk;
                double a = ap[4 * i + u];
                if (E) a = (r < rows_m && c0 + c < cols_n && (!dg || r >= c0 + c)) ? a : 0.0;
                if (ABS) a1[i] += fabs(a);
                p1[i] = MF<double>::mfma(a, xn[u], p1[i]);
            }
        // a-operands of P2, A(rw0 + 4 s + fk, c0 + fr), as at[s]
        double at[4 * RT];
        {
            const unsigned b2 = (unsigned)(size_t)(lds_ptr)&Sb[fr * CS + rw0 + fk];
            if constexpr (RT == 1) {
                double x[4];
                lds_rd4<0, 32>(b2, x);
#pragma unroll
                for (int k = 0; k < 4; k++) at[k] = x[k];
            } else {
#pragma unroll
                for (int s0 = 0; s0 + 7 < 4 * RT; s0 += 8) {
                    double x[8];
                    if (s0 == 0) lds_rd8<0, 128, 32>(b2, x);
                    else lds_rd8<256, 128, 32>(b2, x);
#pragma unroll
                    for (int k = 0; k < 8; k++) at[s0 + k] = x[k];
                }
                if constexpr (RT == 3) {
                    double x[4];
                    lds_rd4<256, 32>(b2, x);
#pragma unroll
                    for (int k = 0; k < 4; k++) at[8 + k] = x[k];
                }
            }
        }
        dx4 p2 = dx4{0.0, 0.0, 0.0, 0.0};
        double a2 = 0.0;
#pragma unroll
        for (int s = 0; s < 4 * RT; s++) {
            const int r = rw0 + 4 * s + fk;
            double a = at[s];
            if (E) a = (r < rows_m && c0 + fr < cols_n && (!dg || r > c0 + fr)) ? a : 0.0;
            if (ABS) a2 += fabs(a);
            p2 = MF<double>::mfma(a, xm[s], p2);
        }
        // element q of p2 in lane (fr, fk): P2(c0 + 4 q + fk, fr)
        double *Rw = red[w];
#pragma unroll
        for (int q = 0; q < 4; q++) Rw[(4 * q + fk) * 16 + fr] = p2[q];
        if (ABS) {
            a2 += __shfl_xor(a2, 16);
            a2 += __shfl_xor(a2, 32);
            if (fk == 0) Rw[CW * 16 + fr] = a2;
        }
    };
    const int nch = (cols_n + CW - 1) / CW;
    // b-operands of P1 for chunk ch, X_n(16 ch + 4 u + fk, fr), read one
    // chunk ahead, before that chunk's own loads
    double xr[4];
    auto read_xn = [&](int ch) {
#pragma unroll
        for (int u = 0; u < 4; u++) xr[u] = Xn[(unsigned)min(ch * CW + 4 * u + fk, cols_n - 1)];
    };
    // P2 (or A2) sum of the previous chunk, held in a register by the
    // threads tid < CW * 16 (+ CW) and stored at the start of the next step,
    // before that step's loads: the wait for those loads then never waits
    // for a store issued just before it
    const bool red_thread = tid < CW * 16 + (ABS ? CW : 0);
    double pend = 0.0;
    auto store_pend = [&](int c0) {
        if (red_thread) {
            if (tid < CW * 16) P2[(int64_t)(c0 + (tid >> 4)) * 16 + (tid & 15)] = pend;
            else A2[c0 + tid - CW * 16] = pend;
        }
    };
    auto step = [&](int ch, const double *Sb, double *Sn) {
        const int c0 = ch * CW;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");      // this wave's loads of chunk ch and its xr
        __syncthreads();                                      // everyone's; the other buffer free
        double xn[4];
#pragma unroll
        for (int u = 0; u < 4; u++) xn[u] = xr[u] * (c0 + 4 * u + fk < cols_n ? jm : 0.0);
        // xn in registers before the loads below are issued, so that no
        // wait for it waits for them
        asm volatile("" : "+v"(xn[0]), "+v"(xn[1]), "+v"(xn[2]), "+v"(xn[3])::"memory");
        if (ch > 0) store_pend(c0 - CW);
        // the next chunk; after the last one, this chunk again (from L2)
        read_xn(min(ch + 1, nch - 1));
        dload(Sn, min(ch + 1, nch - 1));
        if (edge) compute(std::true_type{}, Sb, c0, xn);
        else compute(std::false_type{}, Sb, c0, xn);
        // a barrier without the memory fence of __syncthreads, which would
        // wait for the loads just issued: the LDS writes of red are waited
        // for explicitly, and red is read back by waited asm reads
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (red_thread) {
            double v[8];
            lds_rd8<0, 4 * (CW * 16 + CW) * 8, (CW * 16 + CW) * 8>((unsigned)(size_t)(lds_ptr)&red[0][tid], v);
            pend = ((((((v[0] + v[1]) + v[2]) + v[3]) + v[4]) + v[5]) + v[6]) + v[7];
        }
    };
    read_xn(0);
    dload(S0, 0);
    for (int ch = 0; ch < nch; ch += 2) {
        step(ch, S0, S1);
        if (ch + 1 < nch) step(ch + 1, S1, S0);
    }
    if (nch > 0) store_pend((nch - 1) * CW);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");          // the last, unused loads
    // element q of p1[i] in lane (fr, fk): P1(rw0 + 16 i + 4 q + fk, fr)
#pragma unroll
    for (int i = 0; i < RT; i++) {
#pragma unroll
        for (int q = 0; q < 4; q++) P1[(int64_t)(rw0 + 16 * i + MF<double>::row(q, fk)) * 16 + fr] = p1[i][q];
        if (ABS) {
            double v = a1[i];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (fk == 0) A1[rw0 + 16 * i + fr] = v;
        }
    }
}

// R (ldr) = B (ldb) - (P1 + P2 sums) for nr right-hand sides; thread = (row,
// right-hand side), partials summed over n = 0..m, then m' = m..MT-1.
__global__ __launch_bounds__(256) void k_symm_sum(const double *__restrict__ P, int N, int NB, int MT,
                                                 const double *__restrict__ B, int ldb, double *__restrict__ R,
                                                 int ldr, int nr, double *__restrict__ rowabs)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)N * 16) return;
    const int row = (int)(e >> 4), j = (int)(e & 15);
    const int m = row / NB, r = row - m * NB;
    const int64_t stride = (int64_t)SY_PART * NB;
    double s = 0.0;
    for (int n = 0; n <= m; n++) s += P[tile_index(MT, m, n) * stride + (int64_t)r * 16 + j];
    for (int mm = m; mm < MT; mm++) s += P[tile_index(MT, mm, m) * stride + 16 * (int64_t)NB + (int64_t)r * 16 + j];
    if (j < nr) R[(int64_t)j * ldr + row] = B[(int64_t)j * ldb + row] - s;
    if (rowabs && j == 0) {
        double a = 0.0;
        for (int n = 0; n <= m; n++) a += P[tile_index(MT, m, n) * stride + 32 * (int64_t)NB + r];
        for (int mm = m; mm < MT; mm++) a += P[tile_index(MT, mm, m) * stride + 33 * (int64_t)NB + r];
        rowabs[row] = a;
    }
}

// out[j] = max_i |V(i, j)| for the nc columns of V (M x nc, ld); NaN wins.
// Grid (ceil(M / 4096), nc): each workgroup reduces 4096 rows of one column
// and folds its maximum into out[j] (cleared by k_amax_clear) with a 64-bit
// atomic maximum on the bit pattern: for non-negative values the integer
// order of IEEE doubles is the numeric order, and a NaN with its sign
// cleared lies above +Inf.  The maximum does not depend on the order.
constexpr int AMAX_ROWS = 4096;
__global__ __launch_bounds__(256) void k_amax_clear(double *__restrict__ out, int nc)
{
    for (int j = threadIdx.x; j < nc; j += 256) out[j] = 0.0;
}
__global__ __launch_bounds__(256) void k_colamax(const double *__restrict__ V, int ld, int M, int nc,
                                                double *__restrict__ out)
{
    __shared__ unsigned long long sm[4];
    const int i0 = blockIdx.x * AMAX_ROWS;
    const int i1 = min(M, i0 + AMAX_ROWS);
    for (int j = blockIdx.y; j < nc; j += gridDim.y) {
        const double *v = V + (int64_t)j * ld;
        unsigned long long mx = 0;
        for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(v[i]) & 0x7FFFFFFFFFFFFFFFull;
            mx = b > mx ? b : mx;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long u = __shfl_xor(mx, o);
            mx = u > mx ? u : mx;
        }
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < 4; w++) mx = sm[w] > mx ? sm[w] : mx;
            atomicMax((unsigned long long *)&out[j], mx);
        }
        __syncthreads();
    }
}

// out[0 .. nc) = column maxima of |V| (M x nc, ld), enqueued on s
inline void colamax(hipStream_t s, const double *V, int ld, int M, int nc, double *out)
{
    k_amax_clear<<<1, 256, 0, s>>>(out, nc);
    k_colamax<<<dim3((unsigned)((M + AMAX_ROWS - 1) / AMAX_ROWS), (unsigned)std::min(nc, 65535)), 256, 0, s>>>(
        V, ld, M, nc, out);
}

__global__ void k_flag_clear(int *p) { *p = 0; }

// Waits for s by polling: the host decides the next refinement step as soon
// as the norms are there (a blocking wait that has slept through the
// factorization wakes up ~0.1 ms late).  Every kernel queued here ends by
// itself (the solve's hand-off waits are bounded), so the poll ends.
static hipError_t spin_sync(hipStream_t s)
{
    hipError_t e;
    while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
    }
    return e;
}

// Status words of one refinement step, as doubles next to the norms (one
// copy to the host): the rounding of A overflowed, the rounding of B or of a
// residual overflowed, the FP32 factorization's info, a hand-off wait of the
// solve gave up (abort may be null: no wavefront solve).
__global__ void k_mx_status(double *st, const int *finfo, const int *a_ovf, const int *r_ovf, const int *abort)
{
    st[0] = (double)*a_ovf;
    st[1] = (double)*r_ovf;
    st[2] = (double)*finfo;
    st[3] = abort ? (double)*abort : 0.0;
}

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


}  // namespace

namespace mi {

int g_symm_kernel = 1;   // 1: k_symm_part2 when NB % 128 == 0; 0: k_symm_part

// R = B - A X for NRHS columns (chunks of 16), enqueued on s.
int dsymm_resid(Ctx *c, hipStream_t s, int N, int NB, const double *dA, int NRHS, const double *dX, int ldx,
                const double *dB, int ldb, double *dR, int ldr, double *rowabs)
{
    const int MT = (N + NB - 1) / NB;
    // below NB = 64 the partials outweigh the second read; the row-pair
    // loads need an even NB (16-byte aligned pairs)
    if (NB >= 64 && NB <= 512 && NB % 2 == 0) {
        const int64_t ntile = (int64_t)MT * (MT + 1) / 2;
        double *P = (double *)ws_get(c, WS_MX_PART, (size_t)ntile * SY_PART * NB * sizeof(double));
        if (!P) return -3;
        for (int j0 = 0; j0 < NRHS; j0 += 16) {
            const int nr = NRHS - j0 < 16 ? NRHS - j0 : 16;
            double *ra = j0 == 0 ? rowabs : nullptr;
            const double *Xj = dX + (int64_t)j0 * ldx;
            if (NB % 128 == 0 && g_symm_kernel) {
                // whole-chunk LDS loads (k_symm_part2), NB = 128 .. 512
                const double fl = 2.0 * N * (double)N * nr, by = 8.0 * (double)ntile * NB * NB;
                const unsigned gr = (unsigned)ntile;
#define MI_SYMM2(RT)                                                                                             \
    do {                                                                                                         \
        if (ra)                                                                                                  \
            MI_LAUNCH(c, s, "dsymm_part", fl, by,                                                                \
                      k_symm_part2<true, RT><<<gr, 512, 0, s>>>(dA, N, MT, Xj, ldx, nr, P));                    \
        else                                                                                                     \
            MI_LAUNCH(c, s, "dsymm_part", fl, by,                                                                \
                      k_symm_part2<false, RT><<<gr, 512, 0, s>>>(dA, N, MT, Xj, ldx, nr, P));                   \
    } while (0)
                switch (NB / 128) {
                case 1: MI_SYMM2(1); break;
                case 2: MI_SYMM2(2); break;
                case 3: MI_SYMM2(3); break;
                default: MI_SYMM2(4); break;
                }
#undef MI_SYMM2
            } else if (ra)
                MI_LAUNCH(c, s, "dsymm_part", 2.0 * N * (double)N * nr, 8.0 * (double)ntile * NB * NB,
                          k_symm_part<true><<<(unsigned)ntile, 256, 0, s>>>(dA, N, NB, MT, Xj, ldx, nr, P));
            else
                MI_LAUNCH(c, s, "dsymm_part", 2.0 * N * (double)N * nr, 8.0 * (double)ntile * NB * NB,
                          k_symm_part<false><<<(unsigned)ntile, 256, 0, s>>>(dA, N, NB, MT, Xj, ldx, nr, P));
            MI_LAUNCH(c, s, "dsymm_sum", 0.0, 8.0 * 2.0 * (double)ntile * 16 * NB,
                      k_symm_sum<<<(unsigned)(((int64_t)N * 16 + 255) / 256), 256, 0, s>>>(
                          P, N, NB, MT, dB + (int64_t)j0 * ldb, ldb, dR + (int64_t)j0 * ldr, ldr, nr, ra));
        }
        return 0;
    }
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

int mi_set_symm_kernel(int which)
{
    if (which != 0 && which != 1) { set_error("residual kernel must be 0 (register operands) or 1 (LDS chunks)"); return -2; }
    g_symm_kernel = which;
    return 0;
}

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

    // workspaces of the context, kept across calls so that a repeated solve
    // allocates nothing: FP32 tiles, FP32 right-hand sides, FP64 residual +
    // norms + row sums, FP32 solve workspace
    float *L32 = (float *)ws_get(c, WS_MX_L32, (size_t)ntile_el * sizeof(float));
    float *R32 = L32 ? (float *)ws_get(c, WS_MX_R32, (size_t)N * NRHS * sizeof(float)) : nullptr;
    const int NST = 2 * NRHS + 5;                 // norms and status words
    double *R = R32 ? (double *)ws_get(c, WS_MX_R64, ((size_t)N * NRHS + NST + (size_t)N) * sizeof(double))
                    : nullptr;
    float *W32 = R ? (float *)ws_get(c, WS_MX_W32, trsm_ws_elems(N, NB) * sizeof(float)) : nullptr;
    if (!W32) return -3;
    // [0, NRHS): ||r_j||, [NRHS, 2 NRHS): ||x_j||, ||A||, then the status
    // words of k_mx_status
    double *nrm = R + (size_t)N * NRHS;
    double *rs = nrm + NST;                       // row sums of |A|
    // page-locked host copy of nrm, so that one asynchronous copy and one
    // synchronization per refinement step bring everything the host decides on
    if ((size_t)NST > c->h_stat_n) {
        if (c->h_stat) MI_CHECK(hipHostFree(c->h_stat));
        c->h_stat = nullptr;
        c->h_stat_n = 0;
        MI_CHECK(hipHostMalloc((void **)&c->h_stat, (size_t)NST * sizeof(double), hipHostMallocDefault));
        c->h_stat_n = (size_t)NST;
    }
    const double *h = c->h_stat;
    // device flags: d_info[0] the FP32 factorization's info, d_info[3] the
    // rounding of A overflowed, d_info[4] the rounding of B or of a residual
    int *a_ovf = c->d_info + 3, *r_ovf = c->d_info + 4;
    double anrm = 0.0, best = INFINITY;
    int stall = 0, h_ovf = 0;
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
    // A -> FP32 tiles (the lower triangle; the rest of the image is zeroed).
    // Nothing below waits for the host before the first residual: the
    // overflow of this rounding and the info of the FP32 factorization are
    // read there, with the first norms, and decided on in the order
    // dsposv does (rounding, then factorization).
    k_flag_clear<<<1, 1, 0, s>>>(a_ovf);
    {
        // four runs of four elements per thread and tile part
        const int per = (int)std::max<int64_t>(1, ((int64_t)NB * NB + 4095) / 4096);
        const int64_t nblk = (int64_t)MT * (MT + 1) / 2 * per;
        if (nblk > 0x7fffffff) { set_error("dlag2s grid too large"); rc = -3; goto done; }
        MI_LAUNCH(c, s, "dlag2s", 0.0, 12.0 * (double)ntile_el,
                  k_d2s_lower<<<(unsigned)nblk, 256, 0, s>>>(dA, L32, N, NB, MT, per, a_ovf));
    }
    {
        int sinfo = 0;
        rc = spotrf_tiles_split(c, N, NB, L32, &sinfo, prec == 1, false);
        if (rc) goto done;
    }
    {
        auto solve32 = [&](bool reset_done) -> int {
            int r;
            if ((r = trsm_enqueue<float>(c, s, 111, N, NB, L32, W32, NRHS, R32, N, reset_done))) return r;
            return trsm_enqueue<float>(c, s, 112, N, NB, L32, W32, NRHS, R32, N);
        };
        double cte = 0.0;                         // ||A||inf eps sqrt(N), from the first residual
        int it = 0;
        bool conv = false;
        if ((rc = trsm_begin(c, s))) goto done;
        if ((rc = trsm_prepare<float>(c, s, 0, N, NB, L32, W32))) goto done;
        // X0 = L32^-T L32^-1 B; the overflow flag covers the rounding of B
        // here and of every residual below (dlag2s: ITER = -2)
        k_flag_clear<<<1, 1, 0, s>>>(r_ovf);
        k_d2s_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(dB, ldb, R32, N, N, NRHS, r_ovf);
        if ((rc = solve32(false))) goto done;
        k_s2d_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(R32, N, dX, ldx, N, NRHS, 0);
        for (it = 0; it <= ITERMAX; it++) {
            // the first residual also yields the row sums of |A|, hence ||A||inf
            if ((rc = dsymm_resid(c, s, N, NB, dA, NRHS, dX, ldx, dB, ldb, R, N, it == 0 ? rs : nullptr)))
                goto done;
            colamax(s, R, N, N, NRHS, nrm);
            colamax(s, dX, ldx, N, NRHS, nrm + NRHS);
            if (it == 0) colamax(s, rs, N, N, 1, nrm + 2 * NRHS);
            k_mx_status<<<1, 1, 0, s>>>(nrm + 2 * NRHS + 1, c->d_info, a_ovf, r_ovf, trsm_abort_word(c));
            MX_CHECK(hipMemcpyAsync(c->h_stat, nrm, (size_t)NST * sizeof(double), hipMemcpyDeviceToHost, s), done);
            // the next correction's rounding and hand-off reset, queued
            // before the wait (unused when this step ends the iteration; the
            // status copied above does not include this rounding)
            if (it < ITERMAX) {
                k_d2s_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(R, N, R32, N, N, NRHS, r_ovf);
                if ((rc = trsm_reset<float>(c, s, N, NB))) goto done;
            }
            MX_CHECK(spin_sync(s), done);
            if (it == 0) {
                // decided in dsposv's order: the rounding of A, then the
                // FP32 factorization (what followed them is discarded)
                if (h[2 * NRHS + 1] != 0.0) { rc = fall_back(-2); goto done; }
                if (h[2 * NRHS + 3] != 0.0) { rc = fall_back(-3); goto done; }
            }
            if (h[2 * NRHS + 4] != 0.0) { rc = trsm_abort_error(); goto done; }
            h_ovf = h[2 * NRHS + 2] != 0.0;
            if (h_ovf) break;                     // B or a residual overflowed FP32
            if (it == 0) {
                anrm = h[2 * NRHS];
                cte = anrm * eps * std::sqrt((double)N);
            }
            double worst = 0.0;
            conv = true;
            for (int j = 0; j < NRHS; j++) {
                // non-finite norms never count as converged (inf <= inf)
                if (!(std::isfinite(h[j]) && std::isfinite(h[NRHS + j]) && h[j] <= h[NRHS + j] * cte))
                    conv = false;
                worst = std::isfinite(h[j]) ? std::max(worst, h[j]) : INFINITY;
            }
            if (conv || it == ITERMAX) break;
            // stall: the largest residual failed to halve twice in a row
            stall = (worst > 0.5 * best) ? stall + 1 : 0;
            best = std::min(best, worst);
            if (stall >= 2 || !(worst < INFINITY)) break;
            if ((rc = solve32(true))) goto done;
            k_s2d_cm<<<grid_of((int64_t)N * NRHS), 256, 0, s>>>(R32, N, dX, ldx, N, NRHS, 1);
        }
        if (h_ovf) {
            rc = fall_back(-2);
        } else if (conv) {
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
