// mi_solve.hip -- tile triangular solves on gfx950 behind dplasma_dtrsm /
// dplasma_dpotrs (Left, Lower; NoTrans = src/ztrsm_LLN.jdf, Trans =
// src/ztrsm_LLT.jdf).  The right-hand sides are column-major N x NRHS in HBM
// and are processed in chunks of up to 16 columns.
//
// Per tile step k (forward; backward mirrors it):
//   k_trsv<NR, false>  X_k = L_kk^{-1} B_k       (ztrsm task, LLN.jdf:29-60)
//   k_gemv_fwd<NR>     B_m -= L_mk X_k, m > k    (zgemm task, LLN.jdf:79-174)
// NR = 1 for a single right-hand side, 16 for a chunk of 2..16.
// The diagonal solve walks 32-row blocks and applies the inverse of each
// 32x32 diagonal block (computed for all k by k_inv_diag_blocks); the update
// streams each L tile from HBM exactly once per sweep, which is the bound of
// this HBM-bound phase.
#include <algorithm>

#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

constexpr int RHSB = 16;         // right-hand-side chunk
constexpr int TSV_THREADS = 512; // diagonal solve: one block, one row (or column) per thread
constexpr int TSV_MAX_NB = 960;  // two rows per thread, and the 16-RHS LDS budget
constexpr int GV_ROWS = 64;      // GEMV strip / column group

// Inverse of every 32x32 diagonal block of every diagonal tile.
// grid = MT * nblk, 32 threads (thread j solves L w = e_j).
// W(i, j) of block cb of tile k is at W[(k * nblk + cb) * 1024 + j * 32 + i].
template <class T>
__global__ __launch_bounds__(32) void k_inv_diag_blocks(const T *__restrict__ Tt, int N, int NB,
                                                       int MT, int unit, T *__restrict__ W)
{
    const int nblk = (NB + 31) / 32;
    const int k = blockIdx.x / nblk, cb = blockIdx.x % nblk;
    const int n = tile_rows(N, NB, MT, k);
    const int c0 = cb * 32;
    T *Wb = W + ((int64_t)k * nblk + cb) * 1024;
    const int j = threadIdx.x;
    if (c0 >= n) {                        // padding block: identity
        for (int i = 0; i < 32; i++) Wb[j * 32 + i] = (i == j) ? T(1) : T(0);
        return;
    }
    const int jw = min(32, n - c0);
    const T *L = Tt + tile_index(MT, k, k) * (int64_t)NB * NB + (int64_t)c0 * NB + c0;
    __shared__ T Ls[32][33];
    for (int c = 0; c < 32; c++) Ls[c][j] = (j < jw && c < jw) ? L[(int64_t)c * NB + j] : T(0);
    __syncthreads();
    T w[32];
#pragma unroll
    for (int i = 0; i < 32; i++) {
        if (i < jw) {
            T s = (i == j) ? T(1) : T(0);
#pragma unroll
            for (int p = 0; p < 32; p++)
                if (p < i) s -= Ls[p][i] * w[p];
            w[i] = (i >= j) ? (unit ? s : s / Ls[i][i]) : T(0);
        } else {
            w[i] = (i == j) ? T(1) : T(0);
        }
    }
#pragma unroll
    for (int i = 0; i < 32; i++) Wb[j * 32 + i] = w[i];
}

// Diagonal solve of tile k for nr <= NR right-hand sides, one block of 512
// threads.  The tile is walked in 32-row blocks cb; thread t owns rows
// t + 512 q (q < RPT) of the right-hand side: in registers for NR = 1, in
// LDS rows of stride NR + 1 for NR = 16.  Per block:
//   publish W = inv(L_cb,cb) (and, NR = 1, y of the block's rows) to LDS, sync,
//   x_cb = W y_cb (forward) or W^T y_cb (backward), written to B, sync,
//   every owner of a row still to be solved subtracts its 32-column strip of
//   L times x_cb.
// The strip of L and the next W block are loaded while the next block is
// published and solved (RPT = 1), so a block costs LDS traffic and two
// barriers rather than an HBM round trip.
// Forward (TRANS = false):  X = L_kk^{-1} B_k, rows r >= c0 + 32 use L(r, c0..c0+31).
// Backward (TRANS = true):  X = L_kk^{-T} B_k, walked from the last block;
//   rows c < c0 use L(c0..c0+31, c), which is column c of the tile (contiguous).
template <class T, int NR, bool TRANS, int RPT>
__global__ __launch_bounds__(TSV_THREADS) void k_trsv(const T *__restrict__ Tt, int N, int NB, int MT,
                                                      int k, const T *__restrict__ W,
                                                      T *__restrict__ B, int ldb, int nr)
{
    constexpr int YS = NR + 1;                            // LDS row stride of y (NR = 16)
    extern __shared__ __align__(16) unsigned char sm_raw[];
    T *Yl = reinterpret_cast<T *>(sm_raw);                        // [RPT * 512][YS] (NR = 16)
    __shared__ T Ys[32][NR];
    __shared__ T Xb[32][NR];
    __shared__ T Ws[32][33];                         // Ws[i][q] = W(i, q)
    const int tid = threadIdx.x;
    const int n = tile_rows(N, NB, MT, k);
    const int nblk = (NB + 31) / 32;
    const int ncb = (n + 31) / 32;
    const T *L = Tt + tile_index(MT, k, k) * (int64_t)NB * NB;
    const T *Wk = W + (int64_t)k * nblk * 1024;
    T *Bk = B + (int64_t)k * NB;

    T y[RPT];                                        // NR = 1
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = tid + q * TSV_THREADS;
        if (NR == 1) {
            y[q] = (r < n && nr > 0) ? Bk[r] : T(0);
        } else {
            for (int j = 0; j < NR; j++)
                Yl[r * YS + j] = (r < n && j < nr) ? Bk[(int64_t)j * ldb + r] : T(0);
        }
    }

    // strip of L used after block cb is solved, for the rows this thread owns;
    // branch-free: every load stays inside the tile, masked rows get 0
    T Lc[RPT][32];
    auto load_strip = [&](int cb) {
        const int c0 = cb * 32;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = tid + q * TSV_THREADS;
            const int rc = min(r, NB - 1);
            if (!TRANS) {
                const bool on = r >= c0 + 32 && r < n;
#pragma unroll
                for (int c = 0; c < 32; c++) {
                    const T v = L[(int64_t)min(c0 + c, NB - 1) * NB + rc];
                    Lc[q][c] = on ? v : T(0);
                }
            } else {
                const bool on = r < c0;
                const T *col = L + (int64_t)rc * NB;
#pragma unroll
                for (int p = 0; p < 32; p++) {
                    const T v = col[min(c0 + p, NB - 1)];
                    Lc[q][p] = (on && c0 + p < n) ? v : T(0);
                }
            }
        }
    };

    int cb = TRANS ? ncb - 1 : 0;
    load_strip(cb);
    T w0 = Wk[(int64_t)cb * 1024 + tid], w1 = Wk[(int64_t)cb * 1024 + tid + TSV_THREADS];

    for (int it = 0; it < ncb; it++) {
        const int c0 = cb * 32;
        const int nx = TRANS ? cb - 1 : cb + 1;
        // (a) publish this block's W (and right-hand side)
        if (NR == 1) {
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                const int r = tid + q * TSV_THREADS;
                if (r >= c0 && r < c0 + 32) Ys[r - c0][0] = y[q];
            }
        } else {
            // rows of this block were last updated by other threads in (d)
            __syncthreads();
            for (int idx = tid; idx < 32 * NR; idx += TSV_THREADS)
                Ys[idx / NR][idx % NR] = Yl[(c0 + idx / NR) * YS + idx % NR];
        }
        Ws[tid & 31][tid >> 5] = w0;                      // W(i, q) at q * 32 + i
        Ws[tid & 31][(tid >> 5) + 16] = w1;
        __syncthreads();
        // (b) the next block's W
        if (it + 1 < ncb) {
            w0 = Wk[(int64_t)nx * 1024 + tid];
            w1 = Wk[(int64_t)nx * 1024 + tid + TSV_THREADS];
        }
        // (c) x_cb = W y_cb (or W^T y_cb); rows past n have y = 0 and an identity W
        for (int idx = tid; idx < 32 * NR; idx += TSV_THREADS) {
            const int i = idx / NR, j = idx % NR;
            T s = T(0);
#pragma unroll
            for (int q = 0; q < 32; q++) s += (TRANS ? Ws[q][i] : Ws[i][q]) * Ys[q][j];
            Xb[i][j] = s;
            if (c0 + i < n && j < nr) Bk[(int64_t)j * ldb + c0 + i] = s;
        }
        __syncthreads();
        // (d) rows still to be solved subtract L strip * x_cb (zero strip elsewhere)
        if (NR == 1) {
#pragma unroll
            for (int c = 0; c < 32; c++) {
                const T xv = Xb[c][0];
#pragma unroll
                for (int q = 0; q < RPT; q++) y[q] -= Lc[q][c] * xv;
            }
        } else {
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                const int r = tid + q * TSV_THREADS;
#pragma unroll 1
                for (int j = 0; j < NR; j++) {
                    T s = T(0);
#pragma unroll
                    for (int c = 0; c < 32; c++) s += Lc[q][c] * Xb[c][j];
                    Yl[r * YS + j] -= s;
                }
            }
        }
        // the next strip loads while the next block is published and solved
        if (it + 1 < ncb) load_strip(nx);
        cb = nx;
    }
}

// Diagonal solve of tile k for 2..16 right-hand sides (NB <= 512) on MFMA.
// One block of 512 threads; wave w keeps y for the 16-row tiles g = w + 8t
// (t < 4) in MFMA accumulator layout (element q of lane l = y(16g +
// row(q, l/16), l % 16), the right-hand side in l % 16).  Per 32-row block:
//   the owners of its two row tiles publish y, W = inv(L_cb,cb) is staged,
//   x_cb = W y_cb (W^T y_cb backward) by all threads from LDS, written to B,
//   every row tile strictly past the block (forward: below, backward: above)
//   accumulates -L_strip x_cb with 8 MFMAs (16 rows x 32 x 16 right-hand sides),
//   its strip of L loaded while the block was published and solved.
// Forward: strip = L(rows, c0..c0+31); backward: strip = L(c0..c0+31, rows)^T.
template <class T, bool TRANS>
__global__ __launch_bounds__(TSV_THREADS) void k_trsv_mf(const T *__restrict__ Tt, int N, int NB, int MT, int k,
                                                         const T *__restrict__ W, T *__restrict__ B, int ldb,
                                                         int nr)
{
    typedef typename MF<T>::v4 V4;
    __shared__ T Ys[32][17];
    __shared__ T Xb[32][17];
    __shared__ T Ws[32][33];                          // Ws[i][q] = W(i, q)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const int n = tile_rows(N, NB, MT, k);
    const int nblk = (NB + 31) / 32;
    const int ncb = (n + 31) / 32;
    const T *L = Tt + tile_index(MT, k, k) * (int64_t)NB * NB;
    const T *Wk = W + (int64_t)k * nblk * 1024;
    T *Bk = B + (int64_t)k * NB;

    V4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int g = wave + 8 * t;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = 16 * g + MF<T>::row(q, fk);
            acc[t][q] = (r < n && fr < nr) ? Bk[(int64_t)fr * ldb + r] : T(0);
        }
    }
    // strip operands of the row tiles this wave updates after block cb:
    // A(i, kk) = -L(16g + i, c0 + kk) forward, -L(c0 + kk, 16g + i) backward
    T sa[4][8];
    auto load_strip = [&](int cb) {
        const int c0 = cb * 32;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int g = wave + 8 * t;
            const int r = 16 * g + fr;
            const bool on = TRANS ? (16 * g + 16 <= c0) : (16 * g >= c0 + 32);
            const int rc = min(r, NB - 1);
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int kk = c0 + 4 * s + fk;
                const int kc = min(kk, NB - 1);
                const T v = TRANS ? L[(int64_t)rc * NB + kc] : L[(int64_t)kc * NB + rc];
                sa[t][s] = (on && r < n && kk < n) ? -v : T(0);
            }
        }
    };

    int cb = TRANS ? ncb - 1 : 0;
    load_strip(cb);
    T w0 = Wk[(int64_t)cb * 1024 + tid], w1 = Wk[(int64_t)cb * 1024 + tid + TSV_THREADS];
    for (int it = 0; it < ncb; it++) {
        const int c0 = cb * 32;
        const int nx = TRANS ? cb - 1 : cb + 1;
        // (a) owners publish y of this block's two row tiles; W to LDS
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int g = wave + 8 * t;
            if (g == 2 * cb || g == 2 * cb + 1) {
#pragma unroll
                for (int q = 0; q < 4; q++) Ys[16 * (g - 2 * cb) + MF<T>::row(q, fk)][fr] = acc[t][q];
            }
        }
        Ws[tid & 31][tid >> 5] = w0;
        Ws[tid & 31][(tid >> 5) + 16] = w1;
        __syncthreads();
        if (it + 1 < ncb) {
            w0 = Wk[(int64_t)nx * 1024 + tid];
            w1 = Wk[(int64_t)nx * 1024 + tid + TSV_THREADS];
        }
        // (c) x = W y (forward) or W^T y (backward): thread = (row i, rhs j)
        {
            const int i = tid >> 4, j = tid & 15;
            T x = T(0);
#pragma unroll
            for (int q = 0; q < 32; q++) x += (TRANS ? Ws[q][i] : Ws[i][q]) * Ys[q][j];
            Xb[i][j] = x;
            if (c0 + i < n && j < nr) Bk[(int64_t)j * ldb + c0 + i] = x;
        }
        __syncthreads();
        // (d) row tiles past the block: acc -= L_strip x_cb
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int g = wave + 8 * t;
            const bool on = TRANS ? (16 * g + 16 <= c0) : (16 * g >= c0 + 32);
            if (on && 16 * g < n) {
#pragma unroll
                for (int s = 0; s < 8; s++) acc[t] = MF<T>::mfma(sa[t][s], Xb[4 * s + fk][fr], acc[t]);
            }
        }
        if (it + 1 < ncb) load_strip(nx);
        cb = nx;
    }
}

// Forward update B_m -= L_mk X_k for m = k+1 .. MT-1.
// Grid = (MT-k-1) * ceil(NB/64) strips of 64 rows; 4 waves split the NB
// columns of L_mk, lane = row, 16 independent loads in flight per lane;
// the four partial sums meet in LDS.  X_k is staged in LDS.
template <class T, int NR>
__global__ __launch_bounds__(256) void k_gemv_fwd(const T *__restrict__ Tt, int N, int NB, int MT,
                                                 int k, T *__restrict__ B, int ldb, int nr)
{
    extern __shared__ __align__(16) unsigned char sm_raw[];
    T *sm = reinterpret_cast<T *>(sm_raw);
    T *Xs = sm;                         // [NB][NR]
    T *red = sm + (size_t)NB * NR;      // [3][64][NR]
    const int strips = (NB + GV_ROWS - 1) / GV_ROWS;
    const int m = k + 1 + blockIdx.x / strips;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = (blockIdx.x % strips) * GV_ROWS + lane;
    const int rows_m = tile_rows(N, NB, MT, m);
    const T *Bk = B + (int64_t)k * NB;   // tile k is full (k < m)
    for (int idx = threadIdx.x; idx < NB * NR; idx += 256) {
        const int p = idx / NR, j = idx % NR;
        Xs[idx] = (j < nr) ? Bk[(int64_t)j * ldb + p] : T(0);
    }
    __syncthreads();
    const int pw = (NB + 3) / 4;
    const int p0 = wave * pw, p1 = min(NB, p0 + pw);
    T acc[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) acc[j] = T(0);
    if (r < rows_m) {
        const T *Lr = Tt + tile_index(MT, m, k) * (int64_t)NB * NB + r;
        int p = p0;
        for (; p + 16 <= p1; p += 16) {
            T l[16];
#pragma unroll
            for (int u = 0; u < 16; u++) l[u] = Lr[(int64_t)(p + u) * NB];
#pragma unroll
            for (int u = 0; u < 16; u++)
#pragma unroll
                for (int j = 0; j < NR; j++) acc[j] += l[u] * Xs[(p + u) * NR + j];
        }
        for (; p < p1; p++) {
            const T l = Lr[(int64_t)p * NB];
#pragma unroll
            for (int j = 0; j < NR; j++) acc[j] += l * Xs[p * NR + j];
        }
    }
    if (wave > 0)
#pragma unroll
        for (int j = 0; j < NR; j++) red[((wave - 1) * 64 + lane) * NR + j] = acc[j];
    __syncthreads();
    if (wave == 0 && r < rows_m) {
        T *Bm = B + (int64_t)m * NB + r;
#pragma unroll
        for (int j = 0; j < NR; j++) {
            const T s = acc[j] + red[lane * NR + j] + red[(64 + lane) * NR + j] + red[(128 + lane) * NR + j];
            if (j < nr) Bm[(int64_t)j * ldb] -= s;
        }
    }
}

// Backward update B_m -= L_km^T X_k for m = 0 .. k-1.  Grid = k tiles *
// ceil(NB/64) groups of 64 columns of L_km (= 64 rows of B_m).  Chunks of 64
// rows of L_km are read coalesced (lane = row), transposed through LDS and
// reduced per column; the next chunk is loaded while the current one is used.
//   NR == 1:  the 4 waves split each chunk's rows, partial sums meet in LDS;
//   NR == 16: wave w owns right-hand sides 4w .. 4w+3.
template <class T, int NR>
__global__ __launch_bounds__(256) void k_gemv_bwd(const T *__restrict__ Tt, int N, int NB, int MT,
                                                 int k, T *__restrict__ B, int ldb, int nr)
{
    extern __shared__ __align__(16) unsigned char sm_raw[];
    T *sm = reinterpret_cast<T *>(sm_raw);
    const int NBp = (NB + GV_ROWS - 1) / GV_ROWS * GV_ROWS;
    T *Xs = sm;                          // [NBp][NR]
    T *Ls = Xs + (size_t)NBp * NR;       // [64][65]: Ls[i * 65 + p]
    T *red = Ls + 64 * 65;               // [4][64]
    const int groups = NBp / GV_ROWS;
    const int m = blockIdx.x / groups;
    const int col0 = (blockIdx.x % groups) * GV_ROWS;
    const int ncols = min(GV_ROWS, NB - col0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rows_k = tile_rows(N, NB, MT, k);
    const T *Bk = B + (int64_t)k * NB;
    for (int idx = threadIdx.x; idx < NBp * NR; idx += 256) {
        const int p = idx / NR, j = idx % NR;
        Xs[idx] = (p < rows_k && j < nr) ? Bk[(int64_t)j * ldb + p] : T(0);
    }
    const T *L = Tt + tile_index(MT, k, m) * (int64_t)NB * NB + (int64_t)col0 * NB;
    T lr[16];
    auto load = [&](int p0) {
        const int p = p0 + lane;
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const int i = wave + 4 * s;
            lr[s] = (i < ncols && p < rows_k) ? L[(int64_t)i * NB + p] : T(0);
        }
    };
    load(0);
    constexpr int JPW = NR == 1 ? 1 : NR / 4;
    T acc[JPW];
#pragma unroll
    for (int j = 0; j < JPW; j++) acc[j] = T(0);
    for (int p0 = 0; p0 < rows_k; p0 += 64) {
#pragma unroll
        for (int s = 0; s < 16; s++) Ls[(wave + 4 * s) * 65 + lane] = lr[s];
        __syncthreads();
        if (p0 + 64 < rows_k) load(p0 + 64);
        if (NR == 1) {
#pragma unroll
            for (int pp = 0; pp < 16; pp++) {
                const int p = wave * 16 + pp;
                acc[0] += Ls[lane * 65 + p] * Xs[p0 + p];
            }
        } else {
#pragma unroll 4
            for (int p = 0; p < 64; p++) {
                const T l = Ls[lane * 65 + p];
#pragma unroll
                for (int jj = 0; jj < JPW; jj++) acc[jj] += l * Xs[(p0 + p) * NR + wave * JPW + jj];
            }
        }
        __syncthreads();
    }
    T *Bm = B + (int64_t)m * NB + col0 + lane;  // tile m is full (m < k)
    if (NR == 1) {
        red[wave * 64 + lane] = acc[0];
        __syncthreads();
        if (wave == 0 && lane < ncols) Bm[0] -= red[lane] + red[64 + lane] + red[128 + lane] + red[192 + lane];
    } else if (lane < ncols) {
#pragma unroll
        for (int jj = 0; jj < JPW; jj++) {
            const int j = wave * JPW + jj;
            if (j < nr) Bm[(int64_t)j * ldb] -= acc[jj];
        }
    }
}

// ---------------------------------------------------------------------------
// One-launch wavefront solve (up to 16 right-hand sides, NB % 32 == 0).
// Workgroup g (a logical index from an atomic ticket, so every lower index
// is already running) owns rows [128 g, 128 g + 128) of X, 32 rows per wave.
// Forward it subtracts L(rows, block h) X(block h) for every block h < g - 1
// as X(block h) appears; the L operands of the next block are loaded while
// the wave waits.  Then z = W_g y with W_g the inverse of the 128 x 128
// diagonal triangle (k_inv128, staged in LDS at the start), and, once the
// last block appears, x_g = z - M_g X(block g - 1) with the link matrix
// M_g = W_g L(g, g - 1) (k_wave_link), so the chain between two workgroups
// is one 128 x 128 x 16 product.  Backward mirrors it from the last block
// with L^T, W_g^T and M_g = W_g^T L(g + 1, g)^T.  Each L element is read
// once per sweep.  Every workgroup waits only on lower logical indices, so
// the launch always drains.
constexpr int WV_RB = 128;
constexpr int WV_WS = 136;   // LDS column stride of W_g (128 + 8: fragment reads in two passes)

template <class T>
__device__ inline const T *sym_elem(const T *Tt, int MT, int NB, int row, int col)
{
    const int m = row / NB, n = col / NB;
    return Tt + tile_index(MT, m, n) * (int64_t)NB * NB + (int64_t)(col - n * NB) * NB + (row - m * NB);
}

// Hand-off between workgroups: the staging copy xs of X lives in uncached
// device memory (every access reaches memory, so values stored on one XCD
// are visible on the others without cache maintenance) and is filled with a
// signalling NaN before the launch.  Arithmetic never produces a signalling
// NaN, so a staged value that differs from it is final: the producer stores
// its x values (single-copy atomic, one element per store) and the consumer
// polls the values themselves, with no separate flag and no wait for store
// completion.  Block g occupies xs + g * 2048 in the order given by wv_xs,
// so that every load instruction of a consumer covers whole cache lines.
template <class T> struct WvSent;
template <> struct WvSent<double> {
    typedef unsigned long long U;
    static constexpr U v = 0x7FF4A5A5A5A5A5A5ull;
};
template <> struct WvSent<float> {
    typedef unsigned int U;
    static constexpr U v = 0x7FA5A5A5u;
};

template <class T>
__device__ inline typename WvSent<T>::U wv_load(const T *p)
{
    typedef typename WvSent<T>::U U;
    const U *q = (const U *)p;
    asm volatile("" : "+v"(q));             // vector-memory access
    return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// a value equal to the sentinel (only possible if a signalling NaN came in
// unquieted) is stored as a quiet NaN, so that every wait ends
template <class T>
__device__ inline void wv_store(T *p, T v)
{
    typedef typename WvSent<T>::U U;
    U *q = (U *)p;
    asm volatile("" : "+v"(q));
    U b = __builtin_bit_cast(U, v);
    if (b == WvSent<T>::v) b |= (U)1 << (sizeof(U) == 8 ? 51 : 22);
    __hip_atomic_store(q, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inverse of each 128 x 128 diagonal triangle of L (column-major, 128 x 128
// per block g, at W128 + g * 16384) from the 32 x 32 inverses W32 of
// k_inv_diag_blocks by doubling: W21 = -W22 L21 W11, for 32 then 64 rows.
// Rows beyond N act as identity.  One workgroup per block.
template <class T>
__global__ __launch_bounds__(256) void k_inv128(const T *__restrict__ Tt, int N, int NB, int MT,
                                               const T *__restrict__ W32, T *__restrict__ W128)
{
    __shared__ T tmp[64 * 65];
    const int g = blockIdx.x, tid = threadIdx.x;
    const int r0 = g * WV_RB, nblk = NB / 32;
    T *Wg = W128 + (int64_t)g * 16384;
    auto L = [&](int i, int j) -> T {       // L(r0 + i, r0 + j), i > j, 0 beyond N
        return (r0 + i < N) ? *sym_elem(Tt, MT, NB, r0 + i, r0 + j) : T(0);
    };
    for (int e = tid; e < 16384; e += 256) {
        const int i = e & 127, j = e >> 7;
        const int bi = i >> 5, bj = j >> 5;
        T v = T(0);
        if (bi == bj) {
            const int rc = r0 + 32 * bi;
            if (rc < N) {
                const int tk = rc / NB, cb = (rc - tk * NB) / 32;
                v = W32[((int64_t)tk * nblk + cb) * 1024 + (j & 31) * 32 + (i & 31)];
            } else {
                v = (i == j) ? T(1) : T(0);
            }
        }
        Wg[j * 128 + i] = v;
    }
    __syncthreads();
    for (int h = 32; h < 128; h *= 2) {
        for (int p = 0; p < 128; p += 2 * h) {
            const int q = p + h;
            // tmp = L(q.., p..) W(p.., p..)   (h x h)
            for (int e = tid; e < h * h; e += 256) {
                const int i = e % h, j = e / h;
                T a = T(0);
                for (int k = j; k < h; k++) a += L(q + i, p + k) * Wg[(p + j) * 128 + p + k];
                tmp[j * 65 + i] = a;
            }
            __syncthreads();
            // W(q.., p..) = -W(q.., q..) tmp
            for (int e = tid; e < h * h; e += 256) {
                const int i = e % h, j = e / h;
                T a = T(0);
                for (int k = 0; k <= i; k++) a += Wg[(q + k) * 128 + q + i] * tmp[j * 65 + k];
                Wg[(p + j) * 128 + q + i] = -a;
            }
            __syncthreads();
        }
    }
}

// Link matrices M_g = op(W_g) op(L)(g, h), h the block solved just before g
// (g - 1 forward, g + 1 backward): 128 x 128 at M + g * 16384, column-major
// forward and row-major backward (the order in which the solve reads it).
// op(L)(g, h)(j, k) = L(128 g + j, 128 h + k) forward, L(128 h + k, 128 g + j)
// backward; rows beyond N are zero.  One workgroup per block; thread (k =
// tid % 128) accumulates rows i = tid / 128 + 2 e of column k.
template <class T, bool TRANS>
__global__ __launch_bounds__(256) void k_wave_link(const T *__restrict__ Tt, int N, int NB, int MT,
                                                  const T *__restrict__ W128, T *__restrict__ M)
{
    __shared__ T sw[128][33];     // op(W_g)(i, j0 + jj)
    __shared__ T sl[32][129];     // op(L)(j0 + jj, k)
    const int g = blockIdx.x, nwg = gridDim.x, tid = threadIdx.x;
    if (TRANS ? g + 1 >= nwg : g == 0) return;
    const int h = TRANS ? g + 1 : g - 1;
    const T *Wg = W128 + (int64_t)g * 16384;
    const int k = tid & 127, i0 = tid >> 7;
    T acc[64];
#pragma unroll
    for (int e = 0; e < 64; e++) acc[e] = T(0);
    for (int j0 = 0; j0 < 128; j0 += 32) {
        for (int e = tid; e < 128 * 32; e += 256) {
            const int i = e & 127, jj = e >> 7;
            sw[i][jj] = TRANS ? Wg[i * 128 + j0 + jj] : Wg[(j0 + jj) * 128 + i];
        }
        for (int e = tid; e < 32 * 128; e += 256) {
            const int jj = e & 31, kk = e >> 5;
            const int row = TRANS ? 128 * h + kk : 128 * g + j0 + jj;
            const int col = TRANS ? 128 * g + j0 + jj : 128 * h + kk;
            sl[jj][kk] = row < N ? *sym_elem(Tt, MT, NB, row, col) : T(0);
        }
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < 32; jj++) {
            const T l = sl[jj][k];
#pragma unroll
            for (int e = 0; e < 64; e++) acc[e] += sw[i0 + 2 * e][jj] * l;
        }
        __syncthreads();
    }
    T *Mg = M + (int64_t)g * 16384;            // column-major forward, row-major backward
#pragma unroll
    for (int e = 0; e < 64; e++) Mg[TRANS ? (i0 + 2 * e) * 128 + k : k * 128 + i0 + 2 * e] = acc[e];
}

// Index mappings of the wavefront solve.  The NW waves of a workgroup own
// 128 / NW rows each, as TPW = 8 / NW MFMA row tiles of 16 rows.  With two
// tiles (NW = 4), MFMA row i (0..15) of tile t (0, 1) is row wv_row(t, i) of
// the wave's rows: forward, two consecutive rows belong to one lane (2 i + t),
// so that the column-major L operands load as pairs.  With one tile (NW = 8)
// row i is row i.  Element kq (0..7) of lane group f (0..3) of a 32-column
// slice of the contraction is column wv_k(kq, f) of it: backward, four
// consecutive columns belong to one lane, so that the rows of L (the
// columns of L^T) load as runs of four.
template <bool TRANS, int TPW> __device__ inline int wv_row(int t, int i)
{
    return TPW == 1 ? i : (TRANS ? 16 * t + i : 2 * i + t);
}
template <bool TRANS> __device__ inline int wv_k(int kq, int f) { return TRANS ? 16 * (kq >> 2) + 4 * f + (kq & 3) : 4 * kq + f; }

// Four consecutive elements from p (16-byte aligned): one 16-byte load in
// FP32, two in FP64.
template <class T> __device__ inline void wv_load4(const T *p, T (&v)[4])
{
    if constexpr (sizeof(T) == 4) {
        const fx4 w = *(const fx4 *)p;
        v[0] = w[0]; v[1] = w[1]; v[2] = w[2]; v[3] = w[3];
    } else {
        const double2 a = *(const double2 *)p, b = *(const double2 *)(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
}

// Staging position of X(row R of the block, right-hand side j).  FP64: [row]
// [rhs], read one element per lane and load.  FP32: consumer order, two
// values per 64-bit load: the element kq of lane (j, f) of slice hb sits at
// ((hb * 4 + kq / 2) * 64 + 16 f + j) * 2 + kq % 2.
template <class T, bool TRANS> __device__ inline int wv_xs(int R, int j)
{
    if (sizeof(T) == 8) return R * RHSB + j;
    const int hb = R >> 5, r = R & 31;
    const int kq = TRANS ? 4 * (r >> 4) + (r & 3) : r >> 2;
    const int f = TRANS ? (r >> 2) & 3 : r & 3;
    return ((hb * 4 + (kq >> 1)) * 64 + 16 * f + j) * 2 + (kq & 1);
}

// Every wait is bounded: a consumer that has polled for more than
// wait_ticks of the 100 MHz real-time counter (0: gives up at once) stores 1
// into the abort word flags[1] and goes on with whatever it read; every other
// wait ends as soon as it sees that word set, so the launch always drains and
// the host reports the failure.
template <class T, bool TRANS, int NW>
__global__ __launch_bounds__(64 * NW, 1) void k_trsm_wave(const T *__restrict__ Tt, int N, int NB, int MT,
                                                         const T *__restrict__ W128, const T *__restrict__ Mlink,
                                                         T *__restrict__ B, int ldb, int nr, T *__restrict__ xs,
                                                         int *__restrict__ ticket, long long wait_ticks)
{
    typedef typename MF<T>::v4 V4;
    typedef typename MF<T>::v2 V2;
    typedef typename WvSent<T>::U U;
    constexpr int NT = 64 * NW;                   // threads
    constexpr int RW = WV_RB / NW;                // rows per wave
    constexpr int TPW = RW / 16;                  // MFMA row tiles per wave
    static_assert(TPW == 1 || TPW == 2, "16 or 32 rows per wave");
    __shared__ T Ws[128 * WV_WS];    // a-operands of the diagonal product: Ws[k * WV_WS + i] = op(W)(i, k)
    __shared__ T Ys[128][17];
    __shared__ int s_g;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const int nwg = (N + WV_RB - 1) / WV_RB;
    if (tid == 0) {
        int *tp = ticket;
        asm volatile("" : "+v"(tp));
        s_g = atomicAdd(tp, 1);
    }
    __syncthreads();
    const int g = TRANS ? nwg - 1 - s_g : s_g;
    const int r0 = g * WV_RB;
    const int rw = r0 + RW * wave;               // this wave's rows
    // Loads below always use in-bounds addresses (indices clamped into the
    // tile or matrix) and mask the value afterwards, so they issue together
    // instead of one branch and one wait each.
    {
        const T *Wg = W128 + (int64_t)g * 16384;
        for (int e0 = 0; e0 < 16384; e0 += NT * 16) {
            T v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int e = e0 + NT * u + tid, i = e & 127, k = e >> 7;
                v[u] = TRANS ? Wg[i * 128 + k] : Wg[k * 128 + i];
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int e = e0 + NT * u + tid, i = e & 127, k = e >> 7;
                Ws[k * WV_WS + i] = v[u];
            }
        }
    }
    const int frc = min(fr, max(nr, 1) - 1);
    // Catch-up layout: wave (rg, kh) = (wave / 2, wave % 2) takes rows
    // [32 rg, 32 rg + 32) of the block (two MFMA row tiles, rows
    // wv_row<TRANS, 2>) against one half of each earlier block's columns
    // (slices hb = 2 kh, 2 kh + 1): two rows per lane give 16-byte loads of L
    // forward, and each wave reads only its half of X.  The two halves of y
    // meet in LDS before z = W_g y, which (like the link product and the
    // stores) uses the layout of RW rows per wave.
    static_assert(NW == 8, "the catch-up layout pairs eight waves");
    const int rg = wave >> 1, kh = wave & 1;
    const int rc = r0 + 32 * rg;                 // catch-up rows of this wave
    V4 acc[2];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = rc + wv_row<TRANS, 2>(t, MF<T>::row(q, fk));
            const T v = B[(int64_t)frc * ldb + min(row, N - 1)];
            acc[t][q] = (kh == 0 && row < N && fr < nr) ? v : T(0);
        }

    // a-operands of one other block h: op(L)(row rc + wv_row(t, fr), column
    // 32 (2 kh + hl) + wv_k(kq, fk) of the block) as la[hl][kq][t]
    T la[2][8][2];
    // (the addresses are clamped into the matrix).  FP64 zeroes what lies
    // outside the matrix as the values arrive; FP32, whose loads run two
    // blocks ahead, leaves that to mask_a, once the values are needed.
    constexpr bool MASK_AT_LOAD = false;
    auto load_a = [&](T (&la)[2][8][2], int h) {
#pragma unroll
        for (int hl = 0; hl < 2; hl++) {
            const int c0 = h * WV_RB + 32 * (2 * kh + hl);
            const int cc = min(c0, N - 1) & ~31;     // a valid 32-aligned block when c0 >= N
            const int rr = min(rc, N - 1) & ~31;
            auto put = [&](int kq, int t, T v) {
                if (MASK_AT_LOAD) {
                    const int kk = wv_k<TRANS>(kq, fk), row = rc + wv_row<TRANS, 2>(t, fr);
                    v = (c0 < N && c0 + kk < N && row < N) ? v : T(0);
                }
                la[hl][kq][t] = v;
            };
            if (!TRANS) {
                const T *pA = sym_elem(Tt, MT, NB, rr, cc) + 2 * fr;
#pragma unroll
                for (int kq = 0; kq < 8; kq++) {
                    const V2 v = *(const V2 *)(pA + (int64_t)(4 * kq + fk) * NB);
                    put(kq, 0, v.x);
                    put(kq, 1, v.y);
                }
            } else {
                const T *pA = sym_elem(Tt, MT, NB, cc, rr) + 4 * fk;
#pragma unroll
                for (int t = 0; t < 2; t++)
#pragma unroll
                    for (int kg = 0; kg < 2; kg++) {
                        T v[4];
                        wv_load4(pA + (int64_t)(16 * t + fr) * NB + 16 * kg, v);
#pragma unroll
                        for (int e = 0; e < 4; e++) put(4 * kg + e, t, v[e]);
                    }
            }
        }
    };
    auto mask_a = [&](T (&la)[2][8][2], int h) {
        if ((h + 1) * WV_RB <= N && rc + 32 <= N) return;     // wave-uniform
#pragma unroll
        for (int hl = 0; hl < 2; hl++) {
            const int c0 = h * WV_RB + 32 * (2 * kh + hl);
#pragma unroll
            for (int kq = 0; kq < 8; kq++)
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int kk = wv_k<TRANS>(kq, fk), row = rc + wv_row<TRANS, 2>(t, fr);
                    if (!(c0 < N && c0 + kk < N && row < N)) la[hl][kq][t] = T(0);
                }
        }
    };
    // a-operands of the link matrix: M_g(RW w + wv_row(t, fr), 32 hb + wv_k(kq, fk)),
    // column-major forward, row-major backward
    T lm[4][8][TPW];
    auto load_m = [&]() {
        const T *Mg = Mlink + (int64_t)g * 16384;
#pragma unroll
        for (int hb = 0; hb < 4; hb++) {
            if (!TRANS) {
#pragma unroll
                for (int kq = 0; kq < 8; kq++) {
                    if constexpr (TPW == 2) {
                        const V2 v = *(const V2 *)(Mg + (32 * hb + 4 * kq + fk) * 128 + RW * wave + 2 * fr);
                        lm[hb][kq][0] = v.x;
                        lm[hb][kq][1] = v.y;
                    } else {
                        lm[hb][kq][0] = Mg[(32 * hb + 4 * kq + fk) * 128 + RW * wave + fr];
                    }
                }
            } else {
#pragma unroll
                for (int t = 0; t < TPW; t++)
#pragma unroll
                    for (int kg = 0; kg < 2; kg++) {
                        T v[4];
                        wv_load4(Mg + (RW * wave + 16 * t + fr) * 128 + 32 * hb + 16 * kg + 4 * fk, v);
#pragma unroll
                        for (int e = 0; e < 4; e++) lm[hb][4 * kg + e][t] = v[e];
                    }
            }
        }
    };
    // -X(block h), slices [hb0, hb0 + NH), as b-operands, once every value
    // has been stored: the wave loads its values, and loads them again while
    // any of them is still the sentinel (for the last block too: a cheaper
    // poll of one value per producer wave first costs one more round trip
    // per link, 1-2 % of a sweep)
    // true once this wait has to end without its values: the abort word is
    // set, or this wait outlived its budget (then it sets the word)
    int *abort_word = ticket + 1;
    // (global-memory pointers: a generic one would make every later wait
    // for a vector load a wait for all of them)
    typedef __attribute__((address_space(1))) int GI;
    auto give_up = [&](long long t_end) -> bool {
        GI *ap = (GI *)abort_word;
        asm volatile("" : "+v"(ap));
        if (__hip_atomic_load(ap, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return true;
        if (wait_ticks > 0 && (long long)__builtin_amdgcn_s_memrealtime() < t_end) return false;
        GI *wp = (GI *)abort_word;
        asm volatile("" : "+v"(wp));
        __hip_atomic_store(wp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    };
    auto wait_x = [&](int h, int hb0, auto &bx) {
        constexpr int NH = sizeof(bx) / sizeof(bx[0]);
        const T *xh = xs + (int64_t)h * (WV_RB * RHSB);
        const long long t_end = (long long)__builtin_amdgcn_s_memrealtime() + wait_ticks;
        if (wait_ticks == 0) (void)give_up(t_end);   // a zero budget ends every wait at once
        for (;;) {
            int ok = 1;
            if constexpr (sizeof(T) == 8) {
#pragma unroll
                for (int hl = 0; hl < NH; hl++)
#pragma unroll
                    for (int kq = 0; kq < 8; kq++) {
                        const U b = wv_load(xh + (32 * (hb0 + hl) + wv_k<TRANS>(kq, fk)) * RHSB + fr);
                        ok &= (b != WvSent<T>::v);
                        bx[hl][kq] = __builtin_bit_cast(T, b);
                    }
            } else {
#pragma unroll
                for (int hl = 0; hl < NH; hl++)
#pragma unroll
                    for (int kp = 0; kp < 4; kp++) {
                        const unsigned long long *p =
                            (const unsigned long long *)(xh + (((hb0 + hl) * 4 + kp) * 64 + lane) * 2);
                        asm volatile("" : "+v"(p));
                        const unsigned long long b = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
                        ok &= (lo != (unsigned)WvSent<float>::v) & (hi != (unsigned)WvSent<float>::v);
                        bx[hl][2 * kp] = __builtin_bit_cast(T, lo);
                        bx[hl][2 * kp + 1] = __builtin_bit_cast(T, hi);
                    }
            }
            if (__all(ok)) break;
            if (__any(give_up(t_end))) break;
            __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int hl = 0; hl < NH; hl++)
#pragma unroll
            for (int kq = 0; kq < 8; kq++) {
                const int r = h * WV_RB + 32 * (hb0 + hl) + wv_k<TRANS>(kq, fk);
                bx[hl][kq] = (fr < nr && r < N) ? -bx[hl][kq] : T(0);
            }
    };
    const int nprev = TRANS ? nwg - 1 - g : g;
    auto blk = [&](int s) { return TRANS ? nwg - 1 - s : s; };   // the s-th earlier block
    auto mfma_block = [&](const T (&l)[2][8][2], const T (&bx)[2][8]) {
#pragma unroll
        for (int hl = 0; hl < 2; hl++)
#pragma unroll
            for (int kq = 0; kq < 8; kq++)
#pragma unroll
                for (int t = 0; t < 2; t++) acc[t] = MF<T>::mfma(l[hl][kq][t], bx[hl][kq], acc[t]);
    };
    if constexpr (sizeof(T) == 4) {
        // Every earlier block but the last is long finished by the time it
        // is needed, so its X is read ahead like its L: the L operands and
        // one read of X of blocks s + 1 and s + 2 are issued before block s
        // is waited for and multiplied, three blocks in flight (the counter
        // of outstanding loads, at most 63, holds about two and a half).
        // They are issued after the loads of block s and vector-memory loads
        // complete in order, so waiting for block s leaves them outstanding.
        // The loads are unconditional (the last block is read again at the
        // end) so
        // that no register holding a load in flight is copied.  A read-ahead
        // X that still holds a sentinel is read again by the bounded wait.
        typedef unsigned long long X2[2][4];
        typedef __attribute__((address_space(1))) const unsigned long long GU;
        T lb[2][8][2];
        X2 xa, xb;
        auto read_x = [&](int h, X2 &xr) {
            const T *xh = xs + (int64_t)h * (WV_RB * RHSB);
#pragma unroll
            for (int hl = 0; hl < 2; hl++)
#pragma unroll
                for (int kp = 0; kp < 4; kp++) {
                    GU *p = (GU *)(xh + (((2 * kh + hl) * 4 + kp) * 64 + lane) * 2);
                    xr[hl][kp] = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
        };
        auto ready = [&](const X2 &xr) {
            int ok = 1;
#pragma unroll
            for (int hl = 0; hl < 2; hl++)
#pragma unroll
                for (int kp = 0; kp < 4; kp++)
                    ok &= ((unsigned)xr[hl][kp] != (unsigned)WvSent<float>::v) &
                          ((unsigned)(xr[hl][kp] >> 32) != (unsigned)WvSent<float>::v);
            return __all(ok);
        };
        auto consume = [&](int s, T (&lc)[2][8][2], X2 &xc) {
            if (!ready(xc)) {
                // the bounded wait of wait_x, reading into the same registers
                const long long t_end = (long long)__builtin_amdgcn_s_memrealtime() + wait_ticks;
                if (wait_ticks == 0) (void)give_up(t_end);
                for (;;) {
                    if (__any(give_up(t_end))) break;
                    __builtin_amdgcn_s_sleep(1);
                    read_x(blk(s), xc);
                    if (ready(xc)) break;
                }
            }
            T bx[2][8];
#pragma unroll
            for (int hl = 0; hl < 2; hl++)
#pragma unroll
                for (int kp = 0; kp < 4; kp++) {
                    bx[hl][2 * kp] = __builtin_bit_cast(T, (unsigned)xc[hl][kp]);
                    bx[hl][2 * kp + 1] = __builtin_bit_cast(T, (unsigned)(xc[hl][kp] >> 32));
                }
#pragma unroll
            for (int hl = 0; hl < 2; hl++)
#pragma unroll
                for (int kq = 0; kq < 8; kq++) {
                    const int r = blk(s) * WV_RB + 32 * (2 * kh + hl) + wv_k<TRANS>(kq, fk);
                    bx[hl][kq] = (fr < nr && r < N) ? -bx[hl][kq] : T(0);
                }
            mask_a(lc, blk(s));
            mfma_block(lc, bx);
        };
        if (nprev > 1 && rc < N) {
            const int last = nprev - 2;           // the last block of this loop
            T lc[2][8][2];
            X2 xc;
            load_a(la, blk(0));
            read_x(blk(0), xa);
            load_a(lb, blk(min(1, last)));
            read_x(blk(min(1, last)), xb);
            for (int s = 0;;) {
                load_a(lc, blk(min(s + 2, last)));
                read_x(blk(min(s + 2, last)), xc);
                consume(s, la, xa);
                if (++s > last) break;
                load_a(la, blk(min(s + 2, last)));
                read_x(blk(min(s + 2, last)), xa);
                consume(s, lb, xb);
                if (++s > last) break;
                load_a(lb, blk(min(s + 2, last)));
                read_x(blk(min(s + 2, last)), xb);
                consume(s, lc, xc);
                if (++s > last) break;
            }
        }
    } else {
        if (nprev > 1 && rc < N) load_a(la, blk(0));
        for (int s = 0; s + 1 < nprev; s++) {
            if (rc < N) {
                T bx[2][8];
                wait_x(blk(s), 2 * kh, bx);
                mask_a(la, blk(s));
                mfma_block(la, bx);
                if (s + 2 < nprev) load_a(la, blk(s + 1));
            }
        }
    }
    if (nprev > 0 && rw < N) load_m();

    // y = the two column halves summed (the second half added to the first
    // in LDS, one order for every run), then z = op(W_g) y
    if (kh == 1)
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) Ys[32 * rg + wv_row<TRANS, 2>(t, MF<T>::row(q, fk))][fr] = acc[t][q];
    __syncthreads();
    if (kh == 0)
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                T &y = Ys[32 * rg + wv_row<TRANS, 2>(t, MF<T>::row(q, fk))][fr];
                y = acc[t][q] + y;
            }
    __syncthreads();
    V4 x[TPW];
#pragma unroll
    for (int t = 0; t < TPW; t++) x[t] = V4{T(0), T(0), T(0), T(0)};
#pragma unroll 8
    for (int kq = 0; kq < 32; kq++) {
        const int kk = 4 * kq + fk;
        const T b = Ys[kk][fr];
#pragma unroll
        for (int t = 0; t < TPW; t++)
            x[t] = MF<T>::mfma(Ws[kk * WV_WS + RW * wave + wv_row<TRANS, TPW>(t, fr)], b, x[t]);
    }
    // x_g = z - M_g X(last block)
    if (nprev > 0 && rw < N) {
        T bx[4][8];
        wait_x(TRANS ? g + 1 : g - 1, 0, bx);
#pragma unroll
        for (int hb = 0; hb < 4; hb++)
#pragma unroll
            for (int kq = 0; kq < 8; kq++)
#pragma unroll
                for (int t = 0; t < TPW; t++) x[t] = MF<T>::mfma(lm[hb][kq][t], bx[hb][kq], x[t]);
    }
    T *xg = xs + (int64_t)g * (WV_RB * RHSB);
#pragma unroll
    for (int t = 0; t < TPW; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rb = RW * wave + wv_row<TRANS, TPW>(t, MF<T>::row(q, fk));
            wv_store(xg + wv_xs<T, TRANS>(rb, fr), x[t][q]);
            if (r0 + rb < N && fr < nr) B[(int64_t)fr * ldb + r0 + rb] = x[t][q];
        }
}

// Before each wavefront launch: staging filled with the sentinel, ticket 0
// (the abort word flags[1] is cleared once per solve, by k_wave_abort_clear).
template <class T>
__global__ void k_wave_reset(T *__restrict__ xs, int64_t n, int *__restrict__ ticket)
{
    typedef typename WvSent<T>::U U;
    U *q = (U *)xs;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        q[i] = WvSent<T>::v;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int *tp = ticket;
        asm volatile("" : "+v"(tp));          // vector store
        *tp = 0;
    }
}

template <class T>
__global__ void k_scale(T *__restrict__ B, int N, int nrhs, int ldb, T alpha)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < (int64_t)N * nrhs) {
        const int64_t c = i / N, r = i % N;
        B[c * ldb + r] *= alpha;
    }
}

__global__ void k_wave_abort_clear(int *flags)
{
    int *wp = flags + 1;
    asm volatile("" : "+v"(wp));           // vector store
    *wp = 0;
}

// The hand-off memory of the wavefront solve (the ticket and abort word, and
// the staging copy of X) is the context's uncached pair wv_ticket / wv_stage:
// every access reaches memory, so a value stored by one XCD is visible to
// the others without cache maintenance.
int g_wave_solve = 1;   // 1: one-launch wavefront solve when NB % 32 == 0; 0: per-tile launches
// bound on one hand-off wait of the wavefront solve, in ticks of the 100 MHz
// real-time counter (default 10 s; 0 = give up at once)
long long g_wave_wait_ticks = 1000000000LL;

int set_lds(const void *f, size_t bytes)
{
    if (bytes > 65536) MI_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return 0;
}

template <class T> struct SN;
template <> struct SN<double> {
    static constexpr const char *diag = "dtrsm_diag", *gemv = "dtrsm_gemv", *inv = "dtrsm_inv_diag",
                                *scale = "dtrsm_scale", *wave = "dtrsm_wave", *link = "dtrsm_link";
};
template <> struct SN<float> {
    static constexpr const char *diag = "strsm_diag", *gemv = "strsm_gemv", *inv = "strsm_inv_diag",
                                *scale = "strsm_scale", *wave = "strsm_wave", *link = "strsm_link";
};

template <class T, int NR>
int solve_chunk(Ctx *c, hipStream_t s, int trans, int N, int NB, int MT, const T *dT,
                const T *W, T *Bj, int ldb, int nr)
{
    const int strips = (NB + GV_ROWS - 1) / GV_ROWS;
    const size_t lds_f = ((size_t)NB * NR + 3 * 64 * NR) * sizeof(T);
    const size_t lds_b = ((size_t)strips * GV_ROWS * NR + 64 * 65 + 256) * sizeof(T);
    int rc, lim = 0;
    MI_CHECK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, 0));
    const size_t lds_t = (size_t)(NB <= TSV_THREADS ? 1 : 2) * TSV_THREADS * (NR + 1) * sizeof(T) +
                         (64 * NR + 32 * 33) * sizeof(T);   // diagonal solve, static part included
    const size_t need = std::max(std::max(lds_f, lds_b), NR == 1 ? (size_t)0 : lds_t);
    if (need > (size_t)lim) {
        set_error("tile size NB=%d needs %zu B of LDS for the solve (limit %d)", NB, need, lim);
        return -4;
    }
    // y rows of the diagonal solve (16 right-hand sides): RPT * 512 rows of NR + 1
    // (the one-RHS kernel keeps its strip and y in registers and exists for
    // NB <= 512 only; larger tiles take the 16-RHS kernel)
    constexpr int RPT2 = NR == 1 ? 1 : 2;
    const int rpt = NB <= TSV_THREADS ? 1 : 2;
    if (NR == 1 && rpt > 1) { set_error("one-RHS solve needs NB <= %d", TSV_THREADS); return -4; }
    const size_t lds_y = NR == 1 ? 0 : (size_t)rpt * TSV_THREADS * (NR + 1) * sizeof(T);
    if (lds_y) {
        if ((rc = set_lds(rpt == 1 ? (const void *)k_trsv<T, NR, false, 1> : (const void *)k_trsv<T, NR, false, RPT2>, lds_y))) return rc;
        if ((rc = set_lds(rpt == 1 ? (const void *)k_trsv<T, NR, true, 1> : (const void *)k_trsv<T, NR, true, RPT2>, lds_y))) return rc;
    }
    if ((rc = set_lds((const void *)k_gemv_fwd<T, NR>, lds_f))) return rc;
    if ((rc = set_lds((const void *)k_gemv_bwd<T, NR>, lds_b))) return rc;
    if (trans == 111) {
        for (int k = 0; k < MT; k++) {
            const int n = tile_rows(N, NB, MT, k);
            MI_LAUNCH(c, s, SN<T>::diag, (double)n * n * nr, (double)n * n * sizeof(T) / 2.0,
                      (NB > TSV_THREADS
                       ? k_trsv<T, NR, false, RPT2><<<1, TSV_THREADS, lds_y, s>>>(dT, N, NB, MT, k, W, Bj, ldb, nr)
                       : NR > 1
                       ? k_trsv_mf<T, false><<<1, TSV_THREADS, 0, s>>>(dT, N, NB, MT, k, W, Bj, ldb, nr)
                       : k_trsv<T, NR, false, 1><<<1, TSV_THREADS, lds_y, s>>>(dT, N, NB, MT, k, W, Bj, ldb, nr)));
            if (k + 1 < MT) {
                const double rows = (double)(N - (k + 1) * NB);
                MI_LAUNCH(c, s, SN<T>::gemv, 2.0 * rows * NB * nr, rows * NB * sizeof(T),
                          k_gemv_fwd<T, NR><<<(MT - k - 1) * strips, 256, lds_f, s>>>(dT, N, NB, MT, k, Bj,
                                                                                     ldb, nr));
            }
        }
    } else {
        for (int k = MT - 1; k >= 0; k--) {
            const int n = tile_rows(N, NB, MT, k);
            MI_LAUNCH(c, s, SN<T>::diag, (double)n * n * nr, (double)n * n * sizeof(T) / 2.0,
                      (NB > TSV_THREADS
                       ? k_trsv<T, NR, true, RPT2><<<1, TSV_THREADS, lds_y, s>>>(dT, N, NB, MT, k, W, Bj, ldb, nr)
                       : NR > 1
                       ? k_trsv_mf<T, true><<<1, TSV_THREADS, 0, s>>>(dT, N, NB, MT, k, W, Bj, ldb, nr)
                       : k_trsv<T, NR, true, 1><<<1, TSV_THREADS, lds_y, s>>>(dT, N, NB, MT, k, W, Bj, ldb, nr)));
            if (k > 0) {
                const double cols = (double)k * NB;
                MI_LAUNCH(c, s, SN<T>::gemv, 2.0 * cols * n * nr, cols * n * sizeof(T),
                          k_gemv_bwd<T, NR><<<k * strips, 256, lds_b, s>>>(dT, N, NB, MT, k, Bj, ldb, nr));
            }
        }
    }
    return 0;
}

}  // namespace

namespace mi {

size_t trsm_ws_elems(int N, int NB)
{
    const int MT = (N + NB - 1) / NB;
    const size_t w32 = (size_t)MT * ((NB + 31) / 32) * 1024;
    // plus the 128 x 128 inverses and the forward and backward link matrices
    // of the wavefront solve
    return w32 + (NB % 32 == 0 ? (size_t)((N + WV_RB - 1) / WV_RB) * 3 * 16384 : 0);
}

// Inverses of the 32x32 diagonal blocks of every diagonal tile into W
// (trsm_ws_elems elements), enqueued on s.
template <class T>
int trsm_prepare(Ctx *c, hipStream_t s, int unit, int N, int NB, const T *dT, T *W)
{
    const int MT = (N + NB - 1) / NB;
    const int nblk = (NB + 31) / 32;
    MI_LAUNCH(c, s, SN<T>::inv, (double)MT * NB * 32 * 32 / 3.0, (double)MT * NB * 32 * 2.0 * sizeof(T),
              k_inv_diag_blocks<T><<<MT * nblk, 32, 0, s>>>(dT, N, NB, MT, unit, W));
    if (NB % 32 == 0 && g_wave_solve) {
        const int nwg = (N + WV_RB - 1) / WV_RB;
        T *W128 = W + (size_t)MT * nblk * 1024;
        MI_LAUNCH(c, s, SN<T>::inv, (double)nwg * 128 * 128 * 128 / 3.0, (double)nwg * 16384 * 2.0 * sizeof(T),
                  k_inv128<T><<<nwg, 256, 0, s>>>(dT, N, NB, MT, W, W128));
        T *Mf = W128 + (size_t)nwg * 16384, *Mb = Mf + (size_t)nwg * 16384;
        MI_LAUNCH(c, s, SN<T>::link, (double)nwg * 128 * 128 * 128, (double)nwg * 16384 * 3.0 * sizeof(T),
                  k_wave_link<T, false><<<nwg, 256, 0, s>>>(dT, N, NB, MT, W128, Mf));
        MI_LAUNCH(c, s, SN<T>::link, (double)nwg * 128 * 128 * 128, (double)nwg * 16384 * 3.0 * sizeof(T),
                  k_wave_link<T, true><<<nwg, 256, 0, s>>>(dT, N, NB, MT, W128, Mb));
    }
    return 0;
}

// B <- op(L)^{-1} B for all NRHS columns, enqueued on s (W from trsm_prepare).
template <class T>
int trsm_enqueue(Ctx *c, hipStream_t s, int trans, int N, int NB, const T *dT, const T *W, int NRHS, T *dB,
                 int ldb, bool first_reset_done)
{
    const int MT = (N + NB - 1) / NB;
    if (NB % 32 == 0 && g_wave_solve) {
        // one wavefront launch per chunk of 16 right-hand sides
        const int nwg = (N + WV_RB - 1) / WV_RB;
        int rc;
        if ((rc = ws_uncached(c, c->wv_ticket, 2 * sizeof(int)))) return rc;
        if ((rc = ws_uncached(c, c->wv_stage, (size_t)nwg * WV_RB * RHSB * sizeof(T)))) return rc;
        int *ticket = (int *)c->wv_ticket.p;
        T *xs = (T *)c->wv_stage.p;
        const T *W128 = W + (size_t)MT * ((NB + 31) / 32) * 1024;
        const T *Mf = W128 + (size_t)nwg * 16384, *Mb = Mf + (size_t)nwg * 16384;
        const int64_t nst = (int64_t)nwg * WV_RB * RHSB;
        for (int j0 = 0; j0 < NRHS; j0 += RHSB) {
            const int nr = (NRHS - j0 < RHSB) ? NRHS - j0 : RHSB;
            T *Bj = dB + (int64_t)j0 * ldb;
            if (j0 > 0 || !first_reset_done)
                k_wave_reset<T><<<(unsigned)std::min<int64_t>((nst + 255) / 256, 1024), 256, 0, s>>>(xs, nst, ticket);
            MI_LAUNCH(c, s, SN<T>::wave, 1.0 * N * N * nr, 0.5 * N * (double)N * sizeof(T),
                      (trans == 111
                       ? k_trsm_wave<T, false, 8><<<nwg, 512, 0, s>>>(dT, N, NB, MT, W128, Mf, Bj, ldb, nr, xs,
                                                                      ticket, g_wave_wait_ticks)
                       : k_trsm_wave<T, true, 8><<<nwg, 512, 0, s>>>(dT, N, NB, MT, W128, Mb, Bj, ldb, nr, xs,
                                                                     ticket, g_wave_wait_ticks)));
        }
        return 0;
    }
    if (NB > TSV_MAX_NB) { set_error("tile size NB=%d above the solve limit %d", NB, TSV_MAX_NB); return -4; }
    for (int j0 = 0; j0 < NRHS; j0 += RHSB) {
        const int nr = (NRHS - j0 < RHSB) ? NRHS - j0 : RHSB;
        T *Bj = dB + (int64_t)j0 * ldb;
        int rc = (nr == 1 && NB <= TSV_THREADS) ? solve_chunk<T, 1>(c, s, trans, N, NB, MT, dT, W, Bj, ldb, nr)
                                                : solve_chunk<T, RHSB>(c, s, trans, N, NB, MT, dT, W, Bj, ldb, nr);
        if (rc) return rc;
    }
    return 0;
}

// The reset that trsm_enqueue does before its first chunk, enqueued ahead
// (no-op when the solve does not use the wavefront kernels).
template <class T>
int trsm_reset(Ctx *c, hipStream_t s, int N, int NB)
{
    if (!(NB % 32 == 0 && g_wave_solve)) return 0;
    const int nwg = (N + WV_RB - 1) / WV_RB;
    int rc;
    if ((rc = ws_uncached(c, c->wv_ticket, 2 * sizeof(int)))) return rc;
    if ((rc = ws_uncached(c, c->wv_stage, (size_t)nwg * WV_RB * RHSB * sizeof(T)))) return rc;
    const int64_t nst = (int64_t)nwg * WV_RB * RHSB;
    k_wave_reset<T><<<(unsigned)std::min<int64_t>((nst + 255) / 256, 1024), 256, 0, s>>>(
        (T *)c->wv_stage.p, nst, (int *)c->wv_ticket.p);
    MI_CHECK(hipGetLastError());
    return 0;
}

// Clears the abort word of the wavefront solve on s (before the first
// trsm_enqueue of a blocking call).
int trsm_begin(Ctx *c, hipStream_t s)
{
    int rc;
    if ((rc = ws_uncached(c, c->wv_ticket, 2 * sizeof(int)))) return rc;
    k_wave_abort_clear<<<1, 1, 0, s>>>((int *)c->wv_ticket.p);
    MI_CHECK(hipGetLastError());
    return 0;
}

// After the context's streams are synchronized: -5 when a wait of the
// wavefront solve gave up (its results are not valid), else 0.
int trsm_abort_error()
{
    set_error("triangular solve: a hand-off wait between workgroups exceeded its time limit");
    return -5;
}

int trsm_status(Ctx *c)
{
    int ab = 0;
    if (!c->wv_ticket.p) return 0;
    MI_CHECK(hipMemcpy(&ab, (int *)c->wv_ticket.p + 1, sizeof(int), hipMemcpyDeviceToHost));
    return ab ? trsm_abort_error() : 0;
}

const int *trsm_abort_word(Ctx *c) { return c->wv_ticket.p ? (const int *)c->wv_ticket.p + 1 : nullptr; }

template int trsm_prepare<double>(Ctx *, hipStream_t, int, int, int, const double *, double *);
template int trsm_prepare<float>(Ctx *, hipStream_t, int, int, int, const float *, float *);
template int trsm_enqueue<double>(Ctx *, hipStream_t, int, int, int, const double *, const double *, int,
                                  double *, int, bool);
template int trsm_enqueue<float>(Ctx *, hipStream_t, int, int, int, const float *, const float *, int, float *,
                                 int, bool);
template int trsm_reset<double>(Ctx *, hipStream_t, int, int);
template int trsm_reset<float>(Ctx *, hipStream_t, int, int);

}  // namespace mi

namespace {

// B <- alpha op_last(L)^{-1} ... op_first(L)^{-1} B for the nt sweeps
// trans[0..nt) (111 NoTrans / 112 Trans), with one preparation of the block
// inverses for all of them; blocking.
template <class T>
int trsm_run(mi_ctx_t *c, int nt, const int *trans, int diag, T alpha, int N, int NB, const T *dT, int NRHS,
             T *dB, int ldb)
{
    if (N <= 0 || NRHS <= 0) return 0;
    if (NB <= 0 || ldb < N) { set_error("bad sizes"); return -2; }
    if (NB > TSV_MAX_NB) { set_error("tile size NB=%d above the solve limit %d", NB, TSV_MAX_NB); return -4; }
    int rc;
    T *W = (T *)ws_get(c, WS_TRSM, trsm_ws_elems(N, NB) * sizeof(T));
    if (!W) return -3;
    hipStream_t s = c->s_main;
    if ((rc = trsm_begin(c, s))) return rc;
    MI_CHECK(hipEventRecord(c->t0, s));
    if (alpha != T(1)) {
        const int64_t tot = (int64_t)N * NRHS;
        MI_LAUNCH(c, s, SN<T>::scale, (double)tot, 2.0 * sizeof(T) * tot,
                  k_scale<T><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(dB, N, NRHS, ldb, alpha));
    }
    if ((rc = trsm_prepare<T>(c, s, diag == 132, N, NB, dT, W))) return rc;
    for (int i = 0; i < nt; i++)
        if ((rc = trsm_enqueue<T>(c, s, trans[i], N, NB, dT, W, NRHS, dB, ldb))) return rc;
    MI_CHECK(hipEventRecord(c->t1, s));
    if ((rc = mi_synchronize(c))) return rc;
    if ((rc = trsm_status(c))) return rc;
    float ms = 0.f;
    MI_CHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_ms = ms;
    return 0;
}

template <class T>
int trsm_LL(mi_ctx_t *c, int trans, int diag, T alpha, int N, int NB, const T *dT, int NRHS, T *dB, int ldb)
{
    return trsm_run<T>(c, 1, &trans, diag, alpha, N, NB, dT, NRHS, dB, ldb);
}

// L^{-T} L^{-1} B: both sweeps on one preparation (src/zpotrs_wrapper.c:91-94)
template <class T>
int potrs_tiles(mi_ctx_t *c, int N, int NB, const T *dT, int NRHS, T *dB, int ldb)
{
    const int tr[2] = {111, 112};
    return trsm_run<T>(c, 2, tr, 131, T(1), N, NB, dT, NRHS, dB, ldb);
}

}  // namespace

extern "C" int mi_dtrsm_LL(mi_ctx_t *c, int trans, int diag, double alpha, int N, int NB,
                           const double *dT, int NRHS, double *dB, int ldb)
{
    return trsm_LL<double>(c, trans, diag, alpha, N, NB, dT, NRHS, dB, ldb);
}

extern "C" int mi_strsm_LL(mi_ctx_t *c, int trans, int diag, float alpha, int N, int NB,
                           const float *dT, int NRHS, float *dB, int ldb)
{
    return trsm_LL<float>(c, trans, diag, alpha, N, NB, dT, NRHS, dB, ldb);
}

extern "C" int mi_dpotrs_tiles(mi_ctx_t *c, int N, int NB, const double *dT, int NRHS, double *dB, int ldb)
{
    return potrs_tiles<double>(c, N, NB, dT, NRHS, dB, ldb);
}

extern "C" int mi_spotrs_tiles(mi_ctx_t *c, int N, int NB, const float *dT, int NRHS, float *dB, int ldb)
{
    return potrs_tiles<float>(c, N, NB, dT, NRHS, dB, ldb);
}

// 1 (default): the one-launch wavefront solve whenever NB % 32 == 0;
// 0: one diagonal-solve and one update launch per tile.  Process-wide.
extern "C" int mi_set_solve_wait_limit(double seconds)
{
    if (!(seconds >= 0.0) || seconds > 3600.0) { set_error("wait limit must be in [0, 3600] s"); return -2; }
    g_wave_wait_ticks = (long long)(seconds * 1e8);
    return 0;
}

extern "C" int mi_set_solve_engine(int which)
{
    if (which != 0 && which != 1) { set_error("solve engine must be 0 or 1"); return -2; }
    g_wave_solve = which;
    return 0;
}
