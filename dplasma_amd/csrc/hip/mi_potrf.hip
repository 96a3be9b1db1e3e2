// mi_potrf.hip -- tile Cholesky (lower) on gfx950, tiles resident in HBM.
//
// The four task classes of src/zpotrf_L.jdf become three kernels launched
// per step k:
//   k_potrf_diag   POTRF(k)        :91-208   one workgroup factors the
//                                            diagonal tile with a 32-wide
//                                            right-looking blocking, the
//                                            current panel held in LDS; it also
//                                            writes the inverses of the 32x32
//                                            diagonal blocks for the panel solve
//                                            and records info = k*NB + j.
//   k_trsm_panel   TRSM(m,k), m>k  :214-343  C <- C L_kk^{-T}; one workgroup per
//                                            64-row strip of a panel tile,
//                                            blocked by 32 columns, diagonal
//                                            blocks applied through their
//                                            inverses.
//   k_update_mfma  SYRK(k,m) and   :349-633  C(m,n) -= A(m,k) A(n,k)^T for every
//                  GEMM(m,n,k)               trailing tile in a column range, on
//                                            v_mfma_f64_16x16x4_f64 with
//                                            double-buffered LDS staging; SYRK
//                                            is the m == n case restricted to
//                                            the lower triangle.
// The host loop runs the panel (POTRF+TRSM of step k+1) on a second stream
// as soon as column k+1 has received its step-k update, so it overlaps the
// rest of the step-k trailing update (one-panel look-ahead).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

constexpr int PJB = 32;            // panel width inside the diagonal factorization
constexpr int POTRF_THREADS = 512;

// v = *p, kept where c holds and +0 elsewhere; p must be a valid address
// whatever c is.  The load is unconditional, so that the compiler issues a
// group of them together instead of branching around each and waiting for
// it on its own.
template <class T> __device__ __forceinline__ T load_if(const T *p, bool c)
{
    const T v = *p;
    return c ? v : T(0);
}

// ---------------------------------------------------------------------------
// One wave: acc[ct] += X[r0+i][0:K] * Y[c0+16ct+j][0:K]^T for the 16x16 output
// tiles ct = 0, 1 (i, j < 16) on v_mfma_f64_16x16x4_f64, operands read straight
// from column-major global memory (element (r, k) at X[k*ldx + r]); rows at or
// beyond xr_lim / yr_lim read as zero.  K is a multiple of 32.  Operand lane
// layout: A(i, k) and B(k, j) sit in lane 16k + i (resp. 16k + j); result
// element q of lane l is D(4q + l/16, l%16).
template <class T>
__device__ inline void mfma_panel_2(const T *__restrict__ X, int ldx, int r0, int xr_lim,
                                    const T *__restrict__ Y, int ldy, int c0, int yr_lim, int K,
                                    typename MF<T>::v4 acc[2])
{
    typedef typename MF<T>::v4 V4;
    const int lane = threadIdx.x & 63, fr = lane & 15, fk = lane >> 4;
    const int xr = r0 + fr, y0 = c0 + fr, y1 = c0 + 16 + fr;
    const bool xv = xr < xr_lim, yv0 = y0 < yr_lim, yv1 = y1 < yr_lim;
    // rows past the limits read the last valid row (xr_lim, yr_lim >= 1)
    const int xrc = min(xr, xr_lim - 1), y0c = min(y0, yr_lim - 1), y1c = min(y1, yr_lim - 1);
    V4 e0 = {0, 0, 0, 0}, e1 = e0, o0 = e0, o1 = e0;
    // 32 columns of K per pass, all 24 loads of a pass issued together; in
    // FP32 the operands of pass p+1 are in flight while the 16 MFMAs of pass
    // p run
    T a[8], b0[8], b1[8];
    auto load = [&](int k, T *pa, T *pb0, T *pb1) {
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int64_t kk = k + 4 * s + fk;
            pa[s] = load_if(&X[kk * ldx + xrc], xv);
            pb0[s] = load_if(&Y[kk * ldy + y0c], yv0);
            pb1[s] = load_if(&Y[kk * ldy + y1c], yv1);
        }
    };
    auto mul = [&]() {
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
            e0 = MF<T>::mfma(a[s], b0[s], e0);
            e1 = MF<T>::mfma(a[s], b1[s], e1);
            o0 = MF<T>::mfma(a[s + 1], b0[s + 1], o0);
            o1 = MF<T>::mfma(a[s + 1], b1[s + 1], o1);
        }
    };
    if constexpr (sizeof(T) == 8) {
        // FP64: one chunk of 24 loads in flight (two would spill 45
        // registers in k_trsm_tile)
        for (int k = 0; k < K; k += 32) {
            load(k, a, b0, b1);
            mul();
        }
    } else {
        if (K > 0) load(0, a, b0, b1);
        for (int k = 0; k < K; k += 32) {
            T na[8], nb0[8], nb1[8];
            const bool more = k + 32 < K;
            if (more) load(k + 32, na, nb0, nb1);
            mul();
            if (more) {
#pragma unroll
                for (int s = 0; s < 8; s++) { a[s] = na[s]; b0[s] = nb0[s]; b1[s] = nb1[s]; }
            }
        }
    }
    acc[0] += e0 + o0;
    acc[1] += e1 + o1;
}


// C(r0:r0+32, c0:c0+32) -= X(r0:, 0:32) X(c0:, 0:32)^T on the lower triangle
// of A (rows, columns < n), X held in LDS as P[k * RC + row]; one wave, 2 x 2
// MFMA 16x16 tiles sharing their operands, K = 32.  The MFMA is fed (X_c, X_r)
// so that it accumulates C^T: lane l then holds rows r0 + 16i + l % 16 of C,
// and the loads and stores of the column-major C are 128-byte runs.
template <class T>
__device__ inline void syrk_unit32(T *__restrict__ A, int ld, int n, const T *__restrict__ P, int RC, int r0,
                                   int c0, int fr, int fk, T *__restrict__ dsm = nullptr)
{
    typedef typename MF<T>::v4 V4;
    V4 acc[2][2];                                    // acc[i][j]: rows r0+16i.., columns c0+16j..
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = r0 + 16 * i + fr, c = c0 + 16 * j + MF<T>::row(q, fk);
                acc[i][j][q] = load_if(&A[(int64_t)min(c, n - 1) * ld + min(r, n - 1)], r < n && c < n && r >= c);
            }
    const int ra0 = min(r0 + fr, RC - 1), ra1 = min(r0 + 16 + fr, RC - 1);
    const int cb0 = min(c0 + fr, RC - 1), cb1 = min(c0 + 16 + fr, RC - 1);
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int kk = (4 * s + fk) * RC;
        const T x0 = -P[kk + ra0], x1 = -P[kk + ra1];
        const T y0 = P[kk + cb0], y1 = P[kk + cb1];
        acc[0][0] = MF<T>::mfma(y0, x0, acc[0][0]);
        acc[0][1] = MF<T>::mfma(y1, x0, acc[0][1]);
        acc[1][0] = MF<T>::mfma(y0, x1, acc[1][0]);
        acc[1][1] = MF<T>::mfma(y1, x1, acc[1][1]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = r0 + 16 * i + fr, c = c0 + 16 * j + MF<T>::row(q, fk);
                if (r < n && c < n && r >= c) A[(int64_t)c * ld + r] = acc[i][j][q];
                // the unit also to LDS, column-major with stride 33 (its lower
                // part is what the diagonal factor reads)
                if (dsm) dsm[(c - c0) * 33 + (r - r0)] = acc[i][j][q];
            }
}

// POTRF(k): right-looking by 32-column blocks in one workgroup of 8 waves,
// with look-ahead inside the tile.  LDS holds the current block column X
// (rows j0..n-1, 32 columns) as P[c * RC + row] (absolute rows), W_jj, and
// the next diagonal block as its update left it (wave 0 updates that block
// itself and reads it back from LDS rather than from the tile).
// Per block j0:
//   wave 0     L_jj = chol(D), W_jj = L_jj^{-1} in one register sweep over the
//              32 columns (lanes 0..31: rows of L, lanes 32..63: columns of
//              W, both by the FMA v[c2] -= m * L(c2, c) with L(c2, c)
//              broadcast by v_readlane); columns past the tile edge get an
//              identity diagonal
//   waves 1-7  meanwhile: the trailing update of the previous block for
//              columns >= j0 + 32 (A -= X X^T, 32x32 MFMA units from LDS)
//   all        X = A(j0+32:n, j0:j0+32) W_jj^T on MFMA -> P and A;
//              the next block column (j0+32 .. j0+63) updated first,
//   wave 0     then loads the next diagonal block D while waves 1-7 go on
//              with the rest of the trailing update.
// Every element is updated in the same order as the right-looking loop, so
// results do not depend on timing.  W_jj goes to Winv for the panel TRSM.
// info = k*NB + j of the first non-positive pivot j (1-based), as
// CORE_dpotrf reports it (src/zpotrf_L.jdf:200-202).
template <class T>
__device__ inline void potrf_tile_dev(T *__restrict__ A, int n, int ld, int RC, int info_base,
                                      int *__restrict__ d_info, T *__restrict__ Winv)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typedef typename MF<T>::v4 V4;
    T *P = reinterpret_cast<T *>(smem_raw);          // 32 columns x RC rows (RC odd, >= n + 1)
    T *Ws = P + 32 * RC;                             // W(i, j) at Ws[j*32 + i]
    T *Dsm = Ws + 1024;                              // next diagonal block, column-major, stride 33
    __shared__ int s_fail;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const bool lrow = lane < 32;
    if (tid == 0) s_fail = 0;

    // v[c]: lanes < 32 hold D(lane, c) of the current diagonal block, lanes
    // >= 32 the identity column that becomes W(:, lane - 32)
    T v[32];
    // from_lds: the block as the next-column update left it in Dsm (lower
    // part; the strict upper part of v is never read)
    auto load_diag = [&](int j0, bool from_lds) {
        int j = lane - 32;
        asm volatile("" : "+v"(j));                  // built per block: keeps the identity out of LICM
        const int lr = min(lane, 31);
#pragma unroll
        for (int c = 0; c < 32; c++) {
            const int r = j0 + lane, cc = j0 + c;
            const T a = from_lds ? Dsm[c * 33 + lr]
                                 : load_if(&A[(int64_t)min(cc, n - 1) * ld + min(r, n - 1)], r < n && cc < n);
            v[c] = lrow ? ((r < n && cc < n) ? a : T(0)) : (c == j ? T(1) : T(0));
        }
    };
    if (wave == 0) load_diag(0, false);
    __syncthreads();

    for (int j0 = 0; j0 < n; j0 += 32) {
        const int jw = min(32, n - j0);
        if (wave == 0) {
            unsigned bad = 0;                        // non-positive pivots, bit c
#pragma unroll
            for (int c = 0; c < 32; c++) {
                T d0 = readlane_t(v[c], c);
                const bool ok = c < jw ? d0 > T(0) : true;
                bad |= (ok ? 0u : 1u) << c;
                d0 = ok && c < jw ? d0 : T(1);
                const T d = sqrt(d0);
                const T rd = T(1) / d;
                const T sc = v[c] * rd;              // L(lane, c) below the diagonal / W(c, j)
                const T m = (!lrow || lane > c) ? sc : T(0);
                v[c] = lrow ? (lane == c ? d : (lane > c ? sc : v[c])) : sc;
#pragma unroll
                for (int c2 = c + 1; c2 < 32; c2++) v[c2] -= m * readlane_t(m, c2);
            }
            if (lrow) {
#pragma unroll
                for (int c = 0; c < 32; c++)
                    if (c <= lane) P[c * RC + j0 + lane] = v[c];
            } else {
#pragma unroll
                for (int c = 0; c < 32; c++) Ws[(lane - 32) * 32 + c] = v[c];
            }
            if (bad && lane == 0) s_fail = j0 + __builtin_ctz(bad) + 1;
        } else if (j0 > 0) {
            // trailing update of block j0-32 beyond the current block column
            const int b0 = j0 + 32, nb = (n - b0 + 31) / 32;
            for (int u = wave - 1; u < nb * (nb + 1) / 2; u += 7) {
                int bi = 0, rem = u;                 // unit u -> (row block bi >= col block bj)
                while (rem > bi) { rem -= bi + 1; bi++; }
                syrk_unit32<T>(A, ld, n, P, RC, b0 + 32 * bi, b0 + 32 * rem, fr, fk);
            }
        }
        __syncthreads();
        if (s_fail) break;

        // X = A(j0+32:n, j0:j0+32) W^T on MFMA (K = 32), accumulated as
        // X^T = W P2^T so that lane l holds row rt + l % 16; rows to P and A
        for (int rt = j0 + 32 + 16 * wave; rt < n; rt += 128) {
            V4 e0 = {T(0), T(0), T(0), T(0)}, e1 = e0;
            const int r = rt + fr;
            const int rc = min(r, n - 1);
#pragma unroll
            for (int s4 = 0; s4 < 8; s4++) {
                const int kk = 4 * s4 + fk;
                const T a = load_if(&A[(int64_t)(j0 + min(kk, jw - 1)) * ld + rc], r < n && kk < jw);
                e0 = MF<T>::mfma(Ws[kk * 32 + fr], a, e0);          // W(c, kk), c = fr
                e1 = MF<T>::mfma(Ws[kk * 32 + 16 + fr], a, e1);     // c = 16 + fr
            }
            if (r < n) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int c = MF<T>::row(q, fk);
                    P[c * RC + r] = e0[q];
                    P[(16 + c) * RC + r] = e1[q];
                    if (c < jw) A[(int64_t)(j0 + c) * ld + r] = e0[q];
                    if (16 + c < jw) A[(int64_t)(j0 + 16 + c) * ld + r] = e1[q];
                }
            }
        }
        // L_jj to A, W_jj to Winv
        for (int idx = tid; idx < 1024; idx += 512) {
            const int c = idx >> 5, r = idx & 31;
            if (r >= c && r < jw && c < jw) A[(int64_t)(j0 + c) * ld + j0 + r] = P[c * RC + j0 + r];
            Winv[(j0 >> 5) * 1024 + idx] = Ws[idx];
        }
        __syncthreads();

        // the next block column first (rows j0+32.., columns j0+32..j0+63)
        const int b1 = j0 + 32;
        if (b1 < n) {
            for (int u = wave; u < (n - b1 + 31) / 32; u += 8)
                syrk_unit32<T>(A, ld, n, P, RC, b1 + 32 * u, b1, fr, fk, u == 0 ? Dsm : nullptr);
        }
        __syncthreads();
        if (wave == 0 && b1 < n) load_diag(b1, true);
    }
    if (tid == 0 && s_fail) atomicCAS(d_info, 0, info_base + s_fail);
}

template <class T>
__global__ __launch_bounds__(512) void k_potrf_tile(T *__restrict__ A, int n, int ld, int RC,
                                                   int info_base, int *__restrict__ d_info,
                                                   T *__restrict__ Winv)
{
    potrf_tile_dev<T>(A, n, ld, RC, info_base, d_info, Winv);
}

// ---------------------------------------------------------------------------
// TRSM(m,k) for m = k+1 .. MT-1: C <- C L^{-T}, one workgroup (4 waves) per
// 64-row strip of a tile, 32 columns at a time:
//   P = C[:, j0:j0+32] - X[:, 0:j0] L[j0:j0+32, 0:j0]^T   (MFMA, wave w: rows 16w..)
//   X[:, j0:j0+32] = P W_jj^T                              (W from POTRF(k))
constexpr int TS_ROWS = 64;
template <class T>
__global__ __launch_bounds__(256, 2) void k_trsm_tile(T *__restrict__ Tt, int N, int NB, int MT, int k,
                                                  const T *__restrict__ Winv, int m0)
{
    __shared__ T P[32][TS_ROWS + 1];
    __shared__ T Ws[1024];
    typedef typename MF<T>::v4 V4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const int strips = (NB + TS_ROWS - 1) / TS_ROWS;
    const int m = m0 + blockIdx.x / strips;
    const int r0 = (blockIdx.x % strips) * TS_ROWS;
    const int rows_m = tile_rows(N, NB, MT, m);
    if (r0 >= rows_m) return;
    const int rv = min(TS_ROWS, rows_m - r0);
    T *C = Tt + tile_index(MT, m, k) * (int64_t)NB * NB + r0;
    const T *L = Tt + tile_index(MT, k, k) * (int64_t)NB * NB;

    for (int j0 = 0; j0 < NB; j0 += 32) {
        const int jw = min(32, NB - j0);
        // this block's columns of C (written only by this step) and W_jj do
        // not depend on the solve so far: in flight during the product
        T cb[2][4], wv[4];
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int rr = wave * 16 + MF<T>::row(q, fk), c = ct * 16 + fr;
                cb[ct][q] = load_if(&C[(int64_t)(j0 + min(c, jw - 1)) * NB + min(rr, rv - 1)], rr < rv && c < jw);
            }
#pragma unroll
        for (int u = 0; u < 4; u++) wv[u] = Winv[(j0 >> 5) * 1024 + tid + 256 * u];
        V4 acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        if (j0 > 0) mfma_panel_2(C, NB, wave * 16, rv, L + j0, NB, 0, jw, j0, acc);
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int rr = wave * 16 + MF<T>::row(q, fk), c = ct * 16 + fr;
                P[c][rr] = (rr < rv && c < jw) ? cb[ct][q] - acc[ct][q] : T(0);
            }
#pragma unroll
        for (int u = 0; u < 4; u++) Ws[tid + 256 * u] = wv[u];
        __syncthreads();
        // X = P W_jj^T on MFMA (W_jj is zero above its diagonal), fed (W, P)
        // so that element q of lane l holds row 16 w + l % 16 and column
        // 16 ct + row(q, l / 16): the stores walk rows
        {
            V4 xo[2] = {{T(0), T(0), T(0), T(0)}, {T(0), T(0), T(0), T(0)}};
#pragma unroll
            for (int s4 = 0; s4 < 8; s4++) {
                const int kk = 4 * s4 + fk;
                const T a = P[kk][wave * 16 + fr];
#pragma unroll
                for (int ct = 0; ct < 2; ct++) xo[ct] = MF<T>::mfma(Ws[kk * 32 + ct * 16 + fr], a, xo[ct]);
            }
            const int rr = wave * 16 + fr;
#pragma unroll
            for (int ct = 0; ct < 2; ct++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int c = ct * 16 + MF<T>::row(q, fk);
                    if (rr < rv && c < jw) C[(int64_t)(j0 + c) * NB + rr] = xo[ct][q];
                }
        }
        __syncthreads();
    }
}

// TRSM(m,k) for few panel tiles (latency-bound steps): one workgroup per
// 16-row strip; the left-looking product X[:, 0:j0] L[j0:j0+32, 0:j0]^T of
// each 32-column block is split over the 4 waves by K (contiguous 32-column
// slices, partial sums met in LDS), so the per-block chain is a quarter of
// the 64-row kernel's; then P = C - sum and X = P W_jj^T on MFMA.  Bounded
// to two workgroups per CU (256 registers a lane) without spills.
constexpr int TSK_ROWS = 16;
template <class T>
__global__ __launch_bounds__(256, 2) void k_trsm_tile_ks(T *__restrict__ Tt, int N, int NB, int MT, int k,
                                                     const T *__restrict__ Winv, int m0)
{
    __shared__ T red[4][32][TSK_ROWS + 1];       // partial products [wave][col][row]
    __shared__ T P[32][TSK_ROWS + 1];
    __shared__ T Ws[1024];
    typedef typename MF<T>::v4 V4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const int strips = (NB + TSK_ROWS - 1) / TSK_ROWS;
    const int m = m0 + blockIdx.x / strips;
    const int r0 = (blockIdx.x % strips) * TSK_ROWS;
    const int rows_m = tile_rows(N, NB, MT, m);
    if (r0 >= rows_m) return;
    const int rv = min(TSK_ROWS, rows_m - r0);
    T *C = Tt + tile_index(MT, m, k) * (int64_t)NB * NB + r0;
    const T *L = Tt + tile_index(MT, k, k) * (int64_t)NB * NB;

    for (int j0 = 0; j0 < NB; j0 += 32) {
        const int jw = min(32, NB - j0);
        // this block's columns of C (written only by this step) and W_jj do
        // not depend on the solve so far: in flight during the product
        T cb[2], wv[4];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int idx = tid + 256 * u, r = idx % TSK_ROWS, c = idx / TSK_ROWS;
            cb[u] = load_if(&C[(int64_t)(j0 + min(c, jw - 1)) * NB + min(r, rv - 1)], r < rv && c < jw);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) wv[u] = Winv[(j0 >> 5) * 1024 + tid + 256 * u];
        // this wave's K slice of the j0 solved columns
        const int np = j0 / 32;
        const int p0 = (np * wave) / 4, p1 = (np * (wave + 1)) / 4;
        V4 acc[2] = {{T(0), T(0), T(0), T(0)}, {T(0), T(0), T(0), T(0)}};
        if (p1 > p0)
            mfma_panel_2(C + (int64_t)(32 * p0) * NB, NB, 0, rv, L + (int64_t)(32 * p0) * NB + j0, NB, 0, jw,
                         32 * (p1 - p0), acc);
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int q = 0; q < 4; q++) red[wave][ct * 16 + fr][MF<T>::row(q, fk)] = acc[ct][q];
#pragma unroll
        for (int u = 0; u < 4; u++) Ws[tid + 256 * u] = wv[u];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int idx = tid + 256 * u, r = idx % TSK_ROWS, c = idx / TSK_ROWS;
            const T sum = red[0][c][r] + red[1][c][r] + red[2][c][r] + red[3][c][r];
            P[c][r] = (r < rv && c < jw) ? cb[u] - sum : T(0);
        }
        __syncthreads();
        // X = P W_jj^T on MFMA (W_jj is zero above its diagonal): waves 0
        // and 1 take 16 columns each, fed (W, P) so that the stores walk rows
        if (wave < 2) {
            V4 xo = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int s4 = 0; s4 < 8; s4++) {
                const int kk = 4 * s4 + fk;
                xo = MF<T>::mfma(Ws[kk * 32 + wave * 16 + fr], P[kk][fr], xo);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int c = wave * 16 + MF<T>::row(q, fk);
                if (fr < rv && c < jw) C[(int64_t)(j0 + c) * NB + fr] = xo[q];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// POTRF(k): factor the n x n diagonal tile A (ld) in place, lower.
// LDS holds the current panel P (rows j0..n-1, PJB columns) column-major with
// row stride RC.  Winv receives, for every 32-column block cb, the inverse of
// the lower-triangular 32x32 diagonal block, column-major 32x32 at
// Winv + cb*1024 (identity-padded past n).
__global__ __launch_bounds__(POTRF_THREADS) void k_potrf_diag(double *__restrict__ A, int n, int ld,
                                                             int RC, int info_base,
                                                             int *__restrict__ d_info,
                                                             double *__restrict__ Winv)
{
    extern __shared__ double smem[];
    double *P = smem;                    // PJB * RC
    double *Wl = smem + PJB * RC;        // 32 x 32 inverse being built
    __shared__ int s_fail;
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (tid == 0) s_fail = 0;
    __syncthreads();

    for (int j0 = 0; j0 < n; j0 += PJB) {
        const int jw = min(PJB, n - j0);
        const int R = n - j0;
        for (int idx = tid; idx < PJB * R; idx += nthr) {
            int c = idx / R, r = idx - c * R;
            P[c * RC + r] = (c < jw) ? A[(int64_t)(j0 + c) * ld + j0 + r] : 0.0;
        }
        __syncthreads();

        // unblocked Cholesky of the jw x jw diagonal block by wave 0 (lane = row)
        if (tid < 64) {
            const int lane = tid;
            for (int c = 0; c < jw; c++) {
                double d = P[c * RC + c];
                if (!(d > 0.0)) {
                    if (lane == 0) s_fail = j0 + c + 1;
                    break;
                }
                d = sqrt(d);
                double lr = 0.0;
                if (lane > c && lane < jw) {
                    lr = P[c * RC + lane] / d;
                    P[c * RC + lane] = lr;
                }
                if (lane == c) P[c * RC + c] = d;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (lane > c && lane < jw)
                    for (int c2 = c + 1; c2 <= lane; c2++) P[c2 * RC + lane] -= lr * P[c * RC + c2];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        __syncthreads();
        if (s_fail) break;

        // inverse of the diagonal block: lane j < 32 solves L w = e_j
        if (tid < 32) {
            const int j = tid;
            double w[PJB];
#pragma unroll
            for (int i = 0; i < PJB; i++) w[i] = 0.0;
#pragma unroll
            for (int i = 0; i < PJB; i++) {
                if (i < jw) {
                    double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
                    for (int p = 0; p < PJB; p++)
                        if (p < i) s -= P[p * RC + i] * w[p];
                    w[i] = (i >= j) ? s / P[i * RC + i] : 0.0;
                } else {
                    w[i] = (i == j) ? 1.0 : 0.0;
                }
            }
#pragma unroll
            for (int i = 0; i < PJB; i++) Wl[j * 32 + i] = w[i];
        }
        __syncthreads();
        for (int idx = tid; idx < 1024; idx += nthr) Winv[(j0 / PJB) * 1024 + idx] = Wl[idx];

        // panel rows below the diagonal block: x = b L11^{-T}, i.e. x = b W^T
        for (int r = jw + tid; r < R; r += nthr) {
            double b[PJB];
#pragma unroll
            for (int c = 0; c < PJB; c++) b[c] = P[c * RC + r];
#pragma unroll
            for (int c = 0; c < PJB; c++) {
                double s = 0.0;
#pragma unroll
                for (int p = 0; p < PJB; p++)
                    if (p <= c) s += b[p] * Wl[p * 32 + c];
                P[c * RC + r] = (c < jw) ? s : 0.0;
            }
        }
        __syncthreads();
        for (int idx = tid; idx < jw * R; idx += nthr) {
            int c = idx / R, r = idx - c * R;
            if (r >= c) A[(int64_t)(j0 + c) * ld + j0 + r] = P[c * RC + r];
        }

        // trailing update inside the tile: A22 -= P2 P2^T (lower), 4x4 per thread
        const int M2 = R - jw;
        if (M2 > 0) {
            const int T4 = (M2 + 3) >> 2;
            const int ntile = T4 * (T4 + 1) / 2;
            for (int t = tid; t < ntile; t += nthr) {
                int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
                while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
                while (ti * (ti + 1) / 2 > t) ti--;
                const int tj = t - ti * (ti + 1) / 2;
                const int r0 = jw + 4 * ti, c0 = jw + 4 * tj;
                double acc[4][4];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
                for (int p = 0; p < jw; p++) {
                    const double *pc = P + p * RC;
                    double a[4], bb[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) { a[i] = pc[r0 + i]; bb[i] = pc[c0 + i]; }
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++) acc[i][j] += a[i] * bb[j];
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int cc = c0 + j;
                    if (cc >= R) continue;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int rr = r0 + i;
                        if (rr < R && rr >= cc) A[(int64_t)(j0 + cc) * ld + j0 + rr] -= acc[i][j];
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0 && s_fail) atomicCAS(d_info, 0, info_base + s_fail);
}

// ---------------------------------------------------------------------------
// TRSM(m,k) for m = k+1 .. MT-1: C <- C L^{-T}.  Grid: (MT-k-1) tiles x
// ceil(rows/64) strips.  L = diagonal tile (ld NB), Winv = its 32x32
// diagonal-block inverses.  256 threads: row = t % 64, column group g = t / 64
// owns 8 of the 32 columns of the current block.
constexpr int TRSM_ROWS = 64;
__global__ __launch_bounds__(256) void k_trsm_panel(double *__restrict__ T, int N, int NB, int MT, int k,
                                                   const double *__restrict__ Winv, int m0)
{
    __shared__ double Xs[32][TRSM_ROWS + 1];   // X chunk, [p][row]
    __shared__ double Ls[32][33];              // L chunk, [p][col]
    __shared__ double Ys[TRSM_ROWS][33];       // block right-hand side [row][col]
    __shared__ double Ws[32][33];              // inverse of the diagonal block [col][row]
    const int tid = threadIdx.x;
    const int row = tid & 63, g = tid >> 6;
    const int strips = (NB + TRSM_ROWS - 1) / TRSM_ROWS;
    const int m = m0 + blockIdx.x / strips;
    const int strip = blockIdx.x % strips;
    const int rows_m = tile_rows(N, NB, MT, m);
    const int r0 = strip * TRSM_ROWS;
    if (r0 >= rows_m) return;
    const int rv = min(TRSM_ROWS, rows_m - r0);
    double *C = T + tile_index(MT, m, k) * (int64_t)NB * NB + r0;     // strip origin
    const double *L = T + tile_index(MT, k, k) * (int64_t)NB * NB;

    for (int c0 = 0; c0 < NB; c0 += 32) {
        const int cw = min(32, NB - c0);
        double acc[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = g * 8 + j;
            acc[j] = (row < rv && c < cw) ? C[(int64_t)(c0 + c) * NB + row] : 0.0;
        }
        for (int p0 = 0; p0 < c0; p0 += 32) {
            __syncthreads();
            for (int idx = tid; idx < 32 * TRSM_ROWS; idx += 256) {
                int p = idx / TRSM_ROWS, r = idx % TRSM_ROWS;
                Xs[p][r] = (r < rv) ? C[(int64_t)(p0 + p) * NB + r] : 0.0;
            }
            for (int idx = tid; idx < 32 * 32; idx += 256) {
                int p = idx / 32, c = idx % 32;
                Ls[p][c] = (c < cw) ? L[(int64_t)(p0 + p) * NB + c0 + c] : 0.0;
            }
            __syncthreads();
#pragma unroll 8
            for (int p = 0; p < 32; p++) {
                const double x = Xs[p][row];
#pragma unroll
                for (int j = 0; j < 8; j++) acc[j] -= x * Ls[p][g * 8 + j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; j++) Ys[row][g * 8 + j] = acc[j];
        for (int idx = tid; idx < 1024; idx += 256) {
            int col = idx / 32, r = idx % 32;
            Ws[col][r] = Winv[(c0 / 32) * 1024 + idx];   // W(r, col)
        }
        __syncthreads();
        // X(row, c) = sum_{p<=c} Y(row, p) W(c, p)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = g * 8 + j;
            double s = 0.0;
            for (int p = 0; p <= c; p++) s += Ys[row][p] * Ws[p][c];
            if (row < rv && c < cw) C[(int64_t)(c0 + c) * NB + row] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Trailing update on MFMA: C(m,n) -= A(m,k) A(n,k)^T for tile columns n in
// [n_begin, n_end), rows m in [n, MT).  One workgroup = one BMX x BMX block of
// one tile (BMX = 128, or 64 for launches too small to fill the chip); four
// waves, each a BMX/2 quadrant as (BMX/32)^2 MFMA 16x16 tiles.
constexpr int BM = 128, BK = 16;

// workgroup L of an update launch -> tile (m, n) and its 128x128 block
// (bm, bn); tiles enumerated column by column from column n_begin, whose
// first m_off tiles (rows n_begin .. n_begin+m_off-1) are skipped; the grid
// size decides how many tiles a launch covers
__device__ inline void decode_block(int L, int nsub, int n_begin, int m_off, int MT, int *m, int *n,
                                    int *bm, int *bn)
{
    const int per_tile = nsub * nsub;
    int t = L / per_tile;
    const int s = L - t * per_tile;
    int nn = n_begin, m0 = n_begin + m_off;
    while (t >= MT - m0) { t -= MT - m0; nn++; m0 = nn; }
    *n = nn;
    *m = m0 + t;
    *bm = s / nsub;
    *bn = s - (*bm) * nsub;
}

// The FP64 MFMA pipe of a gfx950 SIMD issues at full rate only with two
// waves ready on it (one wave alone issues at half rate).  The workgroup's
// waves form a WR x WC grid over its BMX x BMX block, each wave a
// (BMX/WR) x (BMX/WC) sub-block of 16x16 MFMA tiles.  Two workgroups share
// each CU: WR x WC = 2 x 2 holds 256 registers per lane (two waves per
// SIMD); 4 x 2 holds 128 (four waves per SIMD), so that a workgroup waiting
// at a barrier or on memory still leaves two waves per SIMD issuing.
template <class T, int BMX, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC, 2) void k_update_mfma(T *__restrict__ Tt, int N, int NB, int MT, int k,
                                                                int n_begin, int m_off, int nsub)
{
    // rows padded by 128 bytes: the four lane groups of a fragment read (k
    // rows 4s + l/16) then fall on alternating halves of the LDS banks
    constexpr int LP = BMX + 128 / (int)sizeof(T);
    constexpr int NT = 64 * WR * WC;                  // threads
    __shared__ T As[2][BK][LP];
    __shared__ T Bs[2][BK][LP];
    typedef typename MF<T>::v4 V4;
    typedef typename MF<T>::v2 V2;
    constexpr int HR = BMX / WR, HC = BMX / WC;       // rows / columns per wave
    constexpr int TR = HR / 16, TC = HC / 16;         // 16x16 MFMA tiles per wave
    constexpr int PR = BMX / 2;                       // element pairs per k row of a panel
    constexpr int SP = BK * PR / NT;                  // pairs staged per thread and operand
    static_assert(TR % 2 == 0 && SP * NT == BK * PR, "wave grid does not fit the block");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WC, wn = wave % WC;

    // XCD-aware remap: consecutive workgroup ids land on different XCDs
    // (round-robin), so give each XCD a contiguous range of blocks; the
    // blocks of one tile then share their A/B panels in that XCD's L2.
    const int G = gridDim.x, bid = blockIdx.x;
    const int G8 = (G / 8) * 8;
    const int L = (bid < G8) ? (bid % 8) * (G8 / 8) + bid / 8 : bid;

    int m, n, bm, bn;
    decode_block(L, nsub, n_begin, m_off, MT, &m, &n, &bm, &bn);
    if (m == n && bm < bn) return;                 // strictly upper blocks of a diagonal tile

    const int rows_m = tile_rows(N, NB, MT, m);
    const int cols_n = tile_rows(N, NB, MT, n);
    const int r_base = bm * BMX, c_base = bn * BMX;
    if (r_base >= rows_m || c_base >= cols_n) return;
    const bool full = (r_base + BMX <= rows_m) && (c_base + BMX <= cols_n) && (NB % BK == 0) && !(NB & 1);

    const T *Ap = Tt + tile_index(MT, m, k) * (int64_t)NB * NB;
    const T *Bp = Tt + tile_index(MT, n, k) * (int64_t)NB * NB;
    T *Cp = Tt + tile_index(MT, m, n) * (int64_t)NB * NB;

    V4 acc[TR][TC];
#pragma unroll
    for (int i = 0; i < TR; i++)
#pragma unroll
        for (int j = 0; j < TC; j++) acc[i][j] = V4{T(0), T(0), T(0), T(0)};

    // global -> register staging in element pairs: pair e = tid + NT*i,
    // k row kk = e / PR, rows 2*(e % PR) and 2*(e % PR) + 1
    V2 ra[SP], rb[SP];
    auto gload = [&](int k0) {
        if (full) {
#pragma unroll
            for (int i = 0; i < SP; i++) {
                const int e = tid + NT * i;
                const int kk = e / PR, r = (e % PR) * 2;
                ra[i] = *(const V2 *)(Ap + (int64_t)(k0 + kk) * NB + r_base + r);
                rb[i] = *(const V2 *)(Bp + (int64_t)(k0 + kk) * NB + c_base + r);
            }
        } else {
#pragma unroll
            for (int i = 0; i < SP; i++) {
                const int e = tid + NT * i;
                const int kk = e / PR, r = (e % PR) * 2;
                const int kg = k0 + kk;
                T a0 = T(0), a1 = T(0), b0 = T(0), b1 = T(0);
                if (kg < NB) {
                    const T *ac = Ap + (int64_t)kg * NB;
                    const T *bc = Bp + (int64_t)kg * NB;
                    if (r_base + r < rows_m) a0 = ac[r_base + r];
                    if (r_base + r + 1 < rows_m) a1 = ac[r_base + r + 1];
                    if (c_base + r < cols_n) b0 = bc[c_base + r];
                    if (c_base + r + 1 < cols_n) b1 = bc[c_base + r + 1];
                }
                ra[i] = V2{a0, a1};
                rb[i] = V2{b0, b1};
            }
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < SP; i++) {
            const int e = tid + NT * i;
            const int kk = e / PR, r = (e % PR) * 2;
            *(V2 *)&As[buf][kk][r] = V2{-ra[i].x, -ra[i].y};     // -A: the MFMA accumulates C - A B^T
            *(V2 *)&Bs[buf][kk][r] = rb[i];
        }
    };

    // MFMA row tiles 2p and 2p+1 of a wave interleave their rows: row i of
    // tile 2p+t is row 32p + 2i + t of the wave's rows, so that a lane holds
    // two consecutive rows of one column and C moves as element pairs
    // full off-diagonal blocks start from C (loaded with the first operands)
    // and end with a plain store; the others accumulate -A B^T from zero
    const int fr0 = lane & 15, fk0 = lane >> 4;
    const bool pre = full && !((m == n) && (bm == bn));
    if (pre) {
#pragma unroll
        for (int p = 0; p < TR / 2; p++)
#pragma unroll
            for (int j = 0; j < TC; j++) {
                const int row = r_base + wm * HR + 32 * p + 2 * fr0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const V2 c = *(const V2 *)&Cp[(int64_t)(c_base + wn * HC + j * 16 + MF<T>::row(q, fk0)) * NB + row];
                    acc[2 * p][j][q] = c.x;
                    acc[2 * p + 1][j][q] = c.y;
                }
            }
    }
    const int nstage = (NB + BK - 1) / BK;
    gload(0);
    lstore(0);
    __syncthreads();
    const int fr = lane & 15, fk = lane >> 4;
    for (int st = 0; st < nstage; st++) {
        const int buf = st & 1;
        if (st + 1 < nstage) gload((st + 1) * BK);
#pragma unroll
        for (int kq = 0; kq < BK / 4; kq++) {
            T a[TR], b[TC];
#pragma unroll
            for (int p = 0; p < TR / 2; p++) {
                const V2 v = *(const V2 *)&As[buf][kq * 4 + fk][wm * HR + 32 * p + 2 * fr];
                a[2 * p] = v.x;
                a[2 * p + 1] = v.y;
            }
#pragma unroll
            for (int j = 0; j < TC; j++) b[j] = Bs[buf][kq * 4 + fk][wn * HC + j * 16 + fr];
#pragma unroll
            for (int i = 0; i < TR; i++)
#pragma unroll
                for (int j = 0; j < TC; j++)
                    acc[i][j] = MF<T>::mfma(b[j], a[i], acc[i][j]);   // C^T: lanes walk rows
        }
        if (st + 1 < nstage) lstore(buf ^ 1);
        __syncthreads();
    }

    // epilogue: the MFMA was fed (B, A), so element q of lane l holds row
    // 2 (l % 16) + t (tile 2p + t) and column row(q, l / 16) of each 16x16
    // tile: a store walks row pairs
    const bool diag = (m == n) && (bm == bn);
#pragma unroll
    for (int p = 0; p < TR / 2; p++) {
        const int rr = r_base + wm * HR + 32 * p + 2 * fr;
#pragma unroll
        for (int j = 0; j < TC; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int col = c_base + wn * HC + j * 16 + MF<T>::row(q, fk);
                if (pre) {
                    *(V2 *)&Cp[(int64_t)col * NB + rr] = V2{acc[2 * p][j][q], acc[2 * p + 1][j][q]};
                } else {
#pragma unroll
                    for (int t = 0; t < 2; t++) {
                        const int r = rr + t;
                        if (r < rows_m && col < cols_n && (!diag || r >= col)) Cp[(int64_t)col * NB + r] += acc[2 * p + t][j][q];
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// FP32 trailing update on BF16 MFMA with a two-term split of both operands
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): with a = ah + al, ah =
// bf16(a), al = bf16(a - ah), each product is taken as
//     al bh + ah bl + ah bh            (the al bl term, ~2^-16 relative, dropped)
// so the update carries about 16 significant bits per operand instead of 8,
// at three BF16 MFMAs per product.  Used by the mixed-precision solve, whose
// FP64 refinement recovers the rest.  Work decomposition, XCD grouping,
// SYRK masking and the C-preload epilogue as k_update_mfma; K in chunks of
// 32, split once per workgroup while staged into LDS as [row][k] bf16 rows
// (row pitch 40 bf16 = 80 B: 16-byte aligned fragment reads, spread banks).
// MFMA operand layout: A(i, k) and B(k, j) for lane l in i (resp. j) = l % 16,
// k = 8 (l / 16) + e, e < 8; result element q at row 4 (l / 16) + q.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
constexpr int XK = 32, XS = XK + 8;

// Epilogue of the BF16-split updates: the TI x 4 accumulator tiles of a wave
// (rows r0 + 16 i + 4 fk + q, columns c0 + 16 j + fr) stored as C (pre:
// C was the MFMA's start; 16-byte stores of four rows) or added to C with
// the tile and lower-triangle masks (diag: a diagonal block).
template <int TI>
__device__ inline void bf16x3_store(float *Cp, int NB, const fx4 (&acc)[TI][4], int r0, int c0, int fr, int fk,
                                    bool pre, int rows_m, int cols_n, bool diag)
{
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = r0 + i * 16 + 4 * fk, col = c0 + j * 16 + fr;
            float *c = &Cp[(int64_t)col * NB + row];
            if (pre) {
                *(fx4 *)c = acc[i][j];
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (row + q < rows_m && col < cols_n && (!diag || row + q >= col)) c[q] += acc[i][j][q];
            }
        }
}

__global__ __launch_bounds__(256, 2) void k_update_bf16x3(float *__restrict__ Tt, int N, int NB, int MT, int k,
                                                         int n_begin, int m_off, int nsub)
{
    __shared__ __align__(16) __bf16 Ah[BM * XS];
    __shared__ __align__(16) __bf16 Al[BM * XS];
    __shared__ __align__(16) __bf16 Bh[BM * XS];
    __shared__ __align__(16) __bf16 Bl[BM * XS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int G = gridDim.x, bid = blockIdx.x;
    const int G8 = (G / 8) * 8;
    const int L = (bid < G8) ? (bid % 8) * (G8 / 8) + bid / 8 : bid;
    int m, n, bm, bn;
    decode_block(L, nsub, n_begin, m_off, MT, &m, &n, &bm, &bn);
    if (m == n && bm < bn) return;
    const int rows_m = tile_rows(N, NB, MT, m);
    const int cols_n = tile_rows(N, NB, MT, n);
    const int r_base = bm * BM, c_base = bn * BM;
    if (r_base >= rows_m || c_base >= cols_n) return;
    const bool full = (r_base + BM <= rows_m) && (c_base + BM <= cols_n) && (NB % XK == 0);
    const float *Ap = Tt + tile_index(MT, m, k) * (int64_t)NB * NB;
    const float *Bp = Tt + tile_index(MT, n, k) * (int64_t)NB * NB;
    float *Cp = Tt + tile_index(MT, m, n) * (int64_t)NB * NB;
    const int fr = lane & 15, fk = lane >> 4;

    // element q of acc[i][j] in lane (fr, fk): C(row 4 fk + q of row tile i,
    // column fr of column tile j), four consecutive rows of one column
    fx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};
    const bool pre = full && !((m == n) && (bm == bn));
    if (pre) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = *(const fx4 *)&Cp[(int64_t)(c_base + wn * 64 + j * 16 + fr) * NB + r_base + wm * 64 +
                                              i * 16 + 4 * fk];
    }

    // staging: thread (rq, kq) = (tid % 32, tid / 32) holds rows rq + 32 i
    // (i < 4) at k = k0 + 4 kq + u (u < 4) of both operands
    const int rq = tid & 31, kq = tid >> 5;
    float ra[4][4], rb[4][4];                    // [i][u]
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int kg = k0 + 4 * kq + u;
            const float *ac = Ap + (int64_t)kg * NB + r_base + rq;
            const float *bc = Bp + (int64_t)kg * NB + c_base + rq;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (full) {
                    ra[i][u] = ac[32 * i];
                    rb[i][u] = bc[32 * i];
                } else {
                    const bool kin = kg < NB;
                    ra[i][u] = (kin && r_base + rq + 32 * i < rows_m) ? ac[32 * i] : 0.f;
                    rb[i][u] = (kin && c_base + rq + 32 * i < cols_n) ? bc[32 * i] : 0.f;
                }
            }
        }
    };
    auto split_store = [&](__bf16 *H, __bf16 *Lo, const float (&v)[4], int row, float sign) {
        bf4 h, l;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const float x = sign * v[u];
            const __bf16 xh = (__bf16)x;
            h[u] = xh;
            l[u] = (__bf16)(x - (float)xh);
        }
        *(bf4 *)&H[row * XS + 4 * kq] = h;
        *(bf4 *)&Lo[row * XS + 4 * kq] = l;
    };
    auto lstore = [&]() {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            split_store(Ah, Al, ra[i], rq + 32 * i, -1.f);   // -A: the MFMA accumulates C - A B^T
            split_store(Bh, Bl, rb[i], rq + 32 * i, 1.f);
        }
    };

    const int nstage = (NB + XK - 1) / XK;
    gload(0);
    for (int st = 0; st < nstage; st++) {
        lstore();
        __syncthreads();
        if (st + 1 < nstage) gload((st + 1) * XK);
        bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = wm * 64 + i * 16 + fr;
            ah[i] = *(const bf8 *)&Ah[row * XS + 8 * fk];
            al[i] = *(const bf8 *)&Al[row * XS + 8 * fk];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = wn * 64 + j * 16 + fr;
            bh[j] = *(const bf8 *)&Bh[row * XS + 8 * fk];
            bl[j] = *(const bf8 *)&Bl[row * XS + 8 * fk];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        __syncthreads();
    }
    bf16x3_store(Cp, NB, acc, r_base + wm * 64, c_base + wn * 64, fr, fk, pre, rows_m, cols_n, (m == n) && (bm == bn));
}

// ---------------------------------------------------------------------------
// The BF16 two-term split of panel column k, once per step instead of once
// per workgroup and K chunk of the update: tile (m, k), m >= m0, is written
// to P as its hi and lo BF16 images (NB x NB each; rows past the tile's
// last row are zero) in chunk-major order [k / 32][row][k % 32], so that
// the update's staging of one 32-wide K chunk of 128 rows reads 8 KB of
// consecutive bytes (whole cache lines) and does no conversion work.  One
// workgroup per 64 x 64 block of a tile, transposed through LDS.
constexpr int SPT = 64;
__global__ __launch_bounds__(256) void k_split_panel(const float *__restrict__ Tt, int N, int NB, int MT, int k,
                                                     int m0, __bf16 *__restrict__ P)
{
    __shared__ float sb[SPT][SPT + 1];
    const int nb = (NB + SPT - 1) / SPT;
    const int per = nb * nb;
    const int t = blockIdx.x / per, s = blockIdx.x - t * per;
    const int m = m0 + t;
    const int bk = s / nb, br = s - bk * nb;       // block of K columns, block of rows
    const int rows = tile_rows(N, NB, MT, m);
    const float *A = Tt + tile_index(MT, m, k) * (int64_t)NB * NB;
    __bf16 *H = P + (int64_t)t * 2 * NB * NB;
    __bf16 *Lo = H + (int64_t)NB * NB;
    const int tid = threadIdx.x;
    for (int e = tid; e < SPT * SPT; e += 256) {   // coalesced along the rows of a column
        const int j = e / SPT, i = e % SPT;
        const int kk = bk * SPT + j, r = br * SPT + i;
        sb[j][i] = (kk < NB && r < rows) ? A[(int64_t)kk * NB + r] : 0.f;
    }
    __syncthreads();
    for (int e = tid; e < SPT * SPT / 2; e += 256) {   // pairs of K, coalesced along a row
        const int i = e / (SPT / 2), j = 2 * (e % (SPT / 2));
        const int kk = bk * SPT + j, r = br * SPT + i;
        if (r < NB && kk < NB) {                   // NB is even here (a multiple of 32)
            const float x0 = sb[j][i], x1 = sb[j + 1][i];
            const __bf16 h0 = (__bf16)x0, h1 = (__bf16)x1;
            typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
            const int64_t o = ((int64_t)(kk / XK) * NB + r) * XK + kk % XK;
            *(bf2 *)&H[o] = bf2{h0, h1};
            *(bf2 *)&Lo[o] = bf2{(__bf16)(x0 - (float)h0), (__bf16)(x1 - (float)h1)};
        }
    }
}

// k_update_bf16x3 on the split panel P (tiles m0.. of column k, as written
// by k_split_panel; NB a multiple of 32).  The A halves are negated on
// their way into LDS (a sign-bit flip: the split of -a is minus the split
// of a), so the MFMAs see exactly the operands of k_update_bf16x3 and the
// result is the same bit for bit.  Staging: thread t holds rows t / 4 and
// t / 4 + 64 at K offset 8 (t % 4) of the four images (A hi, A lo, B hi,
// B lo), for two chunks ahead (two register sets: two waves per SIMD, the
// loads of a chunk in flight over two chunks of MFMA work).  LDS rows are 64 bytes (four 16-byte chunks of eight K values),
// chunk c of row r stored at position c ^ sw(r), sw(r) = ((r >> 1) & 3) ^
// ((r >> 3) & 1): a fragment read (16 rows at one chunk) and a staging
// store (4 rows x 4 chunks) then touch distinct banks in each pass, with
// 32 banks (8 lanes a pass) and with 64 (16 lanes a pass).
__global__ __launch_bounds__(256, 2) void k_update_bf16x3p(float *__restrict__ Tt, int N, int NB, int MT, int k,
                                                          int n_begin, int m_off, int nsub,
                                                          const __bf16 *__restrict__ P, int m0)
{
    __shared__ __align__(16) __bf16 Ah[BM * XK];
    __shared__ __align__(16) __bf16 Al[BM * XK];
    __shared__ __align__(16) __bf16 Bh[BM * XK];
    __shared__ __align__(16) __bf16 Bl[BM * XK];
    auto lds = [](int r, int c) { return r * XK + 8 * (c ^ (((r >> 1) & 3) ^ ((r >> 3) & 1))); };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int G = gridDim.x, bid = blockIdx.x;
    const int G8 = (G / 8) * 8;
    const int L = (bid < G8) ? (bid % 8) * (G8 / 8) + bid / 8 : bid;
    int m, n, bm, bn;
    decode_block(L, nsub, n_begin, m_off, MT, &m, &n, &bm, &bn);
    if (m == n && bm < bn) return;
    const int rows_m = tile_rows(N, NB, MT, m);
    const int cols_n = tile_rows(N, NB, MT, n);
    const int r_base = bm * BM, c_base = bn * BM;
    if (r_base >= rows_m || c_base >= cols_n) return;
    const bool full = (r_base + BM <= rows_m) && (c_base + BM <= cols_n);
    const int64_t img = (int64_t)NB * NB;
    const __bf16 *AH = P + (int64_t)(m - m0) * 2 * img + (int64_t)r_base * XK;
    const __bf16 *BH = P + (int64_t)(n - m0) * 2 * img + (int64_t)c_base * XK;
    float *Cp = Tt + tile_index(MT, m, n) * (int64_t)NB * NB;
    const int fr = lane & 15, fk = lane >> 4;

    fx4 acc[4][4];                               // as in k_update_bf16x3
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};
    const bool pre = full && !((m == n) && (bm == bn));
    if (pre) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = *(const fx4 *)&Cp[(int64_t)(c_base + wn * 64 + j * 16 + fr) * NB + r_base + wm * 64 +
                                              i * 16 + 4 * fk];
    }

    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    auto neg = [](bf8 v) -> bf8 { return __builtin_bit_cast(bf8, __builtin_bit_cast(u4, v) ^ 0x80008000u); };
    const int rq = tid >> 2, kq = tid & 3;
    typedef bf8 Stage[2][4];                     // [row half][A hi, A lo, B hi, B lo]
    Stage sa, sb;
    auto gload = [&](Stage &x, int k0) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int row = rq + 64 * i;
            const int64_t o = ((int64_t)(k0 / XK) * NB + row) * XK + 8 * kq;
            const bool ain = full || r_base + row < rows_m;
            const bool bin = full || c_base + row < cols_n;
            const bf8 z = bf8{};
            x[i][0] = ain ? *(const bf8 *)(AH + o) : z;
            x[i][1] = ain ? *(const bf8 *)(AH + img + o) : z;
            x[i][2] = bin ? *(const bf8 *)(BH + o) : z;
            x[i][3] = bin ? *(const bf8 *)(BH + img + o) : z;
        }
    };
    auto lstore = [&](const Stage &x) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int o = lds(rq + 64 * i, kq);
            *(bf8 *)&Ah[o] = neg(x[i][0]);             // -A: the MFMA accumulates C - A B^T
            *(bf8 *)&Al[o] = neg(x[i][1]);
            *(bf8 *)&Bh[o] = x[i][2];
            *(bf8 *)&Bl[o] = x[i][3];
        }
    };
    auto compute = [&]() {
        bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = lds(wm * 64 + i * 16 + fr, fk);
            ah[i] = *(const bf8 *)&Ah[o];
            al[i] = *(const bf8 *)&Al[o];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = lds(wn * 64 + j * 16 + fr, fk);
            bh[j] = *(const bf8 *)&Bh[o];
            bl[j] = *(const bf8 *)&Bl[o];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
    };

    // chunks in pairs: the loads of chunk c + 2 are issued while chunk c
    // is computed (NB / 32 chunks; an odd count ends with a single one)
    const int nstage = NB / XK;
    gload(sa, 0);
    if (nstage > 1) gload(sb, XK);
    for (int st = 0; st < nstage; st += 2) {
        lstore(sa);
        __syncthreads();
        if (st + 2 < nstage) gload(sa, (st + 2) * XK);
        compute();
        __syncthreads();
        if (st + 1 < nstage) {
            lstore(sb);
            __syncthreads();
            if (st + 3 < nstage) gload(sb, (st + 3) * XK);
            compute();
            __syncthreads();
        }
    }

    bf16x3_store(Cp, NB, acc, r_base + wm * 64, c_base + wn * 64, fr, fk, pre, rows_m, cols_n, (m == n) && (bm == bn));
}

// k_update_bf16x3p with 256 x 256 blocks (NB a multiple of 256): half the
// panel operand traffic per flop.  Eight waves as a 2 x 4 grid of 128 x 64;
// the four images' K chunks go straight into LDS by global_load_lds_dwordx4
// (no staging registers), double-buffered in 128 KB (one workgroup per CU).
// A direct load cannot negate, so the B fragments are negated in registers
// instead: (-b) a is the same exact BF16 product as b (-a), so the MFMAs
// sum the same values and the result is that of k_update_bf16x3p bit for
// bit.  Lane l of a wave's load of 16 rows lands at LDS offset 16 l (row
// 16 g + l / 4, slot l % 4), so it fetches K chunk (l % 4) ^ sw(row) of
// that row.  Rows up to r_base + 255 lie inside the NB-row tile image,
// zero past the tile's last row, so no load needs a guard.
typedef __attribute__((address_space(3))) void *lds_ptr;
constexpr int WB = 256;
__global__ __launch_bounds__(512, 1) void k_update_bf16x3w(float *__restrict__ Tt, int N, int NB, int MT, int k,
                                                          int n_begin, int m_off, int nsub,
                                                          const __bf16 *__restrict__ P, int m0)
{
    __shared__ __align__(16) __bf16 S[2][4][WB * XK];     // [buffer][A hi, A lo, B hi, B lo]
    auto sw = [](int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); };
    auto lds = [&](int r, int c) { return r * XK + 8 * (c ^ sw(r)); };
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const int G = gridDim.x, bid = blockIdx.x;
    const int G8 = (G / 8) * 8;
    const int L = (bid < G8) ? (bid % 8) * (G8 / 8) + bid / 8 : bid;
    int m, n, bm, bn;
    decode_block(L, nsub, n_begin, m_off, MT, &m, &n, &bm, &bn);
    if (m == n && bm < bn) return;
    const int rows_m = tile_rows(N, NB, MT, m);
    const int cols_n = tile_rows(N, NB, MT, n);
    const int r_base = bm * WB, c_base = bn * WB;
    if (r_base >= rows_m || c_base >= cols_n) return;
    const int64_t img = (int64_t)NB * NB;
    const __bf16 *AH = P + (int64_t)(m - m0) * 2 * img + (int64_t)r_base * XK;
    const __bf16 *BH = P + (int64_t)(n - m0) * 2 * img + (int64_t)c_base * XK;
    float *Cp = Tt + tile_index(MT, m, n) * (int64_t)NB * NB;
    const int fr = lane & 15, fk = lane >> 4;

    // Each wave's 128 x 64 lies inside one 128 x 128 block (pr, pc) of
    // k_update_bf16x3p, and takes the path that kernel takes there (C as
    // the MFMA's start, or C += the products), so that both round alike.
    const int pr = r_base + wm * 128, pc = c_base + (wn >> 1) * 128;
    const bool pre = pr + 128 <= rows_m && pc + 128 <= cols_n && !(m == n && pr <= pc);
    fx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                acc[i][j][q] = 0.f;
    if (pre) {
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = *(const fx4 *)&Cp[(int64_t)(c_base + wn * 64 + j * 16 + fr) * NB + pr + i * 16 + 4 * fk];
    }
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    auto neg = [](bf8 v) -> bf8 { return __builtin_bit_cast(bf8, __builtin_bit_cast(u4, v) ^ 0x80008000u); };
    // wave w fetches rows 32 w .. 32 w + 31 of the four images
    auto dload = [&](int buf, int k0) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int R = wave * 32 + 16 * h, lr = R + (lane >> 2);
            const int lc = (lane & 3) ^ sw(lr);
            const int64_t o = ((int64_t)(k0 / XK) * NB + lr) * XK + 8 * lc;
            __builtin_amdgcn_global_load_lds(AH + o, (lds_ptr)&S[buf][0][R * XK], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(AH + img + o, (lds_ptr)&S[buf][1][R * XK], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(BH + o, (lds_ptr)&S[buf][2][R * XK], 16, 0, 0);
            __builtin_amdgcn_global_load_lds(BH + img + o, (lds_ptr)&S[buf][3][R * XK], 16, 0, 0);
        }
    };
    const int nstage = NB / XK;
    dload(0, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int st = 0; st < nstage; st++) {
        const int b = st & 1;
        if (st + 1 < nstage) dload(b ^ 1, (st + 1) * XK);   // into the buffer read at step st - 1
        bf8 bh[4], bl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = lds(wn * 64 + j * 16 + fr, fk);
            bh[j] = neg(*(const bf8 *)&S[b][2][o]);          // -B: the MFMA accumulates C - A B^T
            bl[j] = neg(*(const bf8 *)&S[b][3][o]);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int o = lds(wm * 128 + i * 16 + fr, fk);
            const bf8 ah = *(const bf8 *)&S[b][0][o], al = *(const bf8 *)&S[b][1][o];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
            }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");      // this wave's loads of step st + 1
        __syncthreads();                                      // everyone's, and buffer b read
    }

    bf16x3_store(Cp, NB, acc, pr, c_base + wn * 64, fr, fk, pre, rows_m, cols_n, m == n);
}

// ---------------------------------------------------------------------------
// The same trailing update on the FP64 VALU (engine 0; the MFMA kernel
// above is the default: with two resident waves per SIMD both reach about
// the same FP64 rate on gfx950, and the MFMA kernel needs fewer registers
// per flop).  Work decomposition as k_update_mfma (one
// 128x128 block of one tile per workgroup, XCD-grouped); thread (tx, ty) =
// (t % 16, t / 16) keeps the 8x8 block of rows {32g + 2tx + i} and columns
// {32h + 2ty + j} (g, h < 4, i, j < 2) in registers.  Every k step reads
// four 16-byte A pairs (16 distinct per wave, contiguous: no bank conflict)
// and four B pairs (4 distinct per wave: broadcast) for 64 FMAs.

__global__ __launch_bounds__(256) void k_update_valu(double *__restrict__ T, int N, int NB, int MT, int k,
                                                    int n_begin, int m_off, int nsub)
{
    __shared__ double As[2][BK][BM];
    __shared__ double Bs[2][BK][BM];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int G = gridDim.x, bid = blockIdx.x;
    const int G8 = (G / 8) * 8;
    const int L = (bid < G8) ? (bid % 8) * (G8 / 8) + bid / 8 : bid;
    int m, n, bm, bn;
    decode_block(L, nsub, n_begin, m_off, MT, &m, &n, &bm, &bn);
    if (m == n && bm < bn) return;
    const int rows_m = tile_rows(N, NB, MT, m);
    const int cols_n = tile_rows(N, NB, MT, n);
    const int r_base = bm * BM, c_base = bn * BM;
    if (r_base >= rows_m || c_base >= cols_n) return;
    const bool full = (r_base + BM <= rows_m) && (c_base + BM <= cols_n) && (NB % BK == 0) && !(NB & 1);

    const double *Ap = T + tile_index(MT, m, k) * (int64_t)NB * NB;
    const double *Bp = T + tile_index(MT, n, k) * (int64_t)NB * NB;
    double *Cp = T + tile_index(MT, m, n) * (int64_t)NB * NB;

    double acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = 0.0;

    double2 ra[4], rb[4];
    auto gload = [&](int k0) {
        if (full) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = tid + 256 * i;
                const int kk = e >> 6, r = (e & 63) * 2;
                ra[i] = *(const double2 *)(Ap + (int64_t)(k0 + kk) * NB + r_base + r);
                rb[i] = *(const double2 *)(Bp + (int64_t)(k0 + kk) * NB + c_base + r);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = tid + 256 * i;
                const int kk = e >> 6, r = (e & 63) * 2;
                const int kg = k0 + kk;
                double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
                if (kg < NB) {
                    const double *ac = Ap + (int64_t)kg * NB;
                    const double *bc = Bp + (int64_t)kg * NB;
                    if (r_base + r < rows_m) a0 = ac[r_base + r];
                    if (r_base + r + 1 < rows_m) a1 = ac[r_base + r + 1];
                    if (c_base + r < cols_n) b0 = bc[c_base + r];
                    if (c_base + r + 1 < cols_n) b1 = bc[c_base + r + 1];
                }
                ra[i] = make_double2(a0, a1);
                rb[i] = make_double2(b0, b1);
            }
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = tid + 256 * i;
            const int kk = e >> 6, r = (e & 63) * 2;
            *(double2 *)&As[buf][kk][r] = ra[i];
            *(double2 *)&Bs[buf][kk][r] = rb[i];
        }
    };

    const int nstage = (NB + BK - 1) / BK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int st = 0; st < nstage; st++) {
        const int buf = st & 1;
        if (st + 1 < nstage) gload((st + 1) * BK);
#pragma unroll 2
        for (int kk = 0; kk < BK; kk++) {
            double a[8], b[8];
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const double2 va = *(const double2 *)&As[buf][kk][32 * g + 2 * tx];
                const double2 vb = *(const double2 *)&Bs[buf][kk][32 * g + 2 * ty];
                a[2 * g] = va.x;
                a[2 * g + 1] = va.y;
                b[2 * g] = vb.x;
                b[2 * g + 1] = vb.y;
            }
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        if (st + 1 < nstage) lstore(buf ^ 1);
        __syncthreads();
    }

    const bool diag = (m == n) && (bm == bn);
#pragma unroll
    for (int h = 0; h < 4; h++)
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int col = c_base + 32 * h + 2 * ty + jj;
            double *cc = Cp + (int64_t)col * NB;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int row = r_base + 32 * g + 2 * tx;
                if (full && !diag) {
                    double2 c = *(double2 *)(cc + row);
                    c.x -= acc[2 * g][2 * h + jj];
                    c.y -= acc[2 * g + 1][2 * h + jj];
                    *(double2 *)(cc + row) = c;
                } else if (col < cols_n) {
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        const int rr = row + i;
                        if (rr < rows_m && (!diag || rr >= col)) cc[rr] -= acc[2 * g + i][2 * h + jj];
                    }
                }
            }
        }
}

// Recursive diagonal factor (hmb): the lower triangle of one n x n tile
// (column-major, ld) <-> a packed lower image of hmb x hmb sub-tiles (the
// layout of the whole matrix, one level down).  One thread per element of
// the image; the strict upper part of the diagonal sub-tiles and padding
// rows are copied out as zeros and never copied back.
template <class T, bool TO_IMG>
__global__ __launch_bounds__(256) void k_subtile_copy(T *__restrict__ tile, int n, int ld, int hmb, int MTs,
                                                      T *__restrict__ img, int64_t nel)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= nel) return;
    const int64_t per = (int64_t)hmb * hmb;
    int64_t t = e / per;
    const int w = (int)(e - t * per), col = w / hmb, row = w - col * hmb;
    int sn = 0;                                  // sub-tile column
    while (t >= MTs - sn) { t -= MTs - sn; sn++; }
    const int sm = sn + (int)t;
    const int r = sm * hmb + row, cc = sn * hmb + col;
    const bool in = r < n && cc < n && r >= cc;
    if (TO_IMG) img[e] = in ? tile[(int64_t)cc * ld + r] : T(0);
    else if (in) tile[(int64_t)cc * ld + r] = img[e];
}

__global__ void k_fill_int(int *p, int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// panel engine: 0 = VALU diagonal factor / solve (FP64 only), 1 = MFMA panel updates (default)
int g_panel_kernel = 1;
// trailing-update engine: 0 = FP64 VALU, 1 = MFMA (default)
int g_update_kernel = -1;
// waves per workgroup of the 128x128 MFMA update: 4 (2 x 2) or 8 (4 x 2);
// 0 = not yet read from the environment (DPLASMA_MI_UPDATE_WAVES)
int g_update_waves = 0;
// FP32 trailing update of mi_spotrf_tiles and the FP32 step launches: 0 =
// FP32 MFMA (default), 1 = BF16 MFMA, two-term split (the mixed solve passes
// its choice as an argument instead)
int g_supdate = 0;
// BF16-split updates of the main stream: 1 = panel column split once per
// step (k_split_panel + k_update_bf16x3p, default), 0 = split by every
// workgroup (k_update_bf16x3)
int g_bf16_panel_split = 1;

// profiler names per precision
template <class T> struct KN;
// (update: launches on the main stream, which never overlap one another;
// update_la: the look-ahead column updates on the panel stream)
template <> struct KN<double> {
    static constexpr const char *potrf = "dpotrf_potrf_diag";
    static constexpr const char *trsm = "dpotrf_trsm_panel";
    static constexpr const char *update = "dpotrf_update";
    static constexpr const char *update_la = "dpotrf_update_la";
};
template <> struct KN<float> {
    static constexpr const char *potrf = "spotrf_potrf_diag";
    static constexpr const char *trsm = "spotrf_trsm_panel";
    static constexpr const char *update = "spotrf_update";
    static constexpr const char *update_la = "spotrf_update_la";
    static constexpr const char *update_bf16 = "spotrf_update_bf16x3";
    static constexpr const char *update_bf16_la = "spotrf_update_bf16x3_la";
    static constexpr const char *split_panel = "spotrf_split_panel";
};

template <class T> bool valu_panel() { return std::is_same<T, double>::value && g_panel_kernel == 0; }

template <class T>
int potrf_lds_bytes(int NB, int *RC)
{
    if (!valu_panel<T>()) {
        *RC = ((NB + 15) & ~15) + 1;
        return (32 * (*RC) + 1024 + 32 * 33) * (int)sizeof(T);
    }
    *RC = ((NB + 3) & ~3) + 4;
    return (PJB * (*RC) + 1024) * (int)sizeof(double);
}

// info_off: added to k*NB + j of a failing pivot (the offset of this tile
// grid inside a larger matrix)
template <class T>
int launch_potrf(Ctx *c, hipStream_t s, T *dT, int N, int NB, int MT, int k, T *Winv, int info_off = 0)
{
    // LDS and its row stride RC for the n rows of this tile (RC >= n + 1)
    const int n = tile_rows(N, NB, MT, k);
    int RC;
    int lds = potrf_lds_bytes<T>(n, &RC);
    T *A = dT + tile_index(MT, k, k) * (int64_t)NB * NB;
    const double fl = (double)n * n * n / 3.0;
    if (!valu_panel<T>()) {
        MI_LAUNCH(c, s, KN<T>::potrf, fl, 2.0 * n * n * sizeof(T),
                  k_potrf_tile<T><<<1, 512, lds, s>>>(A, n, NB, RC, info_off + k * NB, c->d_info, Winv));
    } else {
        MI_LAUNCH(c, s, KN<T>::potrf, fl, 2.0 * n * n * sizeof(T),
                  k_potrf_diag<<<1, POTRF_THREADS, lds, s>>>((double *)A, n, NB, RC, info_off + k * NB,
                                                              c->d_info, (double *)Winv));
    }
    return 0;
}

// TRSM(m, k) for the ntile tiles m = m0 .. m0 + ntile - 1 (by default every
// m > k).
template <class T>
int launch_trsm(Ctx *c, hipStream_t s, T *dT, int N, int NB, int MT, int k, const T *Winv, int m0 = -1,
                int ntile = -1)
{
    if (m0 < 0) m0 = k + 1;
    if (ntile < 0) ntile = MT - m0;
    if (ntile <= 0) return 0;
    double fl = 0.0;
    for (int m = m0; m < m0 + ntile; m++) fl += (double)tile_rows(N, NB, MT, m) * NB * NB;
    if (!valu_panel<T>()) {
        // 64-row strips while there are many tiles; below ~50 tiles (FP64)
        // or ~40 (FP32) of NB = 512 the 16-row K-split kernel is faster alone
        const int strips = (NB + TS_ROWS - 1) / TS_ROWS;
        if (ntile * strips >= (sizeof(T) == 8 ? 400 : 320)) {
            MI_LAUNCH(c, s, KN<T>::trsm, fl, fl / NB * 2.0 * sizeof(T),
                      k_trsm_tile<T><<<ntile * strips, 256, 0, s>>>(dT, N, NB, MT, k, Winv, m0));
        } else {
            // few tiles: 16-row strips with the K range split over the waves
            const int ks = (NB + TSK_ROWS - 1) / TSK_ROWS;
            MI_LAUNCH(c, s, KN<T>::trsm, fl, fl / NB * 2.0 * sizeof(T),
                      k_trsm_tile_ks<T><<<ntile * ks, 256, 0, s>>>(dT, N, NB, MT, k, Winv, m0));
        }
    } else {
        const int strips = (NB + TRSM_ROWS - 1) / TRSM_ROWS;
        MI_LAUNCH(c, s, KN<T>::trsm, fl, fl / NB * 2.0 * sizeof(T),
                  k_trsm_panel<<<ntile * strips, 256, 0, s>>>((double *)dT, N, NB, MT, k,
                                                             (const double *)Winv, m0));
    }
    return 0;
}

// Trailing update of step k over a run of tiles: column n_begin from row
// n_begin + m_off down, then whole columns, ntiles tiles in all (ntiles < 0:
// through the last column).
template <class T>
int launch_update_run(Ctx *c, hipStream_t s, T *dT, int N, int NB, int MT, int k, int n_begin,
                      int m_off, int64_t ntiles, bool bf16_split, bool la = false,
                      const __bf16 *split = nullptr, int split_m0 = 0)
{
    if (n_begin >= MT || n_begin + m_off >= MT) return 0;
    int64_t avail = MT - n_begin - m_off;
    for (int n = n_begin + 1; n < MT; n++) avail += MT - n;
    if (ntiles < 0 || ntiles > avail) ntiles = avail;
    if (ntiles <= 0) return 0;
    int nsub = (NB + BM - 1) / BM;
    int64_t grid = ntiles * nsub * nsub;
    // a launch of fewer than two 128x128 blocks per CU is latency-bound:
    // 64x64 blocks give four times the workgroups at a quarter of the chain
    if (g_update_kernel < 0) {
        const char *e = getenv("DPLASMA_MI_UPDATE");
        g_update_kernel = (e && !strcmp(e, "valu")) ? 0 : 1;
    }
    if (!g_update_waves) {
        const char *e = getenv("DPLASMA_MI_UPDATE_WAVES");
        g_update_waves = (e && atoi(e) == 4) ? 4 : 8;
    }
    const bool valu = std::is_same<T, double>::value && g_update_kernel == 0;
    const bool bf16 = std::is_same<T, float>::value && bf16_split;
    const bool small = grid < 512 && !valu && !bf16;
    if (small) {
        nsub = (NB + 63) / 64;
        grid = ntiles * nsub * nsub;
    }
    if (grid > 0x7fffffff) { set_error("update grid too large"); return -3; }
    // algorithmic flops / bytes of the covered tiles: each C tile read and
    // written once, each panel tile (m, k) the launch needs read once
    double fl = 0.0, by = 0.0;
    int64_t left = ntiles;
    int m_lo = MT;
    for (int n = n_begin; n < MT && left > 0; n++)
        for (int m = (n == n_begin ? n + m_off : n); m < MT && left > 0; m++, left--) {
            const double tm = tile_rows(N, NB, MT, m);
            fl += (m == n) ? (double)NB * tm * (tm + 1.0) : 2.0 * tm * NB * NB;
            by += 2.0 * tm * NB * sizeof(T);
            m_lo = std::min(m_lo, n);
        }
    for (int m = m_lo; m < MT; m++) by += (double)tile_rows(N, NB, MT, m) * NB * sizeof(T);
    if (valu)
        MI_LAUNCH(c, s, "dpotrf_update_valu", fl, by,
                  k_update_valu<<<(unsigned)grid, 256, 0, s>>>((double *)dT, N, NB, MT, k, n_begin, m_off,
                                                               nsub));
    else if (bf16 && split && NB % WB == 0 && ntiles * (NB / WB) * (NB / WB) >= 2 * 256)
        // 256 x 256 blocks once the launch fills the chip twice over (one
        // workgroup per CU)
        MI_LAUNCH(c, s, la ? KN<float>::update_bf16_la : KN<float>::update_bf16, fl, by,
                  k_update_bf16x3w<<<(unsigned)(ntiles * (NB / WB) * (NB / WB)), 512, 0, s>>>(
                      (float *)dT, N, NB, MT, k, n_begin, m_off, NB / WB, split, split_m0));
    else if (bf16 && split)
        MI_LAUNCH(c, s, la ? KN<float>::update_bf16_la : KN<float>::update_bf16, fl, by,
                  k_update_bf16x3p<<<(unsigned)grid, 256, 0, s>>>((float *)dT, N, NB, MT, k, n_begin, m_off,
                                                                  nsub, split, split_m0));
    else if (bf16)
        MI_LAUNCH(c, s, la ? KN<float>::update_bf16_la : KN<float>::update_bf16, fl, by,
                  k_update_bf16x3<<<(unsigned)grid, 256, 0, s>>>((float *)dT, N, NB, MT, k, n_begin, m_off,
                                                                 nsub));
    else if (small)
        MI_LAUNCH(c, s, la ? KN<T>::update_la : KN<T>::update, fl, by,
                  k_update_mfma<T, 64, 2, 2><<<(unsigned)grid, 256, 0, s>>>(dT, N, NB, MT, k, n_begin, m_off,
                                                                            nsub));
    else if (g_update_waves == 8)
        MI_LAUNCH(c, s, la ? KN<T>::update_la : KN<T>::update, fl, by,
                  k_update_mfma<T, BM, 4, 2><<<(unsigned)grid, 512, 0, s>>>(dT, N, NB, MT, k, n_begin, m_off,
                                                                            nsub));
    else
        MI_LAUNCH(c, s, la ? KN<T>::update_la : KN<T>::update, fl, by,
                  k_update_mfma<T, BM, 2, 2><<<(unsigned)grid, 256, 0, s>>>(dT, N, NB, MT, k, n_begin, m_off,
                                                                            nsub));
    return 0;
}

template <class T>
int launch_update(Ctx *c, hipStream_t s, T *dT, int N, int NB, int MT, int k, int n_begin, int n_end,
                  bool bf16_split, bool la = false, const __bf16 *split = nullptr, int split_m0 = 0)
{
    if (n_begin >= n_end) return 0;
    int64_t ntiles = 0;
    for (int n = n_begin; n < n_end; n++) ntiles += MT - n;
    return launch_update_run<T>(c, s, dT, N, NB, MT, k, n_begin, 0, ntiles, bf16_split, la, split, split_m0);
}

// BF16 split of the panel tiles (m, k), m >= m0, into P (k_split_panel)
int launch_split_panel(Ctx *c, hipStream_t s, const float *dT, int N, int NB, int MT, int k, int m0, __bf16 *P)
{
    if (m0 >= MT) return 0;
    const int nb = (NB + SPT - 1) / SPT;
    const int64_t grid = (int64_t)(MT - m0) * nb * nb;
    double by = 0.0;
    for (int m = m0; m < MT; m++) by += (double)tile_rows(N, NB, MT, m) * NB * sizeof(float) + 2.0 * NB * NB * 2;
    MI_LAUNCH(c, s, KN<float>::split_panel, 0.0, by,
              k_split_panel<<<(unsigned)grid, 256, 0, s>>>(dT, N, NB, MT, k, m0, P));
    return 0;
}

template <class T>
int prepare_potrf(int NB)
{
    int RC;
    int lds = potrf_lds_bytes<T>(NB, &RC);
    int dev = 0;
    MI_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MI_CHECK(hipGetDeviceProperties(&prop, dev));
    size_t limit = prop.sharedMemPerBlock;
    if ((size_t)lds > limit) {
        set_error("tile size NB=%d needs %d B of LDS for the diagonal factorization (limit %zu)",
                  NB, lds, limit);
        return -4;
    }
    if (lds > 65536) {
        MI_CHECK(hipFuncSetAttribute(!valu_panel<T>() ? (const void *)k_potrf_tile<T> : (const void *)k_potrf_diag,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    }
    return 0;
}

// Elements of the diagonal-block inverses handed from POTRF(k) to TRSM(k),
// two steps in flight.
constexpr size_t potrf_ws_elems(int NB) { return (size_t)2 * ((NB + 31) / 32) * 1024; }


// Whole factorization of the packed lower tiles dT (N, NB), resident, with a
// two-step look-ahead over two streams (src/zpotrf_L.jdf's priorities give
// the panel the critical path):
//   panel stream, step k:  [wait: column k+1 updated by step k-1]
//                          column k+1 updated by step k (SYRK + GEMMs),
//                          POTRF(k+1), TRSM(k+1)
//   main stream, step k:   [wait: TRSM(k)]
//                          column k+2 updated by step k  -> event for the panel,
//                          columns k+3.. updated by step k (the bulk)
// Every tile gets its updates in step order and no two launches touch the
// same tile at once; the panel chain waits for one column of the bulk of
// step k-1, not for all of it.
// bf16_split: FP32 trailing updates on BF16 MFMA with the two-term split.
// ws: potrf_ws_elems(NB) elements of workspace, or nullptr for the context's.
// ring: cross-stream dependencies use the context's fixed ring of events
// (created with the context) instead of its growing pool; every event is
// waited on within three records of being recorded, far inside the ring.
// POTRF(k) of an n x n diagonal tile (ld NB) as a tile Cholesky of its own
// with hmb x hmb sub-tiles (src/zpotrf_L.jdf:168-192: the RECURSIVE body
// re-tiles the task with smallnb when the tile is larger), enqueued on s:
// the tile is copied into a packed image of sub-tiles, factored there by
// the same POTRF / TRSM / update launches (many workgroups for the panel
// solves and updates inside the tile), and the factor copied back.  hmb is
// a multiple of 32, so the 32 x 32 diagonal-block inverses of sub-step j
// land in Winv at block j * hmb / 32, where the panel TRSM of the enclosing
// step reads them.  img: the packed image, (MTs (MTs + 1) / 2) hmb^2
// elements.
template <class T>
int potrf_rec(Ctx *c, hipStream_t s, T *tile, int n, int NB, int hmb, T *Winv, int info_off, T *img)
{
    const int MTs = (n + hmb - 1) / hmb;
    const int64_t nel = (int64_t)MTs * (MTs + 1) / 2 * hmb * hmb;
    const unsigned grid = (unsigned)((nel + 255) / 256);
    MI_LAUNCH(c, s, "potrf_rec_copy", 0.0, 16.0 * nel,
              (k_subtile_copy<T, true><<<grid, 256, 0, s>>>(tile, n, NB, hmb, MTs, img, nel)));
    int rc;
    for (int j = 0; j < MTs; j++) {
        T *w = Winv + (int64_t)(j * hmb / 32) * 1024;
        if ((rc = launch_potrf<T>(c, s, img, n, hmb, MTs, j, w, info_off))) return rc;
        if ((rc = launch_trsm<T>(c, s, img, n, hmb, MTs, j, w, -1, -1))) return rc;
        if ((rc = launch_update<T>(c, s, img, n, hmb, MTs, j, j + 1, MTs, false, true))) return rc;
    }
    MI_LAUNCH(c, s, "potrf_rec_copy", 0.0, 16.0 * nel,
              (k_subtile_copy<T, false><<<grid, 256, 0, s>>>(tile, n, NB, hmb, MTs, img, nel)));
    return 0;
}

// hmb: recursive diagonal factor with sub-tiles of hmb (rounded down to a
// multiple of 32, at least 32) for the diagonal tiles larger than that; 0:
// one POTRF launch per step.  img_ext: the sub-tile image (nullptr: the
// context's workspace).  rec_from: the first step whose diagonal tile is
// factored recursively (the earlier ones take one POTRF launch).
// Sub-tile size of the recursive diagonal factor used by default: two
// sub-tiles of NB / 2 a side for NB >= 512 (a multiple of 64), none below
// (the image then has 3 (NB / 2)^2 elements).  The FP64 factor uses it for
// every step (N = 32768, NB = 512: 202.2 against 204.2 ms); the BF16-split
// FP32 factor only from step MT / 2 on, where the panel chain is critical
// (59.6 against 61.6-61.9 ms; from step 0 60.7 ms: in the first half the
// extra panel launches compete with the bulk update).
constexpr int half_tile_hmb(int NB) { return (NB >= 512 && NB % 64 == 0) ? NB / 2 : 0; }

template <class T>
int potrf_tiles(Ctx *c, int N, int NB, T *dT, int *info, bool bf16_split, T *ws, bool ring = false, int hmb = 0,
                T *img_ext = nullptr, int rec_from = 0, bool wait = true)
{
    *info = 0;
    if (N < 0 || NB <= 0) { set_error("bad sizes N=%d NB=%d", N, NB); return -2; }
    if (N == 0) return 0;
    const int MT = (N + NB - 1) / NB;
    int rc;
    const size_t wsz = potrf_ws_elems(NB) / 2;      // elements per step
    if (!ws && !(ws = (T *)ws_get(c, WS_POTRF_WINV, potrf_ws_elems(NB) * sizeof(T)))) return -3;
    c->ev_next = 0;
    size_t ring_next = 0;
    auto ev = [&]() -> hipEvent_t { return ring ? c->ev_ring[ring_next++ % c->ev_ring.size()] : next_event(c); };
    if (hmb > 0) hmb = std::max(32, hmb & ~31);
    T *img = nullptr;
    if (hmb > 0 && hmb < NB) {
        const int MTs = (NB + hmb - 1) / hmb;
        const size_t ib = (size_t)MTs * (MTs + 1) / 2 * hmb * hmb * sizeof(T);
        if (!(img = img_ext ? img_ext : (T *)ws_get(c, WS_POTRF_REC, ib))) return -3;
    }
    // the largest diagonal block one POTRF launch factors: a sub-tile when
    // the diagonal tiles are factored recursively (so NB may exceed what one
    // launch's LDS holds), else a whole tile
    if ((rc = prepare_potrf<T>(img ? hmb : NB))) return rc;
    // FP32 factor with BF16 updates: the main stream splits each panel
    // column once (tiles k+2.. of column k; its launches run in order, so one
    // image serves every step)
    __bf16 *split = nullptr;
    if (std::is_same<T, float>::value && bf16_split && g_bf16_panel_split && NB % XK == 0 && MT > 2) {
        split = (__bf16 *)ws_get(c, WS_BF16_PANEL, (size_t)(MT - 2) * 2 * NB * NB * sizeof(__bf16));
        if (!split) return -3;
    }
    auto diag = [&](int k) -> int {
        T *w = ws + (k & 1) * wsz;
        const int n = tile_rows(N, NB, MT, k);
        if (img && n > hmb && k >= rec_from)
            return potrf_rec<T>(c, c->s_panel, dT + tile_index(MT, k, k) * (int64_t)NB * NB, n, NB, hmb, w, k * NB,
                                img);
        return launch_potrf<T>(c, c->s_panel, dT, N, NB, MT, k, w);
    };

    MI_CHECK(hipEventRecord(c->t0, c->s_main));
    k_fill_int<<<1, 64, 0, c->s_main>>>(c->d_info, 0, 1);
    hipEvent_t start = ev();
    MI_CHECK(hipEventRecord(start, c->s_main));
    MI_CHECK(hipStreamWaitEvent(c->s_panel, start, 0));

    // panel 0
    if ((rc = diag(0))) return rc;
    if ((rc = launch_trsm<T>(c, c->s_panel, dT, N, NB, MT, 0, ws))) return rc;
    hipEvent_t pk = ev();                          // TRSM(k) done
    MI_CHECK(hipEventRecord(pk, c->s_panel));
    hipEvent_t col_ready = nullptr;                // column k+1 updated by step k-1

    for (int k = 0; k < MT - 1; k++) {
        // panel stream: column k+1 by step k, then the next panel
        if (col_ready) MI_CHECK(hipStreamWaitEvent(c->s_panel, col_ready, 0));
        if ((rc = launch_update<T>(c, c->s_panel, dT, N, NB, MT, k, k + 1, k + 2, bf16_split, true))) return rc;
        T *w = ws + ((k + 1) & 1) * wsz;
        if ((rc = diag(k + 1))) return rc;
        if ((rc = launch_trsm<T>(c, c->s_panel, dT, N, NB, MT, k + 1, w))) return rc;
        hipEvent_t pk_next = ev();
        MI_CHECK(hipEventRecord(pk_next, c->s_panel));
        // main stream: the look-ahead column k+2, then the bulk of step k
        MI_CHECK(hipStreamWaitEvent(c->s_main, pk, 0));
        col_ready = nullptr;
        const __bf16 *sp = (split && k + 2 < MT) ? split : nullptr;
        if (sp && (rc = launch_split_panel(c, c->s_main, (const float *)dT, N, NB, MT, k, k + 2, split))) return rc;
        if (k + 2 < MT) {
            if ((rc = launch_update<T>(c, c->s_main, dT, N, NB, MT, k, k + 2, k + 3, bf16_split, false, sp, k + 2)))
                return rc;
            col_ready = ev();
            MI_CHECK(hipEventRecord(col_ready, c->s_main));
        }
        if ((rc = launch_update<T>(c, c->s_main, dT, N, NB, MT, k, k + 3, MT, bf16_split, false, sp, k + 2)))
            return rc;
        pk = pk_next;
    }
    MI_CHECK(hipStreamWaitEvent(c->s_main, pk, 0));
    MI_CHECK(hipEventRecord(c->t1, c->s_main));
    if (!wait) return 0;                           // info stays in d_info[0]
    int h_info = 0;
    MI_CHECK(hipMemcpyAsync(&h_info, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->s_main));
    if ((rc = mi_synchronize(static_cast<mi_ctx_t *>(c)))) return rc;
    float ms = 0.f;
    MI_CHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_ms = ms;
    *info = h_info;
    return 0;
}

// single task-class launches, blocking (tests)
template <class T>
int step_potrf(mi_ctx_t *c, int N, int NB, T *dT, int k, int *info)
{
    if (N <= 0) { *info = 0; return 0; }
    const int MT = (N + NB - 1) / NB;
    int rc = prepare_potrf<T>(NB);
    if (rc) return rc;
    T *ws = (T *)ws_get(c, WS_POTRF_WINV, potrf_ws_elems(NB) * sizeof(T));
    if (!ws) return -3;
    k_fill_int<<<1, 64, 0, c->s_main>>>(c->d_info, 0, 1);
    if ((rc = launch_potrf<T>(c, c->s_main, dT, N, NB, MT, k, ws))) return rc;
    MI_CHECK(hipMemcpyAsync(info, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->s_main));
    return mi_synchronize(c);
}

template <class T>
int step_trsm(mi_ctx_t *c, int N, int NB, T *dT, int k)
{
    // the diagonal-block inverses come from the POTRF launch of the same k
    const int MT = (N + NB - 1) / NB;
    const T *ws = (const T *)c->ws[WS_POTRF_WINV].p;
    if (!ws || c->ws[WS_POTRF_WINV].bytes < potrf_ws_elems(NB) * sizeof(T)) {
        set_error("the panel solve of step k needs the POTRF launch of step k first");
        return -2;
    }
    int rc = launch_trsm<T>(c, c->s_main, dT, N, NB, MT, k, ws);
    if (rc) return rc;
    return mi_synchronize(c);
}

template <class T>
int step_update(mi_ctx_t *c, int N, int NB, T *dT, int k, int n_begin, int n_end)
{
    const int MT = (N + NB - 1) / NB;
    if (k < 0 || n_begin <= k || n_end > MT) { set_error("bad update range"); return -2; }
    int rc = launch_update<T>(c, c->s_main, dT, N, NB, MT, k, n_begin, n_end, g_supdate == 1);
    if (rc) return rc;
    return mi_synchronize(c);
}

}  // namespace

namespace mi {

// Single tile tasks on the packed lower image, enqueued on s (the bodies of
// the four task classes of src/zpotrf_L.jdf:91-633, one tile each).
int task_prepare_fp64(int NB) { return prepare_potrf<double>(NB); }
int task_potrf_fp64(Ctx *c, hipStream_t s, double *dT, int N, int NB, int k, double *Winv)
{
    return launch_potrf<double>(c, s, dT, N, NB, (N + NB - 1) / NB, k, Winv);
}
int task_trsm_fp64(Ctx *c, hipStream_t s, double *dT, int N, int NB, int m, int k, const double *Winv)
{
    return launch_trsm<double>(c, s, dT, N, NB, (N + NB - 1) / NB, k, Winv, m, 1);
}
// tile (m, n), m >= n > k: A(m, n) -= A(m, k) A(n, k)^T (SYRK when m == n)
int task_update_fp64(Ctx *c, hipStream_t s, double *dT, int N, int NB, int m, int n, int k)
{
    return launch_update_run<double>(c, s, dT, N, NB, (N + NB - 1) / NB, k, n, m - n, 1, false);
}
size_t task_winv_elems(int NB) { return potrf_ws_elems(NB) / 2; }

int spotrf_tiles_split(Ctx *c, int N, int NB, float *dT, int *info, bool bf16_split, bool wait)
{
    const int MT = NB > 0 ? (N + NB - 1) / NB : 0;
    return potrf_tiles<float>(c, N, NB, dT, info, bf16_split, nullptr, false, bf16_split ? half_tile_hmb(NB) : 0,
                              nullptr, MT / 2, wait);
}
}  // namespace mi

// ---------------------------------------------------------------------------
// Allocation-free FP64 factorization (mi_dpotrf_tiles_noalloc).  Its only
// scratch memory, the 32 x 32 diagonal-block inverses that POTRF(k) hands to
// TRSM(k) for two steps in flight and the sub-tile image of the recursive
// diagonal factor (the same factor as mi_dpotrf_tiles), lives in this
// file-scope device array instead of context workspaces: one copy per GPU,
// held until the process ends, shared by every call on that GPU.  POTRF(k)
// writes each 1024-element block whole before TRSM(k) reads it, and the
// image is copied in from the tile before it is read, so nothing has to be
// cleared between calls.  On the host the array's name is only a stand-in: host code uses
// the device address that noalloc_setup reads once per device.
namespace {
constexpr int NOALLOC_MAX_NB = 1024;
constexpr int NOALLOC_MAX_DEV = 64;
constexpr size_t NOALLOC_WINV = potrf_ws_elems(NOALLOC_MAX_NB);                          // 512 KiB
constexpr size_t NOALLOC_IMG = 3 * (size_t)(NOALLOC_MAX_NB / 2) * (NOALLOC_MAX_NB / 2);   // 6 MiB
__device__ __align__(256) double scratch[NOALLOC_WINV + NOALLOC_IMG];

std::mutex g_noalloc_mu;
bool g_noalloc_done[NOALLOC_MAX_DEV];
hipError_t g_noalloc_err[NOALLOC_MAX_DEV];
double *g_noalloc_ptr[NOALLOC_MAX_DEV];
}  // namespace

namespace mi {
// Called by mi_ctx_create with `device` current: the device address of
// `scratch` on that GPU, read once per device and kept.
void noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEV) return;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc_done[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();
    g_noalloc_err[device] = e;
    g_noalloc_ptr[device] = e == hipSuccess ? static_cast<double *>(p) : nullptr;
    g_noalloc_done[device] = true;
}
}  // namespace mi

extern "C" {

int mi_set_panel_engine(int which)
{
    if (which != 0 && which != 1) { set_error("panel engine must be 0 (VALU) or 1 (MFMA)"); return -2; }
    g_panel_kernel = which;
    return 0;
}

int mi_set_spotrf_update(int which)
{
    if (which != 0 && which != 1) { set_error("FP32 update must be 0 (FP32 MFMA) or 1 (BF16 split)"); return -2; }
    g_supdate = which;
    return 0;
}

int mi_set_bf16_panel_split(int which)
{
    if (which != 0 && which != 1) { set_error("BF16 panel split must be 0 (per workgroup) or 1 (once per step)"); return -2; }
    g_bf16_panel_split = which;
    return 0;
}

int mi_set_update_engine(int which)
{
    if (which != 0 && which != 1) { set_error("update engine must be 0 (VALU) or 1 (MFMA)"); return -2; }
    g_update_kernel = which;
    return 0;
}

int mi_dpotrf_step_potrf(mi_ctx_t *c, int N, int NB, double *dT, int k, int *info)
{
    return step_potrf<double>(c, N, NB, dT, k, info);
}

int mi_dpotrf_step_trsm(mi_ctx_t *c, int N, int NB, double *dT, int k)
{
    return step_trsm<double>(c, N, NB, dT, k);
}

int mi_dpotrf_step_update(mi_ctx_t *c, int N, int NB, double *dT, int k, int n_begin, int n_end)
{
    return step_update<double>(c, N, NB, dT, k, n_begin, n_end);
}

int mi_spotrf_step_potrf(mi_ctx_t *c, int N, int NB, float *dT, int k, int *info)
{
    return step_potrf<float>(c, N, NB, dT, k, info);
}

int mi_spotrf_step_trsm(mi_ctx_t *c, int N, int NB, float *dT, int k)
{
    return step_trsm<float>(c, N, NB, dT, k);
}

int mi_spotrf_step_update(mi_ctx_t *c, int N, int NB, float *dT, int k, int n_begin, int n_end)
{
    return step_update<float>(c, N, NB, dT, k, n_begin, n_end);
}

int mi_dpotrf_tiles(mi_ctx_t *c, int N, int NB, double *dT, int *info)
{
    return potrf_tiles<double>(c, N, NB, dT, info, false, nullptr, false, half_tile_hmb(NB));
}

int mi_dpotrf_tiles_rec(mi_ctx_t *c, int N, int NB, double *dT, int hmb, int *info)
{
    if (hmb <= 0) { *info = 0; set_error("mi_dpotrf_tiles_rec: hmb=%d must be positive", hmb); return -2; }
    return potrf_tiles<double>(c, N, NB, dT, info, false, nullptr, false, hmb);
}

int mi_dpotrf_tiles_noalloc(mi_ctx_t *c, int N, int NB, double *dT, int *info)
{
    *info = 0;
    if (N < 0 || NB <= 0) { set_error("mi_dpotrf_tiles_noalloc: bad sizes N=%d NB=%d", N, NB); return -2; }
    if (NB > NOALLOC_MAX_NB) {
        set_error("mi_dpotrf_tiles_noalloc: NB=%d exceeds the limit of its fixed scratch (NB <= %d)", NB,
                  NOALLOC_MAX_NB);
        return -2;
    }
    double *ptr = nullptr;
    hipError_t e = hipErrorInvalidDevice;
    if (c->device >= 0 && c->device < NOALLOC_MAX_DEV) {
        std::lock_guard<std::mutex> g(g_noalloc_mu);
        if (g_noalloc_done[c->device]) {
            ptr = g_noalloc_ptr[c->device];
            e = g_noalloc_err[c->device];
        }
    }
    if (!ptr) {
        set_error("mi_dpotrf_tiles_noalloc: no scratch address for device %d (%s)", c->device,
                  hipGetErrorString(e));
        return -1;
    }
    // the context's fixed event ring instead of its event pool, which grows
    // on demand; the profiler creates events on demand, so it is off for
    // this call
    const bool prof = c->prof;
    c->prof = false;
    const int hmb = half_tile_hmb(NB);            // image: 3 (NB / 2)^2 <= NOALLOC_IMG elements
    const int rc = potrf_tiles<double>(c, N, NB, dT, info, false, ptr, true, hmb, ptr + NOALLOC_WINV);
    c->prof = prof;
    return rc;
}

int mi_spotrf_tiles(mi_ctx_t *c, int N, int NB, float *dT, int *info)
{
    return spotrf_tiles_split(c, N, NB, dT, info, g_supdate == 1);
}

}  // extern "C"
