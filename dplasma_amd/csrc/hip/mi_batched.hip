// mi_batched.hip -- many independent SPD systems of uneven sizes, factored
// and solved in one persistent launch (captured in a hipGraph for replay).
//
// No counterpart in the reference (its batching growth path: each matrix
// is dplasma_dposv, src/zposv_wrapper.c:63-113, on one tile; the per-matrix
// semantics are LAPACK dpotrf/dpotrs as in src/cores/core_zpotrf.c:68-74:
// info = first column whose pivot is not positive, 1-based, and the matrix
// is left partly factored).
//
// Layout: matrix i is n_i x n_i column-major (lda = n_i) at element offset
// sum_{j<i} n_j^2 of one buffer; its right-hand side (NRHS = 1) at offset
// sum_{j<i} n_j of another.  n_i <= 1024.
//
// Kernel: gridDim = (workgroups per CU) x CUs workgroups of 4 waves; each
// takes matrices from a queue (one atomic counter) in decreasing size order
// (longest first, which bounds the imbalance by the last small matrices) and
// factors one matrix with all its waves, left-looking over 32-column panels:
//   1. panel update C -= L(j:n, 0:j) L(j:j+32, 0:j)^T on MFMA, strips of SH
//      rows staged through LDS in K chunks of 32;
//   2. the 32x32 diagonal block factored and inverted (W = L_jj^-1) by wave
//      0 in one register sweep (one v_readlane broadcast per column);
//   3. the rows below: X = C W^T on MFMA;
// then the two triangular solves, 32 rows at a time (partial sums by all
// waves, the 32-row substitution by wave 0).
// Two strip heights are compiled (SH = 64 and 128): matrices with
// n <= nsplit take SH = 64.  nsplit and the workgroups per CU are the tile
// shape the autotuner picks (mi_batch_autotune).
#include <algorithm>
#include <numeric>
#include <vector>

#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

constexpr int BW = 32;            // diagonal block / panel-solve width
// width of the panel the long update covers: 64 in FP64 (two workgroups of
// 76 KB LDS per CU), 128 in FP32 (42 -> 50 KB; in FP64 the 92 KB it needs
// would leave one workgroup per CU: 36.3 against 28.8 ms for the batch)
template <class T> constexpr int panel_width() { return sizeof(T) == 8 ? 64 : 128; }
constexpr int BMAX = 1024;        // largest matrix order of a batch

// A[i] where c holds, +0 elsewhere; ic is i clamped into the matrix.  In
// FP64 the load reads ic, a valid address, whatever c is: the compiler then
// issues the loads of a whole staging group together instead of branching
// around each one and waiting for it on its own (batch 28.7 -> 21.3 ms).  In
// FP32 that raises the register spills (0 -> 186) and costs more than it
// gains (19.1 -> 26.1 ms), so there the load stays conditional on c.
template <class T> __device__ __forceinline__ T mload(const T *A, int ic, int i, bool c)
{
    if constexpr (sizeof(T) == 8) {
        const T v = A[ic];                        // read whatever c is
        return c ? v : T(0);
    } else {
        return c ? A[i] : T(0);
    }
}

template <class T, int SH, int BWC>
struct BLds {
    static constexpr int LPA = SH + 128 / (int)sizeof(T);
    static constexpr int LPB = BWC + 128 / (int)sizeof(T);
};

template <class T>
struct BShared {
    T sa[BW * (128 + 128 / sizeof(T))];   // -L rows of the strip, [kk][row]
    T sb[BW * (panel_width<T>() + 128 / sizeof(T))];   // L rows of the panel's columns, [kk][col]
    T ws[BW][BW + 1];                     // its inverse
    T ys[BMAX];                           // solve vector
    T red[8][BW];
    int idx;
    int fail;
};

// 1. C(0:M, 0:jb) -= L(j0:n, k0:k1) L(j0:j0+jb, k0:k1)^T, C = A(j0:n, j0:j0+jb),
// jb <= BWC (k0, k1 multiples of 32): the long update of a panel of BWC
// columns over every column left of it (k0 = 0, k1 = j0), or the update of
// the second half of a 64-wide panel by its first half
template <class T, int SH, int BWC>
__device__ void panel_update(T *__restrict__ A, int n, int j0, int jb, int k0, int k1, BShared<T> &sh)
{
    typedef typename MF<T>::v4 V4;
    constexpr int LPA = BLds<T, SH, BWC>::LPA, LPB = BLds<T, SH, BWC>::LPB;
    constexpr int TC = BWC / 16;                // 16-column MFMA tiles
    constexpr int NB_ = BWC * BW / 256;         // staged B elements per thread
    constexpr int TW = SH / 64;                 // 16-row MFMA tiles per wave
    constexpr int NA = SH * BW / 256;           // staged A elements per thread
    // opaque per call: offsets derived from tid are recomputed here instead
    // of being held live across the caller's loops
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const int M = n - j0;
    const int lda = n;   // n <= 1024: 32-bit offsets
    for (int s = 0; s < M; s += SH) {
        V4 acc[TW][TC];
#pragma unroll
        for (int i = 0; i < TW; i++)
#pragma unroll
            for (int jj = 0; jj < TC; jj++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = s + wave * (SH / 4) + i * 16 + fr;
                    const int col = jj * 16 + MF<T>::row(q, fk);
                    acc[i][jj][q] = mload(A, (j0 + min(col, jb - 1)) * lda + j0 + min(row, M - 1),
                                          (j0 + col) * lda + j0 + row, row < M && col < jb && row >= col);
                }
        T ra[NA], rb[NB_];
        auto gload = [&](int kc) {
#pragma unroll
            for (int p = 0; p < NA; p++) {
                const int e = tid + 256 * p, kk = e / SH, r = e - kk * SH;
                ra[p] = -mload(A, (kc + kk) * lda + j0 + min(s + r, M - 1), (kc + kk) * lda + j0 + s + r, s + r < M);
            }
#pragma unroll
            for (int p = 0; p < NB_; p++) {
                const int e = tid + 256 * p, kk = e / BWC, c = e - kk * BWC;
                rb[p] = mload(A, (kc + kk) * lda + j0 + min(c, jb - 1), (kc + kk) * lda + j0 + c, c < jb);
            }
        };
        auto lstore = [&]() {
#pragma unroll
            for (int p = 0; p < NA; p++) {
                const int e = tid + 256 * p, kk = e / SH, r = e - kk * SH;
                sh.sa[kk * LPA + r] = ra[p];
            }
#pragma unroll
            for (int p = 0; p < NB_; p++) {
                const int e = tid + 256 * p, kk = e / BWC, c = e - kk * BWC;
                sh.sb[kk * LPB + c] = rb[p];
            }
        };
        if (k0 < k1) gload(k0);
        for (int kc = k0; kc < k1; kc += BW) {
            __syncthreads();                    // previous chunk consumed
            lstore();
            __syncthreads();
            if (kc + BW < k1) gload(kc + BW);
#pragma unroll
            for (int kq = 0; kq < BW / 4; kq++) {
                const int kk = kq * 4 + fk;
                T a[TW], b[TC];
#pragma unroll
                for (int i = 0; i < TW; i++) a[i] = sh.sa[kk * LPA + wave * (SH / 4) + i * 16 + fr];
#pragma unroll
                for (int jj = 0; jj < TC; jj++) b[jj] = sh.sb[kk * LPB + jj * 16 + fr];
#pragma unroll
                for (int i = 0; i < TW; i++)
#pragma unroll
                    for (int jj = 0; jj < TC; jj++) acc[i][jj] = MF<T>::mfma(b[jj], a[i], acc[i][jj]);
            }
        }
#pragma unroll
        for (int i = 0; i < TW; i++)
#pragma unroll
            for (int jj = 0; jj < TC; jj++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = s + wave * (SH / 4) + i * 16 + fr;
                    const int col = jj * 16 + MF<T>::row(q, fk);
                    if (row < M && col < jb && row >= col) A[(j0 + col) * lda + j0 + row] = acc[i][jj][q];
                }
    }
}

// 2. factor the diagonal block and invert it in one register sweep of wave
// 0: lanes 0-31 hold rows of the block and become L_jj, lanes 32-63 hold
// the rows of the identity and become the columns of W = L_jj^-1 (the same
// eliminations applied to [L | I]).  Returns the 1-based column of the first
// non-positive pivot within the block, or 0.
template <class T>
__device__ __forceinline__ int diag_block(T *__restrict__ A, int n, int j0, int jb, BShared<T> &sh)
{
    const int tid = threadIdx.x;
    const int lda = n;   // n <= 1024: 32-bit offsets
    if (tid < 64) {
        int fail = 0;
        int r = tid;
        // opaque to the optimizer: keeps the 32 identity tests below from
        // being hoisted out of the caller's panel loop into live registers
        asm volatile("" : "+v"(r));
        const bool lrow = r < BW;
        const int jw = r - BW;
        T v[BW];
        const int rc = min(r, jb - 1);
#pragma unroll
        for (int c = 0; c < BW; c++) {
            const bool ld = lrow && r < jb && c < jb && c <= r;
            const T a = mload(A, (j0 + min(c, jb - 1)) * lda + j0 + rc, (j0 + c) * lda + j0 + r, ld);
            const T id = lrow ? ((r == c) ? T(1) : T(0)) : ((c == jw) ? T(1) : T(0));
            v[c] = ld ? a : id;
        }
#pragma unroll
        for (int c = 0; c < BW; c++) {
            const T piv = readlane_t(v[c], c);
            if (!(piv > T(0)) && fail == 0) fail = c + 1;
            const T d = sqrt(piv > T(0) ? piv : T(1));
            const T rd = T(1) / d;
            const T sc = v[c] * rd;
            const T m = (!lrow || r > c) ? sc : T(0);
            v[c] = lrow ? (r == c ? d : (r > c ? sc : v[c])) : sc;
#pragma unroll
            for (int c2 = c + 1; c2 < BW; c2++) {
                v[c2] -= m * readlane_t(m, c2);
                // at most four broadcasts in flight: keeps the scalar
                // registers they land in from piling up (fewer spills)
                if ((c2 & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (r == 0) sh.fail = (fail > 0 && fail <= jb) ? fail : 0;   // wave-uniform
        if (lrow) {
            if (r < jb) {
#pragma unroll
                for (int c = 0; c < BW; c++)
                    if (c <= r && c < jb) A[(j0 + c) * lda + j0 + r] = v[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < BW; c++) sh.ws[c][jw] = v[c];      // W(c, jw)
        }
    }
    __syncthreads();
    const int f = sh.fail;
    __syncthreads();
    return f;
}

// 3. rows below the diagonal block: X = C W^T, in place
template <class T, int SH>
__device__ void panel_solve(T *__restrict__ A, int n, int j0, int jb, BShared<T> &sh)
{
    typedef typename MF<T>::v4 V4;
    constexpr int TW = SH / 64;
    // opaque per call: offsets derived from tid are recomputed here instead
    // of being held live across the caller's loops
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int fr = lane & 15, fk = lane >> 4;
    const int lda = n;   // n <= 1024: 32-bit offsets
    for (int s = j0 + jb; s < n; s += SH) {
        T a[TW][BW / 4];
#pragma unroll
        for (int i = 0; i < TW; i++)
#pragma unroll
            for (int kq = 0; kq < BW / 4; kq++) {
                const int row = s + wave * (SH / 4) + i * 16 + fr, k = kq * 4 + fk;
                a[i][kq] = mload(A, (j0 + min(k, jb - 1)) * lda + min(row, n - 1), (j0 + k) * lda + row, row < n && k < jb);
            }
        V4 acc[TW][2];
#pragma unroll
        for (int i = 0; i < TW; i++)
#pragma unroll
            for (int jj = 0; jj < 2; jj++) acc[i][jj] = V4{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int kq = 0; kq < BW / 4; kq++) {
            T b[2];
#pragma unroll
            for (int jj = 0; jj < 2; jj++) b[jj] = sh.ws[jj * 16 + fr][kq * 4 + fk];   // W^T(k, c) = W(c, k)
#pragma unroll
            for (int i = 0; i < TW; i++)
#pragma unroll
                for (int jj = 0; jj < 2; jj++) acc[i][jj] = MF<T>::mfma(b[jj], a[i][kq], acc[i][jj]);
        }
#pragma unroll
        for (int i = 0; i < TW; i++)
#pragma unroll
            for (int jj = 0; jj < 2; jj++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = s + wave * (SH / 4) + i * 16 + fr;
                    const int col = jj * 16 + MF<T>::row(q, fk);
                    if (row < n && col < jb) A[(j0 + col) * lda + row] = acc[i][jj][q];
                }
    }
}

// Left-looking over panels of PW = 64 (FP64) or 128 (FP32) columns: one long
// update of the whole panel by every column left of it (the strip of L to
// the left is streamed once per PW columns instead of once per 32), then
// the panel's blocks of BW = 32: the update by the panel's earlier blocks,
// diagonal block, and the rows below it.
template <class T>
__device__ int factor_one(T *__restrict__ A, int n, bool narrow, BShared<T> &sh)
{
    constexpr int PW = panel_width<T>();
    for (int p0 = 0; p0 < n; p0 += PW) {
        const int pb = min(PW, n - p0);
        if (narrow) panel_update<T, 64, PW>(A, n, p0, pb, 0, p0, sh);
        else panel_update<T, 128, PW>(A, n, p0, pb, 0, p0, sh);
        __syncthreads();
        for (int j0 = p0; j0 < p0 + pb; j0 += BW) {
            const int jb = min(BW, n - j0);
            if (j0 > p0) {
                if (narrow) panel_update<T, 64, BW>(A, n, j0, jb, p0, j0, sh);
                else panel_update<T, 128, BW>(A, n, j0, jb, p0, j0, sh);
                __syncthreads();
            }
            const int f = diag_block<T>(A, n, j0, jb, sh);
            if (f) return j0 + f;
            if (narrow) panel_solve<T, 64>(A, n, j0, jb, sh);
            else panel_solve<T, 128>(A, n, j0, jb, sh);
            __syncthreads();
        }
    }
    return 0;
}

// L L^T x = b in place (NRHS = 1)
template <class T>
__device__ void solve_one(const T *__restrict__ A, int n, T *__restrict__ b, BShared<T> &sh)
{
    const int tid = threadIdx.x;
    const int lda = n;   // n <= 1024: 32-bit offsets
    const int nblk = (n + BW - 1) / BW;
    for (int i = tid; i < n; i += 256) sh.ys[i] = b[i];
    __syncthreads();
    // forward: y = L^-1 b
    for (int bj = 0; bj < nblk; bj++) {
        const int j0 = bj * BW, jb = min(BW, n - j0);
        {
            const int r = tid & 31, sl = tid >> 5;
            T acc = T(0);
            if (r < jb) {
                // eight independent loads in flight per step
                int k = sl;
                for (; k + 56 < j0; k += 64) {
                    T a[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) a[u] = A[(k + 8 * u) * lda + j0 + r];
#pragma unroll
                    for (int u = 0; u < 8; u++) acc += a[u] * sh.ys[k + 8 * u];
                }
                for (; k < j0; k += 8) acc += A[k * lda + j0 + r] * sh.ys[k];
            }
            sh.red[sl][r] = acc;
        }
        __syncthreads();
        if (tid < 64) {
            const int r = tid;
            T t = T(0), dg = T(1);
            if (r < jb) {
                t = sh.ys[j0 + r];
                for (int sl = 0; sl < 8; sl++) t -= sh.red[sl][r];
                dg = A[(j0 + r) * lda + j0 + r];
            }
            // the block's row r, loaded before the sequential substitution
            T Lr[BW];
#pragma unroll
            for (int c = 0; c < BW; c++)
                Lr[c] = mload(A, (j0 + min(c, jb - 1)) * lda + j0 + min(r, jb - 1), (j0 + c) * lda + j0 + r, r < jb && c < r);
            const T idg = T(1) / dg;
#pragma unroll
            for (int c = 0; c < BW; c++) {
                if (c < jb) {
                    const T yc = readlane_t(t * idg, c);
                    if (r == c) t = yc;
                    else if (r > c) t -= Lr[c] * yc;
                }
            }
            if (r < jb) sh.ys[j0 + r] = t;
        }
        __syncthreads();
    }
    // backward: x = L^-T y
    for (int bj = nblk - 1; bj >= 0; bj--) {
        const int j0 = bj * BW, jb = min(BW, n - j0);
        {
            const int r = tid >> 3, sl = tid & 7;
            T acc = T(0);
            if (r < jb) {
                int k = j0 + jb + sl;
                for (; k + 56 < n; k += 64) {
                    T a[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) a[u] = A[(j0 + r) * lda + k + 8 * u];
#pragma unroll
                    for (int u = 0; u < 8; u++) acc += a[u] * sh.ys[k + 8 * u];
                }
                for (; k < n; k += 8) acc += A[(j0 + r) * lda + k] * sh.ys[k];
            }
            sh.red[sl][r] = acc;
        }
        __syncthreads();
        if (tid < 64) {
            const int r = tid;
            T t = T(0), dg = T(1);
            if (r < jb) {
                t = sh.ys[j0 + r];
                for (int sl = 0; sl < 8; sl++) t -= sh.red[sl][r];
                dg = A[(j0 + r) * lda + j0 + r];
            }
            // the block's column r (L(c, r), c > r), loaded before the substitution
            T Lc[BW];
#pragma unroll
            for (int c = 0; c < BW; c++)
                Lc[c] = mload(A, (j0 + min(r, jb - 1)) * lda + j0 + min(c, jb - 1), (j0 + r) * lda + j0 + c,
                              r < jb && c > r && c < jb);
            const T idg = T(1) / dg;
#pragma unroll
            for (int c = BW - 1; c >= 0; c--) {
                if (c < jb) {
                    const T xc = readlane_t(t * idg, c);
                    if (r == c) t = xc;
                    else if (r < c) t -= Lc[c] * xc;
                }
            }
            if (r < jb) sh.ys[j0 + r] = t;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += 256) b[i] = sh.ys[i];
}

template <class T>
__global__ __launch_bounds__(256, 2) void k_posv_batched(int count, const int *__restrict__ order,
                                                        const int *__restrict__ nsz,
                                                        const int64_t *__restrict__ aoff,
                                                        const int64_t *__restrict__ boff, T *__restrict__ A,
                                                        T *__restrict__ B, int *__restrict__ info,
                                                        int *__restrict__ counter, int nsplit)
{
    __shared__ BShared<T> sh;
    for (;;) {
        if (threadIdx.x == 0) sh.idx = atomicAdd(counter, 1);
        __syncthreads();
        const int q = sh.idx;
        __syncthreads();
        if (q >= count) break;
        const int m = order[q];
        const int n = nsz[m];
        T *Am = A + aoff[m];
        const int f = factor_one<T>(Am, n, n <= nsplit, sh);
        if (threadIdx.x == 0) info[m] = f;
        if (f == 0 && B) solve_one<T>(Am, n, B + boff[m], sh);
        __syncthreads();
    }
}

__global__ void k_batch_reset(int *counter)
{
    if (threadIdx.x == 0) counter[threadIdx.x] = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
struct mi_batch_s {
    mi_ctx_t *ctx = nullptr;
    int count = 0;
    std::vector<int> n;
    int64_t a_elems = 0, b_elems = 0;
    int *d_n = nullptr, *d_order = nullptr, *d_info = nullptr, *d_counter = nullptr;
    int64_t *d_aoff = nullptr, *d_boff = nullptr;
    int nsplit = 128, wg_per_cu = 2, cus = 0;
    // captured graph of the last (A, B, type, shape) it ran on
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    const void *g_a = nullptr, *g_b = nullptr;
    int g_type = -1, g_nsplit = -1, g_wg = -1;
    double last_ms = 0.0;
};

namespace {

void batch_drop_graph(mi_batch_t *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

template <class T>
int batch_enqueue(mi_batch_t *p, T *dA, T *dB, hipStream_t s)
{
    k_batch_reset<<<1, 1, 0, s>>>(p->d_counter);
    const unsigned grid = (unsigned)std::max(1, std::min(p->count, p->wg_per_cu * p->cus));
    k_posv_batched<T><<<grid, 256, 0, s>>>(p->count, p->d_order, p->d_n, p->d_aoff, p->d_boff, dA, dB, p->d_info,
                                          p->d_counter, p->nsplit);
    MI_CHECK(hipGetLastError());
    return 0;
}

// the same kernel launched once per matrix (one workgroup each), in queue
// order: the launch-loop baseline of the persistent launch
template <class T>
int batch_run_loop(mi_batch_t *p, T *dA, T *dB)
{
    mi_ctx_t *c = p->ctx;
    hipStream_t s = c->s_main;
    if (p->count == 0) return 0;
    MI_CHECK(hipEventRecord(c->t0, s));
    for (int q = 0; q < p->count; q++) {
        k_batch_reset<<<1, 1, 0, s>>>(p->d_counter);
        k_posv_batched<T><<<1, 256, 0, s>>>(1, p->d_order + q, p->d_n, p->d_aoff, p->d_boff, dA, dB, p->d_info,
                                           p->d_counter, p->nsplit);
    }
    MI_CHECK(hipGetLastError());
    MI_CHECK(hipEventRecord(c->t1, s));
    int rc = mi_synchronize(c);
    if (rc) return rc;
    float ms = 0.f;
    MI_CHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
    p->last_ms = ms;
    c->last_ms = ms;
    return 0;
}

template <class T>
int batch_run(mi_batch_t *p, T *dA, T *dB, int use_graph)
{
    mi_ctx_t *c = p->ctx;
    hipStream_t s = c->s_main;
    const int type = sizeof(T) == 8 ? 0 : 1;
    int rc;
    if (p->count == 0) return 0;
    if (use_graph) {
        if (!p->exec || p->g_a != dA || p->g_b != dB || p->g_type != type || p->g_nsplit != p->nsplit ||
            p->g_wg != p->wg_per_cu) {
            batch_drop_graph(p);
            MI_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            rc = batch_enqueue<T>(p, dA, dB, s);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            MI_CHECK(e);
            p->graph = g;
            MI_CHECK(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
            p->g_a = dA;
            p->g_b = dB;
            p->g_type = type;
            p->g_nsplit = p->nsplit;
            p->g_wg = p->wg_per_cu;
        }
        MI_CHECK(hipEventRecord(c->t0, s));
        MI_CHECK(hipGraphLaunch(p->exec, s));
        MI_CHECK(hipEventRecord(c->t1, s));
    } else {
        MI_CHECK(hipEventRecord(c->t0, s));
        if ((rc = batch_enqueue<T>(p, dA, dB, s))) return rc;
        MI_CHECK(hipEventRecord(c->t1, s));
    }
    if ((rc = mi_synchronize(c))) return rc;
    float ms = 0.f;
    MI_CHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
    p->last_ms = ms;
    c->last_ms = ms;
    return 0;
}

template <class T>
int batch_autotune(mi_batch_t *p, const T *dA0, const T *dB0, T *dA, T *dB, int *choice)
{
    mi_ctx_t *c = p->ctx;
    static const int splits[] = {0, 256, 512, BMAX};
    static const int wgs[] = {1, 2};
    double best = 1e300;
    int bs = p->nsplit, bw = p->wg_per_cu, k = 0, bk = 0;
    for (int w : wgs)
        for (int sp : splits) {
            p->nsplit = sp;
            p->wg_per_cu = w;
            double t = 0.0;
            for (int rep = 0; rep < 2; rep++) {     // the first run warms up
                MI_CHECK(hipMemcpyAsync(dA, dA0, (size_t)p->a_elems * sizeof(T), hipMemcpyDeviceToDevice, c->s_main));
                if (dB) MI_CHECK(hipMemcpyAsync(dB, dB0, (size_t)p->b_elems * sizeof(T), hipMemcpyDeviceToDevice, c->s_main));
                int rc = batch_run<T>(p, dA, dB, 0);
                if (rc) return rc;
                t = p->last_ms;
            }
            if (t < best) { best = t; bs = sp; bw = w; bk = k; }
            k++;
        }
    p->nsplit = bs;
    p->wg_per_cu = bw;
    if (choice) *choice = bk;
    return 0;
}

uint64_t sm64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int mi_batch_sizes(uint64_t seed, int count, int lo, int hi, int *n)
{
    if (count < 0 || lo < 1 || hi < lo) { set_error("bad arguments"); return -2; }
    for (int i = 0; i < count; i++)
        n[i] = lo + (int)(sm64(seed + (uint64_t)i * 0x9E3779B97F4A7C15ULL) % (uint64_t)(hi - lo + 1));
    return 0;
}

int mi_batch_create(mi_ctx_t *c, int count, const int *n, mi_batch_t **plan)
{
    *plan = nullptr;
    if (count < 0) { set_error("bad count"); return -2; }
    for (int i = 0; i < count; i++)
        if (n[i] < 1 || n[i] > BMAX) { set_error("matrix %d: order %d outside [1, %d]", i, n[i], BMAX); return -2; }
    mi_batch_t *p = new mi_batch_t();
    p->ctx = c;
    p->count = count;
    p->n.assign(n, n + count);
    std::vector<int64_t> ao(count), bo(count);
    for (int i = 0; i < count; i++) {
        ao[i] = p->a_elems;
        bo[i] = p->b_elems;
        p->a_elems += (int64_t)n[i] * n[i];
        p->b_elems += n[i];
    }
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (p->cus <= 0) p->cus = 256;
    const size_t m = std::max(1, count);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&p->d_n, m * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_order, m * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_info, m * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_counter, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_aoff, m * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&p->d_boff, m * sizeof(int64_t));
    if (e == hipSuccess && count) e = hipMemcpy(p->d_n, n, count * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && count) e = hipMemcpy(p->d_order, order.data(), count * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && count) e = hipMemcpy(p->d_aoff, ao.data(), count * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && count) e = hipMemcpy(p->d_boff, bo.data(), count * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("mi_batch_create: %s", hipGetErrorString(e));
        mi_batch_destroy(p);
        return -1000 - (int)e;
    }
    *plan = p;
    return 0;
}

int mi_batch_destroy(mi_batch_t *p)
{
    if (!p) return 0;
    batch_drop_graph(p);
    (void)hipFree(p->d_n);
    (void)hipFree(p->d_order);
    (void)hipFree(p->d_info);
    (void)hipFree(p->d_counter);
    (void)hipFree(p->d_aoff);
    (void)hipFree(p->d_boff);
    delete p;
    return 0;
}

int mi_batch_extent(const mi_batch_t *p, int64_t *a_elems, int64_t *b_elems)
{
    *a_elems = p->a_elems;
    *b_elems = p->b_elems;
    return 0;
}

int mi_batch_set_shape(mi_batch_t *p, int nsplit, int wg_per_cu)
{
    if (wg_per_cu < 1 || wg_per_cu > 2) { set_error("workgroups per CU must be 1 or 2"); return -2; }
    p->nsplit = nsplit;
    p->wg_per_cu = wg_per_cu;
    return 0;
}

int mi_batch_get_shape(const mi_batch_t *p, int *nsplit, int *wg_per_cu)
{
    *nsplit = p->nsplit;
    *wg_per_cu = p->wg_per_cu;
    return 0;
}

int mi_dposv_batched(mi_batch_t *p, double *dA, double *dB, int use_graph)
{
    return batch_run<double>(p, dA, dB, use_graph);
}

int mi_sposv_batched(mi_batch_t *p, float *dA, float *dB, int use_graph)
{
    return batch_run<float>(p, dA, dB, use_graph);
}

int mi_dposv_batched_loop(mi_batch_t *p, double *dA, double *dB)
{
    return batch_run_loop<double>(p, dA, dB);
}

int mi_sposv_batched_loop(mi_batch_t *p, float *dA, float *dB)
{
    return batch_run_loop<float>(p, dA, dB);
}

int mi_batch_info(const mi_batch_t *p, int *info)
{
    if (p->count) MI_CHECK(hipMemcpy(info, p->d_info, p->count * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int mi_batch_last_ms(const mi_batch_t *p, double *ms)
{
    *ms = p->last_ms;
    return 0;
}

int mi_dbatch_autotune(mi_batch_t *p, const double *dA0, const double *dB0, double *dA, double *dB, int *choice)
{
    return batch_autotune<double>(p, dA0, dB0, dA, dB, choice);
}

int mi_sbatch_autotune(mi_batch_t *p, const float *dA0, const float *dB0, float *dA, float *dB, int *choice)
{
    return batch_autotune<float>(p, dA0, dB0, dA, dB, choice);
}

}  // extern "C"
