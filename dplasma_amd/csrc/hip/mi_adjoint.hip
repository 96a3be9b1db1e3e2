// mi_adjoint.hip -- reverse-mode derivative of the tile Cholesky on gfx950.
//
// Given the factor L (packed lower tiles) of A = L L^T and the gradient Lbar
// of a scalar f with respect to L (lower), the entry points return the
// symmetric gradient Abar (df = sum_ij Abar_ij dA_ij, Abar = Abar^T) in
// place of Lbar, lower tiles.  No counterpart in the reference: this is the
// reverse of its right-looking tile factorization (src/zpotrf_L.jdf:91-633),
// one tile column at a time from the last:
//
//   S1  Lbar_mk -= 2 sum_{n>k} Abar_(m,n) L_nk    (m > k; Abar full symmetric)
//   S2  Abar_mk  = Lbar_mk L_kk^-1 / 2            (m > k)
//   S3  Lbar_kk -= tril(2 sum_{m>k} Abar_mk^T L_mk)
//   S4  Abar_kk  = (S + S^T) / 2,  S = L_kk^-T Phi(L_kk^T Lbar_kk) L_kk^-1
//
// (Phi: lower triangle, diagonal halved).  S1 is the transpose of the
// trailing update and carries 2/3 N^3 of the ~2/3 N^3 + 2 N^2 NB flops.
// The inverses W_k = L_kk^-1 are formed first, for all k at once: 32x32
// diagonal blocks by substitution, then doubling W21 = -W22 L21 W11.
//
// Every product is one launch of k_gemm_jobs over a list of jobs
// C = alpha op(A) B + beta C0 (op(A) = A or A^T, B read down its columns,
// K made of one or more tile segments), MFMA 16x16x4 in 128x128 or 64x64
// blocks as the trailing update, with the K range split over workgroups
// (partials reduced in a fixed order, so results are deterministic) when a
// launch has too few blocks to fill the chip.
#include <algorithm>
#include <vector>

#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

constexpr int GK = 16;   // K depth of one LDS stage

template <class T> struct GSeg {
    const T *A;
    int form;   // 0: op(A) = A (rows of A contiguous), 1: op(A) = A^T
    int pad;
};

// One output matrix C (rows x cols, column-major, ldc) of a launch.
template <class T> struct GJob {
    const T *A;            // segment 0 of A when segs == nullptr
    const GSeg<T> *segs;   // per-segment A, or nullptr (then A + s * a_ss)
    const T *B;            // B(kk, j) at B[j * ldb + kk], segment s at B + s * b_ss
    const T *Cin;          // C0 (may alias Cout; unused when beta == 0)
    T *Cout;
    int64_t a_ss, b_ss;
    T alpha, beta;
    int rows, cols;
    int kseg, klast, nseg; // K = (nseg - 1) * kseg + klast
    int lda, ldb, ldc;
    int aform;             // form of every segment when segs == nullptr
    int amask;             // A stored lower triangular: ignore its strict upper part
    int outmode;           // 0 full, 1 lower (upper set to 0), 2 lower with diagonal halved
    int fast;              // operands 16-byte aligned with even leading dimensions
};

template <class T>
__device__ inline T out_value(T acc, T alpha, T beta, const T *Cin, int64_t at, int rr, int col, int outmode)
{
    T v = alpha * acc;
    if (beta != T(0)) v += beta * Cin[at];
    if (outmode >= 1 && rr < col) v = T(0);
    if (outmode == 2 && rr == col) v *= T(0.5);
    return v;
}

// One workgroup: one BMX x BMX block of one job's C over one split of its
// K stages.  Four waves in a 2x2 arrangement, (BMX/32)^2 MFMA tiles each,
// fed (B, A) so each lane's accumulators walk rows of C (coalesced epilogue).
template <class T, int BMX>
__global__ __launch_bounds__(256, 2) void k_gemm_jobs(const GJob<T> *__restrict__ jobs, int nsub, int KS,
                                                     T *__restrict__ part)
{
    // LDS layouts.  op(A) with rows contiguous in memory (form 0) is staged
    // k-major, As[kk * LP + i] (rows padded by 128 bytes: the four lane
    // groups of a fragment read fall on alternating bank halves).  Operands
    // contiguous along k (B, and op(A) = A^T) are staged row-major with a
    // short pad, Xs[c * KR + kk]: one 16-byte store per element pair, and
    // the 16 rows x 4 k of a fragment read hit distinct banks (KR = 18
    // doubles, 20 floats).
    constexpr int LP = BMX + 128 / (int)sizeof(T);
    constexpr int KR = GK + (sizeof(T) == 8 ? 2 : 4);
    constexpr int SZ = (GK * LP > BMX * KR) ? GK * LP : BMX * KR;
    __shared__ T As[2][SZ];
    __shared__ T Bs[2][SZ];
    typedef typename MF<T>::v4 V4;
    typedef typename MF<T>::v2 V2;
    constexpr int TI = BMX / 32;
    constexpr int HB = BMX / 2;
    constexpr int NP = BMX * GK / 512;     // element pairs per thread and operand
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    const int G = gridDim.x, bid = blockIdx.x;
    const int G8 = (G / 8) * 8;
    const int L = (bid < G8) ? (bid % 8) * (G8 / 8) + bid / 8 : bid;
    const int per_job = nsub * nsub * KS;
    const int j = L / per_job;
    const int rem = L - j * per_job;
    const int blk = rem / KS, s = rem - blk * KS;
    const int bm = blk / nsub, bn = blk - bm * nsub;
    const GJob<T> &J = jobs[j];
    const int rows = J.rows, cols = J.cols;
    const int r_base = bm * BMX, c_base = bn * BMX;
    if (r_base >= rows || c_base >= cols) return;
    const int outmode = J.outmode;
    const bool upper_blk = outmode >= 1 && r_base + BMX <= c_base;
    if (upper_blk) {                            // wholly above the diagonal: zeros
        if (KS == 1) {
            for (int e = tid; e < BMX * BMX; e += 256) {
                const int rr = r_base + (e % BMX), col = c_base + e / BMX;
                if (rr < rows && col < cols) J.Cout[(int64_t)col * J.ldc + rr] = T(0);
            }
        }
        return;
    }

    const int kseg = J.kseg, klast = J.klast, nseg = J.nseg;
    const int lda = J.lda, ldb = J.ldb;
    const int sps = (kseg + GK - 1) / GK, sl = (klast + GK - 1) / GK;
    const int total = nseg > 0 ? (nseg - 1) * sps + sl : 0;
    const int g0 = (int)((int64_t)total * s / KS), g1 = (int)((int64_t)total * (s + 1) / KS);
    const bool rc_full = (r_base + BMX <= rows) && (c_base + BMX <= cols) && J.fast && !J.amask;

    V4 acc[TI][TI];
#pragma unroll
    for (int i = 0; i < TI; i++)
#pragma unroll
        for (int jj = 0; jj < TI; jj++) acc[i][jj] = V4{T(0), T(0), T(0), T(0)};

    V2 ra[NP], rb[NP];
    int st_form = 0;
    auto gload = [&](int g) {
        const int seg = min(g / sps, nseg - 1);
        const int k0 = (g - seg * sps) * GK;
        const int klen = (seg == nseg - 1) ? klast : kseg;
        const T *Ap;
        int form;
        if (J.segs) {
            Ap = J.segs[seg].A;
            form = J.segs[seg].form;
        } else {
            Ap = J.A + seg * J.a_ss;
            form = J.aform;
        }
        const T *Bp = J.B + seg * J.b_ss;
        st_form = form;
        if (rc_full && k0 + GK <= klen) {
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const int e = tid + 256 * p;
                if (form == 0) {
                    const int kk = e / HB, r = (e % HB) * 2;
                    ra[p] = *(const V2 *)(Ap + (int64_t)(k0 + kk) * lda + r_base + r);
                } else {
                    const int r = e / (GK / 2), kk = (e % (GK / 2)) * 2;
                    ra[p] = *(const V2 *)(Ap + (int64_t)(r_base + r) * lda + k0 + kk);
                }
                const int c = e / (GK / 2), kk = (e % (GK / 2)) * 2;
                rb[p] = *(const V2 *)(Bp + (int64_t)(c_base + c) * ldb + k0 + kk);
            }
        } else {
            const bool am = J.amask;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const int e = tid + 256 * p;
                T a0 = T(0), a1 = T(0), b0 = T(0), b1 = T(0);
                if (form == 0) {
                    const int kk = e / HB, r = (e % HB) * 2;
                    const int kg = k0 + kk, rg = r_base + r;
                    if (kg < klen) {
                        if (rg < rows && (!am || rg >= kg)) a0 = Ap[(int64_t)kg * lda + rg];
                        if (rg + 1 < rows && (!am || rg + 1 >= kg)) a1 = Ap[(int64_t)kg * lda + rg + 1];
                    }
                } else {
                    const int r = e / (GK / 2), kk = (e % (GK / 2)) * 2;
                    const int kg = k0 + kk, rg = r_base + r;
                    if (rg < rows) {
                        if (kg < klen && (!am || kg >= rg)) a0 = Ap[(int64_t)rg * lda + kg];
                        if (kg + 1 < klen && (!am || kg + 1 >= rg)) a1 = Ap[(int64_t)rg * lda + kg + 1];
                    }
                }
                const int c = e / (GK / 2), kk = (e % (GK / 2)) * 2;
                const int kg = k0 + kk, cg = c_base + c;
                if (cg < cols) {
                    if (kg < klen) b0 = Bp[(int64_t)cg * ldb + kg];
                    if (kg + 1 < klen) b1 = Bp[(int64_t)cg * ldb + kg + 1];
                }
                ra[p] = V2{a0, a1};
                rb[p] = V2{b0, b1};
            }
        }
    };
    int form0 = 0, form1 = 0;   // layout of the op(A) stage held in buffer 0 / 1
    auto lstore = [&](int buf) {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int e = tid + 256 * p;
            if (st_form == 0) {
                const int kk = e / HB, r = (e % HB) * 2;
                *(V2 *)&As[buf][kk * LP + r] = ra[p];
            } else {
                const int r = e / (GK / 2), kk = (e % (GK / 2)) * 2;
                *(V2 *)&As[buf][r * KR + kk] = ra[p];
            }
            const int c = e / (GK / 2), kk = (e % (GK / 2)) * 2;
            *(V2 *)&Bs[buf][c * KR + kk] = rb[p];
        }
        if (buf) form1 = st_form;
        else form0 = st_form;
    };

    const int fr = lane & 15, fk = lane >> 4;
    if (g0 < g1) {
        gload(g0);
        lstore(0);
        __syncthreads();
        for (int g = g0; g < g1; g++) {
            const int buf = (g - g0) & 1;
            if (g + 1 < g1) gload(g + 1);
#pragma unroll
            for (int kq = 0; kq < GK / 4; kq++) {
                T a[TI], b[TI];
                const int kk = kq * 4 + fk;
                if ((buf ? form1 : form0) == 0) {
#pragma unroll
                    for (int i = 0; i < TI; i++) a[i] = As[buf][kk * LP + wm * HB + i * 16 + fr];
                } else {
#pragma unroll
                    for (int i = 0; i < TI; i++) a[i] = As[buf][(wm * HB + i * 16 + fr) * KR + kk];
                }
#pragma unroll
                for (int jj = 0; jj < TI; jj++) b[jj] = Bs[buf][(wn * HB + jj * 16 + fr) * KR + kk];
#pragma unroll
                for (int i = 0; i < TI; i++)
#pragma unroll
                    for (int jj = 0; jj < TI; jj++) acc[i][jj] = MF<T>::mfma(b[jj], a[i], acc[i][jj]);
            }
            if (g + 1 < g1) lstore(buf ^ 1);
            __syncthreads();
        }
    }

    // element q of lane l: row l % 16, column row(q, l / 16) of each 16x16 tile
    if (KS == 1) {
        const T alpha = J.alpha, beta = J.beta;
        const int ldc = J.ldc;
#pragma unroll
        for (int i = 0; i < TI; i++) {
            const int rr = r_base + wm * HB + i * 16 + fr;
#pragma unroll
            for (int jj = 0; jj < TI; jj++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int col = c_base + wn * HB + jj * 16 + MF<T>::row(q, fk);
                    if (rr < rows && col < cols) {
                        const int64_t at = (int64_t)col * ldc + rr;
                        J.Cout[at] = out_value(acc[i][jj][q], alpha, beta, J.Cin, at, rr, col, outmode);
                    }
                }
        }
    } else {
        T *P = part + ((int64_t)(j * nsub * nsub + blk) * KS + s) * BMX * BMX;
#pragma unroll
        for (int i = 0; i < TI; i++) {
            const int rl = wm * HB + i * 16 + fr;
#pragma unroll
            for (int jj = 0; jj < TI; jj++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int cl = wn * HB + jj * 16 + MF<T>::row(q, fk);
                    P[cl * BMX + rl] = acc[i][jj][q];
                }
        }
    }
}

// Sum of the KS partial blocks in split order, then the job's epilogue.
// RS_SLICE elements of one block per workgroup (8 per thread).
constexpr int RS_SLICE = 2048;

template <class T, int BMX>
__global__ __launch_bounds__(256) void k_gemm_reduce(const GJob<T> *__restrict__ jobs, int nsub, int KS,
                                                    const T *__restrict__ part)
{
    constexpr int NSL = BMX * BMX / RS_SLICE;
    const int L = blockIdx.x / NSL, sl = blockIdx.x - (blockIdx.x / NSL) * NSL;
    const int j = L / (nsub * nsub), blk = L - j * nsub * nsub;
    const int bm = blk / nsub, bn = blk - bm * nsub;
    const GJob<T> &J = jobs[j];
    const int r_base = bm * BMX, c_base = bn * BMX;
    if (r_base >= J.rows || c_base >= J.cols) return;
    const bool upper_blk = J.outmode >= 1 && r_base + BMX <= c_base;
    const T *P = part + (int64_t)(j * nsub * nsub + blk) * KS * BMX * BMX;
    for (int e = sl * RS_SLICE + threadIdx.x; e < (sl + 1) * RS_SLICE; e += 256) {
        const int rl = e % BMX, cl = e / BMX;
        const int rr = r_base + rl, col = c_base + cl;
        if (rr >= J.rows || col >= J.cols) continue;
        const int64_t at = (int64_t)col * J.ldc + rr;
        if (upper_blk) { J.Cout[at] = T(0); continue; }
        T a = T(0);
        for (int s = 0; s < KS; s++) a += P[(int64_t)s * BMX * BMX + e];
        J.Cout[at] = out_value(a, J.alpha, J.beta, J.Cin, at, rr, col, J.outmode);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ p, int64_t n, T v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Inverse of each 32x32 diagonal block of every L_kk into W_k (tile k of W,
// NB x NB, ld NB).  One wave per block; lane c < 32 forms column c by
// forward substitution.  Rows beyond a ragged tile act as identity.
template <class T>
__global__ __launch_bounds__(64) void k_inv32(const T *__restrict__ Lt, int N, int NB, int MT, T *__restrict__ W)
{
    __shared__ T Ls[32][33];
    __shared__ T Ws[32][33];
    const int nblk = (NB + 31) / 32;
    const int k = blockIdx.x / nblk, b = blockIdx.x - k * nblk;
    const int n = tile_rows(N, NB, MT, k);
    const int r0 = 32 * b;
    if (r0 >= n) return;
    const int nb = min(32, n - r0);
    const T *Lk = Lt + tile_index(MT, k, k) * (int64_t)NB * NB;
    T *Wk = W + (int64_t)k * NB * NB;
    for (int e = threadIdx.x; e < 1024; e += 64) {
        const int r = e & 31, c = e >> 5;
        T v = (r == c) ? T(1) : T(0);
        if (r < nb && c < nb) v = (r >= c) ? Lk[(int64_t)(r0 + c) * NB + r0 + r] : T(0);
        Ls[r][c] = v;
    }
    __syncthreads();
    const int c = threadIdx.x;
    if (c < 32) {
        for (int i = 0; i < 32; i++) {
            T w = T(0);
            if (i >= c) {
                T sacc = (i == c) ? T(1) : T(0);
                for (int p = c; p < i; p++) sacc -= Ls[i][p] * Ws[p][c];
                w = sacc / Ls[i][i];
            }
            Ws[i][c] = w;
        }
        if (c < nb)
            for (int i = 0; i < nb; i++) Wk[(int64_t)(r0 + c) * NB + r0 + i] = Ws[i][c];
    }
}

// Abar_kk = (S + S^T) / 2: lower part into the tile (k,k) of Abar, the full
// symmetric tile into D (the trailing operand of later steps).
template <class T>
__global__ __launch_bounds__(256) void k_sym_half(const T *__restrict__ S, int n, int NB, T *__restrict__ Akk,
                                                 T *__restrict__ D)
{
    const int64_t tot = (int64_t)n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
        const T v = T(0.5) * (S[(int64_t)c * NB + r] + S[(int64_t)r * NB + c]);
        D[(int64_t)c * NB + r] = v;
        if (r >= c) Akk[(int64_t)c * NB + r] = v;
    }
}

template <class T> struct AN;
template <> struct AN<double> {
    static constexpr const char *gemm = "dpotrf_adjoint_gemm";
    static constexpr const char *reduce = "dpotrf_adjoint_reduce";
    static constexpr const char *aux = "dpotrf_adjoint_aux";
};
template <> struct AN<float> {
    static constexpr const char *gemm = "spotrf_adjoint_gemm";
    static constexpr const char *reduce = "spotrf_adjoint_reduce";
    static constexpr const char *aux = "spotrf_adjoint_aux";
};


constexpr int64_t PART_CAP = (int64_t)96 << 20;   // bytes of split-K partials per launch

// One planned launch: a range of the job array, its largest output side,
// its K stages per job and the chosen block size and split.
struct Launch {
    int first, njobs, bmx, nsub, ks;
    double flops;
    int kind;      // 0 gemm jobs, 1 sym_half (step k), 2 fill, 3 inv32
    int k;
};

template <class T> bool aligned16(const T *p) { return ((uintptr_t)p & 15) == 0; }

template <class T>
struct Planner {
    std::vector<GJob<T>> jobs;
    std::vector<GSeg<T>> segs;
    std::vector<Launch> plan;
    const GSeg<T> *dsegs = nullptr;   // device address of segs[0]

    void add_job(GJob<T> J, const std::vector<GSeg<T>> *sv)
    {
        bool fast = (J.lda % 2 == 0) && (J.ldb % 2 == 0) && aligned16(J.B) && (J.b_ss % 2 == 0);
        if (sv) {
            for (const GSeg<T> &g : *sv) fast = fast && aligned16(g.A);
            J.segs = (const GSeg<T> *)(uintptr_t)(segs.size() + 1);   // index + 1, patched to a device address later
            segs.insert(segs.end(), sv->begin(), sv->end());
        } else {
            fast = fast && aligned16(J.A) && (J.a_ss % 2 == 0);
            J.segs = nullptr;
        }
        J.fast = fast ? 1 : 0;
        jobs.push_back(J);
    }
    // close the jobs added since 'first' into one launch
    void close(int first, int maxdim, int64_t kstages, double flops)
    {
        Launch l{};
        l.first = first;
        l.njobs = (int)jobs.size() - first;
        l.kind = 0;
        l.flops = flops;
        if (l.njobs == 0) return;
        const int n128 = (maxdim + 127) / 128;
        l.bmx = ((int64_t)l.njobs * n128 * n128 >= 256) ? 128 : 64;
        l.nsub = (maxdim + l.bmx - 1) / l.bmx;
        const int64_t blocks = (int64_t)l.njobs * l.nsub * l.nsub;
        int ks = (int)std::min<int64_t>(16, std::max<int64_t>(1, (512 + blocks - 1) / blocks));
        ks = (int)std::min<int64_t>(ks, std::max<int64_t>(1, kstages / 4));
        while (ks > 1 && blocks * ks * l.bmx * l.bmx * (int64_t)sizeof(T) > PART_CAP) ks--;
        l.ks = ks;
        plan.push_back(l);
    }
    void special(int kind, int k)
    {
        Launch l{};
        l.kind = kind;
        l.k = k;
        plan.push_back(l);
    }
};

template <class T>
GJob<T> job(const T *A, int aform, const T *B, const T *Cin, T *Cout, T alpha, T beta, int rows, int cols, int K,
            int ld, int outmode = 0, int amask = 0)
{
    GJob<T> J{};
    J.A = A;
    J.B = B;
    J.Cin = Cin;
    J.Cout = Cout;
    J.alpha = alpha;
    J.beta = beta;
    J.rows = rows;
    J.cols = cols;
    J.kseg = K;
    J.klast = K;
    J.nseg = K > 0 ? 1 : 0;
    J.lda = J.ldb = J.ldc = ld;
    J.aform = aform;
    J.amask = amask;
    J.outmode = outmode;
    return J;
}

template <class T>
int adjoint_tiles(mi_ctx_t *c, int N, int NB, const T *dL, T *dG)
{
    if (N <= 0) return 0;
    if (NB <= 0) { set_error("bad tile size"); return -2; }
    const int MT = (N + NB - 1) / NB;
    const int64_t T2 = (int64_t)NB * NB;
    hipStream_t s = c->s_main;

    // workspace: W (MT tiles), D (MT tiles), Pb (MT tiles; also the scratch
    // of the inversion), Q, P, X (one tile each)
    const size_t ws_elems = (size_t)(3 * MT + 3) * T2;
    T *ws = (T *)ws_get(c, WS_ADJ_W, ws_elems * sizeof(T));
    if (!ws) return -3;
    T *W = ws, *D = W + MT * T2, *Pb = D + MT * T2, *Q = Pb + MT * T2, *P = Q + T2, *X = P + T2;
    auto tile = [&](const T *base, int m, int n) { return const_cast<T *>(base) + tile_index(MT, m, n) * T2; };

    Planner<T> pl;
    // ---- inverses of the diagonal tiles
    pl.special(2, 0);                                  // W = 0
    pl.special(3, 0);                                  // 32x32 diagonal blocks
    for (int h = 32; h < NB; h *= 2) {
        const int f1 = (int)pl.jobs.size();
        double fl = 0.0;
        for (int k = 0; k < MT; k++) {
            const int n = tile_rows(N, NB, MT, k);
            const T *Lk = tile(dL, k, k);
            for (int p = 0; p + h < n; p += 2 * h) {
                const int q = p + h, rq = std::min(h, n - q);
                // T1 = L21 W11 -> Pb tile k at (q, p)
                pl.add_job(job<T>(Lk + (int64_t)p * NB + q, 0, W + k * T2 + (int64_t)p * NB + p, nullptr,
                                  Pb + k * T2 + (int64_t)p * NB + q, T(1), T(0), rq, h, h, NB), nullptr);
                fl += 2.0 * rq * h * h;
            }
        }
        pl.close(f1, h, (h + GK - 1) / GK, fl);
        const int f2 = (int)pl.jobs.size();
        fl = 0.0;
        for (int k = 0; k < MT; k++) {
            const int n = tile_rows(N, NB, MT, k);
            for (int p = 0; p + h < n; p += 2 * h) {
                const int q = p + h, rq = std::min(h, n - q);
                // W21 = -W22 T1
                pl.add_job(job<T>(W + k * T2 + (int64_t)q * NB + q, 0, Pb + k * T2 + (int64_t)p * NB + q, nullptr,
                                  W + k * T2 + (int64_t)p * NB + q, T(-1), T(0), rq, h, rq, NB), nullptr);
                fl += 2.0 * rq * rq * h;
            }
        }
        pl.close(f2, h, (h + GK - 1) / GK, fl);
    }

    // ---- reverse sweep
    std::vector<GSeg<T>> sv;
    for (int k = MT - 1; k >= 0; k--) {
        const int nk = tile_rows(N, NB, MT, k);
        const int nt = MT - k - 1;                    // tiles below the diagonal
        const int klast = tile_rows(N, NB, MT, MT - 1);
        const T *Wk = W + k * T2;
        if (nt > 0) {
            // S1: Pb[t] = Lbar(m,k) - 2 sum_n Abar(m,n) L(n,k)
            int f = (int)pl.jobs.size();
            double fl = 0.0;
            for (int t = 0; t < nt; t++) {
                const int m = k + 1 + t;
                sv.clear();
                for (int u = 0; u < nt; u++) {
                    const int n = k + 1 + u;
                    GSeg<T> g{};
                    if (n < m) { g.A = tile(dG, m, n); g.form = 0; }
                    else if (n == m) { g.A = D + m * T2; g.form = 0; }
                    else { g.A = tile(dG, n, m); g.form = 1; }
                    sv.push_back(g);
                }
                GJob<T> J = job<T>(nullptr, 0, tile(dL, k + 1, k), tile(dG, m, k), Pb + t * T2, T(-2), T(1),
                                   tile_rows(N, NB, MT, m), NB, NB, NB);
                J.b_ss = T2;
                J.nseg = nt;
                J.kseg = NB;
                J.klast = klast;
                pl.add_job(J, &sv);
                fl += 2.0 * tile_rows(N, NB, MT, m) * NB * ((double)(nt - 1) * NB + klast);
            }
            pl.close(f, NB, ((int64_t)(nt - 1) * NB + klast + GK - 1) / GK, fl);
            // S2: Abar(m,k) = Pb[t] W_k / 2
            f = (int)pl.jobs.size();
            fl = 0.0;
            for (int t = 0; t < nt; t++) {
                const int m = k + 1 + t, rm = tile_rows(N, NB, MT, m);
                pl.add_job(job<T>(Pb + t * T2, 0, Wk, nullptr, tile(dG, m, k), T(0.5), T(0), rm, NB, NB, NB),
                           nullptr);
                fl += 2.0 * rm * NB * NB;
            }
            pl.close(f, NB, (NB + GK - 1) / GK, fl);
        }
        // S3: Q = tril(Lbar(k,k) - 2 sum_m Abar(m,k)^T L(m,k))
        {
            const int f = (int)pl.jobs.size();
            GJob<T> J = job<T>(nt > 0 ? tile(dG, k + 1, k) : nullptr, 1, nt > 0 ? tile(dL, k + 1, k) : nullptr,
                               tile(dG, k, k), Q, T(-2), T(1), nk, nk, NB, NB, 1);
            J.a_ss = T2;
            J.b_ss = T2;
            J.nseg = nt;
            J.klast = klast;
            if (nt == 0) { J.A = Q; J.B = Q; }            // no K: C = tril(Lbar_kk)
            pl.add_job(J, nullptr);
            const double K = nt > 0 ? (double)(nt - 1) * NB + klast : 0.0;
            pl.close(f, nk, (int64_t)(K + GK - 1) / GK, 2.0 * nk * nk * K);
        }
        // S4: P = Phi(L_kk^T Q); X = P W_k; S = W_k^T X (into P); Abar_kk = sym(S)
        {
            int f = (int)pl.jobs.size();
            pl.add_job(job<T>(tile(dL, k, k), 1, Q, nullptr, P, T(1), T(0), nk, nk, nk, NB, 2, 1), nullptr);
            pl.close(f, nk, (nk + GK - 1) / GK, 2.0 * nk * nk * nk);
            f = (int)pl.jobs.size();
            pl.add_job(job<T>(P, 0, Wk, nullptr, X, T(1), T(0), nk, nk, nk, NB), nullptr);
            pl.close(f, nk, (nk + GK - 1) / GK, 2.0 * nk * nk * nk);
            f = (int)pl.jobs.size();
            pl.add_job(job<T>(Wk, 1, X, nullptr, P, T(1), T(0), nk, nk, nk, NB), nullptr);
            pl.close(f, nk, (nk + GK - 1) / GK, 2.0 * nk * nk * nk);
            pl.special(1, k);
        }
    }

    // upload the job and segment tables (one synchronous copy each)
    GSeg<T> *dsegs = nullptr;
    if (!pl.segs.empty()) {
        dsegs = (GSeg<T> *)ws_get(c, WS_ADJ_SEGS, pl.segs.size() * sizeof(GSeg<T>));
        if (!dsegs) return -3;
        MI_CHECK(hipMemcpy(dsegs, pl.segs.data(), pl.segs.size() * sizeof(GSeg<T>), hipMemcpyHostToDevice));
    }
    for (GJob<T> &J : pl.jobs)
        if (J.segs) J.segs = dsegs + ((uintptr_t)J.segs - 1);
    GJob<T> *djobs = (GJob<T> *)ws_get(c, WS_ADJ_JOBS, std::max<size_t>(1, pl.jobs.size()) * sizeof(GJob<T>));
    if (!djobs) return -3;
    MI_CHECK(hipMemcpy(djobs, pl.jobs.data(), pl.jobs.size() * sizeof(GJob<T>), hipMemcpyHostToDevice));
    int64_t part_elems = 1;
    for (const Launch &l : pl.plan)
        if (l.kind == 0 && l.ks > 1)
            part_elems = std::max<int64_t>(part_elems, (int64_t)l.njobs * l.nsub * l.nsub * l.ks * l.bmx * l.bmx);
    T *part = (T *)ws_get(c, WS_ADJ_PART, (size_t)part_elems * sizeof(T));
    if (!part) return -3;

    MI_CHECK(hipEventRecord(c->t0, s));
    for (const Launch &l : pl.plan) {
        if (l.kind == 2) {
            const int64_t n = (int64_t)MT * T2;
            MI_LAUNCH(c, s, AN<T>::aux, 0.0, (double)n * sizeof(T),
                      k_fill<T><<<(unsigned)std::min<int64_t>(65536, (n + 255) / 256), 256, 0, s>>>(W, n, T(0)));
        } else if (l.kind == 3) {
            const int nblk = (NB + 31) / 32;
            MI_LAUNCH(c, s, AN<T>::aux, (double)MT * NB * 32 * 32 / 3.0, 0.0,
                      k_inv32<T><<<MT * nblk, 64, 0, s>>>(dL, N, NB, MT, W));
        } else if (l.kind == 1) {
            const int nk = tile_rows(N, NB, MT, l.k);
            const int64_t n = (int64_t)nk * nk;
            MI_LAUNCH(c, s, AN<T>::aux, 0.0, 3.0 * n * sizeof(T),
                      k_sym_half<T><<<(unsigned)std::min<int64_t>(4096, (n + 255) / 256), 256, 0, s>>>(
                          P, nk, NB, tile(dG, l.k, l.k), D + l.k * T2));
        } else {
            const GJob<T> *jj = djobs + l.first;
            const unsigned grid = (unsigned)((int64_t)l.njobs * l.nsub * l.nsub * l.ks);
            if (l.bmx == 128)
                MI_LAUNCH(c, s, AN<T>::gemm, l.flops, 0.0, k_gemm_jobs<T, 128><<<grid, 256, 0, s>>>(jj, l.nsub, l.ks, part));
            else
                MI_LAUNCH(c, s, AN<T>::gemm, l.flops, 0.0, k_gemm_jobs<T, 64><<<grid, 256, 0, s>>>(jj, l.nsub, l.ks, part));
            if (l.ks > 1) {
                const unsigned rg = (unsigned)((int64_t)l.njobs * l.nsub * l.nsub * l.bmx * l.bmx / RS_SLICE);
                if (l.bmx == 128)
                    MI_LAUNCH(c, s, AN<T>::reduce, 0.0, 0.0, k_gemm_reduce<T, 128><<<rg, 256, 0, s>>>(jj, l.nsub, l.ks, part));
                else
                    MI_LAUNCH(c, s, AN<T>::reduce, 0.0, 0.0, k_gemm_reduce<T, 64><<<rg, 256, 0, s>>>(jj, l.nsub, l.ks, part));
            }
        }
    }
    MI_CHECK(hipEventRecord(c->t1, s));
    int rc = mi_synchronize(c);
    if (rc) return rc;
    float ms = 0.f;
    MI_CHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_ms = ms;
    return 0;
}

// Lbar = diag(2 / L_ii) (the gradient of logdet A = 2 sum log L_ii) in
// packed lower tiles, zero elsewhere; *info = 1-based index of the first
// zero diagonal entry of L (atomicMin over a device word holding N + 1).
__global__ __launch_bounds__(256) void k_logdet_seed(const double *__restrict__ L, int N, int NB, int MT,
                                                    double *__restrict__ G, int *__restrict__ zero_at)
{
    // one pass, one value per element (no ordering between workgroups needed)
    const int64_t tot = (int64_t)MT * (MT + 1) / 2 * NB * NB;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e / ((int64_t)NB * NB);
        const int64_t w = e - t * NB * NB;
        const int c = (int)(w / NB), r = (int)(w - (int64_t)c * NB);
        int n = 0;
        while (t >= MT - n) { t -= MT - n; n++; }
        const bool diag = (t == 0) && (r == c) && (n * NB + r < N);
        double v = 0.0;
        if (diag) {
            const double d = L[e];
            if (d == 0.0) atomicMin(zero_at, n * NB + r + 1);
            v = 2.0 / d;
        }
        G[e] = v;
    }
}

// D = S on the lower part of the packed lower tiles (diagonal tiles: lower
// triangle only; padding rows untouched)
__global__ __launch_bounds__(256) void k_copy_lower_tiles(const double *__restrict__ S, double *__restrict__ D, int N,
                                                         int NB, int MT)
{
    const int64_t tot = (int64_t)MT * (MT + 1) / 2 * NB * NB;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e / ((int64_t)NB * NB);
        const int64_t w = e - t * NB * NB;
        const int c = (int)(w / NB), r = (int)(w - (int64_t)c * NB);
        int n = 0;
        while (t >= MT - n) { t -= MT - n; n++; }
        const int m = n + (int)t;
        if (m * NB + r >= N || n * NB + c >= N) continue;
        if (m == n && r < c) continue;
        D[e] = S[e];
    }
}


}  // namespace

extern "C" {

// In-place inverse of the SPD matrix from its Cholesky factor: the adjoint
// of logdet A = 2 sum log L_ii, whose gradient with respect to A is A^-1.
int mi_dpotri_tiles(mi_ctx_t *c, int N, int NB, double *dT, int *info)
{
    *info = 0;
    if (N <= 0) return 0;
    if (NB <= 0) { set_error("bad tile size"); return -2; }
    const int MT = (N + NB - 1) / NB;
    const int64_t tot = (int64_t)MT * (MT + 1) / 2 * NB * NB;
    double *G = (double *)ws_get(c, WS_POTRI, (size_t)tot * sizeof(double));
    if (!G) return -3;
    int h_zero = 0;
    k_fill<int><<<1, 64, 0, c->s_main>>>(c->d_info, 1, N + 1);
    k_logdet_seed<<<(unsigned)std::min<int64_t>(65536, (tot + 255) / 256), 256, 0, c->s_main>>>(dT, N, NB, MT, G,
                                                                                              c->d_info);
    MI_CHECK(hipGetLastError());
    MI_CHECK(hipMemcpyAsync(&h_zero, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->s_main));
    int rc = mi_synchronize(c);
    if (rc) return rc;
    if (h_zero <= N) { *info = h_zero; return 0; }      // singular factor: LAPACK dtrtri's info
    if ((rc = adjoint_tiles<double>(c, N, NB, dT, G))) return rc;
    const double ms = c->last_ms;
    k_copy_lower_tiles<<<(unsigned)std::min<int64_t>(65536, (tot + 255) / 256), 256, 0, c->s_main>>>(G, dT, N, NB, MT);
    MI_CHECK(hipGetLastError());
    rc = mi_synchronize(c);
    c->last_ms = ms;
    return rc;
}

int mi_dpotrf_adjoint_tiles(mi_ctx_t *c, int N, int NB, const double *dL, double *dLbar)
{
    return adjoint_tiles<double>(c, N, NB, dL, dLbar);
}

int mi_spotrf_adjoint_tiles(mi_ctx_t *c, int N, int NB, const float *dL, float *dLbar)
{
    return adjoint_tiles<float>(c, N, NB, dL, dLbar);
}

}  // extern "C"
