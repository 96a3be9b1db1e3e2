// mi_util.hip -- generators, layout transforms and the residual checks of
// the gfx950 kernel library.
//
// Generators: the LCG of src/cores/random.h:19-41 (X_{t+1} = A X_t + C mod
// 2^64, jump-ahead by binary powering) and the entry formula of
// src/cores/core_zplgsy.c:89 / core_zplrnt.c:81, 0.5f - (float)X * 2^-64 in
// single precision, widened to double; each thread jumps once to the start
// of a 64-row column segment and then steps the recurrence, as the
// reference walks a tile column.
//
// Checks: check_dpotrf / check_daxmb of src/dplasma_zcheck.c:63-133,179-224
// computed on the device (R = L L^T - A0 on the lower tiles, R = B - A0 X by
// the strip residual kernel of mi_mixed.hip, infinity norms of symmetric and
// general matrices by row-sum atomics).
#include <algorithm>

#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

__device__ __host__ inline uint64_t rnd_jump(uint64_t n, uint64_t seed)
{
    uint64_t a = 6364136223846793005ULL, c = 1ULL, x = seed;
    while (n) {
        if (n & 1ULL) x = a * x + c;
        c = c * (a + 1ULL);
        a = a * a;
        n >>= 1;
    }
    return x;
}

__device__ inline double rnd_entry(uint64_t x)
{
    // float(x) rounds to nearest; the scale by 2^-64 is exact in float
    const float f = (float)x;
    const float s = __fmul_rn(f, 5.4210108624275222e-20f);
    return (double)__fsub_rn(0.5f, s);
}

constexpr int SEG = 64;

// dplgsy on packed lower tiles.  Grid: (MT*(MT+1)/2 tiles) x (NB columns);
// blockDim = 64 threads, thread = a 64-row segment... (loop over segments).
__global__ __launch_bounds__(256) void k_dplgsy(double *__restrict__ T, int N, int NB, int MT,
                                               double bump, uint64_t seed)
{
    const int64_t ntiles = (int64_t)MT * (MT + 1) / 2;
    const int nseg = (NB + SEG - 1) / SEG;
    const int64_t work = ntiles * NB * nseg;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < work;
         w += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = w / ((int64_t)NB * nseg);
        const int rem = (int)(w - t * (int64_t)NB * nseg);
        const int c = rem / nseg, sg = rem % nseg;
        int n = 0;
        while (t >= MT - n) { t -= MT - n; n++; }
        const int m = n + (int)t;
        const int rows = tile_rows(N, NB, MT, m), cols = tile_rows(N, NB, MT, n);
        if (c >= cols) continue;
        double *tile = T + tile_index(MT, m, n) * (int64_t)NB * NB + (int64_t)c * NB;
        const int r0 = sg * SEG, r1 = min(rows, r0 + SEG);
        const uint64_t gi = (uint64_t)m * NB, gj = (uint64_t)n * NB + c;
        if (m != n) {
            uint64_t x = rnd_jump(gi + r0 + gj * (uint64_t)N, seed);
            for (int r = r0; r < r1; r++) {
                tile[r] = rnd_entry(x);
                x = 6364136223846793005ULL * x + 1ULL;
            }
        } else {
            // lower part (r >= c) walks the column; the diagonal gets the bump;
            // the upper part (r < c) mirrors entry (c, r)
            int rl = max(r0, c);
            if (rl < r1) {
                uint64_t x = rnd_jump(gi + rl + gj * (uint64_t)N, seed);
                for (int r = rl; r < r1; r++) {
                    double v = rnd_entry(x);
                    if (r == c) v += bump;
                    tile[r] = v;
                    x = 6364136223846793005ULL * x + 1ULL;
                }
            }
            for (int r = r0; r < min(r1, c); r++) {
                const uint64_t idx = (gi + c) + ((uint64_t)n * NB + r) * (uint64_t)N;
                tile[r] = rnd_entry(rnd_jump(idx, seed));
            }
        }
    }
}

// dplrnt into a column-major M x Ncols block at (m0, n0) of a gM-row matrix.
__global__ __launch_bounds__(256) void k_dplrnt(double *__restrict__ A, int M, int Nc, int lda, int gM,
                                               int m0, int n0, uint64_t seed)
{
    const int nseg = (M + SEG - 1) / SEG;
    const int64_t work = (int64_t)Nc * nseg;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < work;
         w += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(w / nseg), sg = (int)(w % nseg);
        const int r0 = sg * SEG, r1 = min(M, r0 + SEG);
        uint64_t x = rnd_jump((uint64_t)(m0 + r0) + (uint64_t)(n0 + c) * (uint64_t)gM, seed);
        double *col = A + (int64_t)c * lda;
        for (int r = r0; r < r1; r++) {
            col[r] = rnd_entry(x);
            x = 6364136223846793005ULL * x + 1ULL;
        }
    }
}

// Packed upper tiles <-> packed lower tiles holding the transposes.
// Grid: tiles x (NB/32 x NB/32 blocks) handled by a strided loop; 32x32 LDS transpose.
__global__ __launch_bounds__(256) void k_sym_transpose(const double *__restrict__ S, double *__restrict__ D,
                                                      int N, int NB, int MT, int upper_to_lower)
{
    __shared__ double buf[32][33];
    const int nb32 = (NB + 31) / 32;
    const int64_t ntiles = (int64_t)MT * (MT + 1) / 2;
    const int64_t work = ntiles * nb32 * nb32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int64_t w = blockIdx.x; w < work; w += gridDim.x) {
        int64_t t = w / (nb32 * nb32);
        const int b = (int)(w % (nb32 * nb32));
        const int bi = b / nb32, bj = b % nb32;
        // lower tile (m, n), m >= n, enumerated column by column
        int n = 0;
        while (t >= MT - n) { t -= MT - n; n++; }
        const int m = n + (int)t;
        // upper tile (n, m) sits at index m*(m+1)/2 + n
        const int64_t li = tile_index(MT, m, n), ui = (int64_t)m * (m + 1) / 2 + n;
        const double *src = S + (upper_to_lower ? ui : li) * (int64_t)NB * NB;
        double *dst = D + (upper_to_lower ? li : ui) * (int64_t)NB * NB;
        // dst(r, c) = src(c, r); block (bi, bj) of dst
        const int r0 = bi * 32, c0 = bj * 32;
        __syncthreads();
        for (int y = ty; y < 32; y += 8) {
            const int sr = c0 + tx, sc = r0 + y;       // src(row = c0+tx, col = r0+y)
            buf[y][tx] = (sr < NB && sc < NB) ? src[(int64_t)sc * NB + sr] : 0.0;
        }
        __syncthreads();
        for (int y = ty; y < 32; y += 8) {
            const int dr = r0 + tx, dc = c0 + y;
            if (dr < NB && dc < NB) dst[(int64_t)dc * NB + dr] = buf[tx][y];
        }
    }
}

__global__ __launch_bounds__(256) void k_tiles_colmajor(const double *__restrict__ S, double *__restrict__ D,
                                                       int M, int N, int MB, int NB, int lda, int to_cm)
{
    const int64_t tot = (int64_t)M * N;
    const int MT = (M + MB - 1) / MB;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < tot;
         w += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(w / M), i = (int)(w % M);
        const int m = i / MB, n = j / NB;
        const int64_t toff = ((int64_t)n * MT + m) * MB * NB + (int64_t)(j - n * NB) * MB + (i - m * MB);
        if (to_cm) D[(int64_t)j * lda + i] = S[toff];
        else D[toff] = S[(int64_t)j * lda + i];
    }
}

// ---------------------------------------------------------------- checks
// R(m,n) block = sum_{p <= n} L(m,p) L(n,p)^T - A0(m,n) for one 64x64 block
// of a lower tile; row sums of |R| (symmetric) accumulated into rs[N].
// L's diagonal tiles are read as lower triangular (their strict upper part
// still holds A).
__global__ __launch_bounds__(256) void k_check_llt(const double *__restrict__ L, const double *__restrict__ A0,
                                                  int N, int NB, int MT, double *__restrict__ rs)
{
    __shared__ double As[16][65], Bs[16][65];
    const int nsub = (NB + 63) / 64;
    int64_t t = blockIdx.x / (nsub * nsub);
    const int s = blockIdx.x % (nsub * nsub);
    int n = 0;
    while (t >= MT - n) { t -= MT - n; n++; }
    const int m = n + (int)t;
    const int bm = s / nsub, bn = s % nsub;
    if (m == n && bm < bn) return;
    const int rows = tile_rows(N, NB, MT, m), cols = tile_rows(N, NB, MT, n);
    const int r0 = bm * 64, c0 = bn * 64;
    if (r0 >= rows || c0 >= cols) return;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // 16 x 16, 4x4 each
    double acc[4][4] = {};
    for (int p = 0; p <= n; p++) {
        const double *Lm = L + tile_index(MT, m, p) * (int64_t)NB * NB;
        const double *Ln = L + tile_index(MT, n, p) * (int64_t)NB * NB;
        const int kp = tile_rows(N, NB, MT, p);
        for (int k0 = 0; k0 < kp; k0 += 16) {
            __syncthreads();
            for (int idx = threadIdx.x; idx < 16 * 64; idx += 256) {
                const int kk = idx / 64, r = idx % 64;
                const int kg = k0 + kk, ra = r0 + r, rb = c0 + r;
                double va = 0.0, vb = 0.0;
                if (kg < kp) {
                    if (ra < rows && (p != m || ra >= kg)) va = Lm[(int64_t)kg * NB + ra];
                    if (rb < cols && (p != n || rb >= kg)) vb = Ln[(int64_t)kg * NB + rb];
                }
                As[kk][r] = va;
                Bs[kk][r] = vb;
            }
            __syncthreads();
#pragma unroll 4
            for (int kk = 0; kk < 16; kk++) {
                double a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; i++) { a[i] = As[kk][ty * 4 + i]; b[i] = Bs[kk][tx * 4 + i]; }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
            }
        }
    }
    const double *A = A0 + tile_index(MT, m, n) * (int64_t)NB * NB;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = r0 + ty * 4 + i, c = c0 + tx * 4 + j;
            if (r >= rows || c >= cols) continue;
            if (m == n && r < c) continue;
            const double v = fabs(acc[i][j] - A[(int64_t)c * NB + r]);
            atomicAdd(&rs[m * NB + r], v);
            if (!(m == n && r == c)) atomicAdd(&rs[n * NB + c], v);
        }
}

// Row sums of |A| for a symmetric matrix in packed lower tiles.
__global__ __launch_bounds__(256) void k_sym_rowsum(const double *__restrict__ A, int N, int NB, int MT,
                                                   double *__restrict__ rs)
{
    const int64_t ntiles = (int64_t)MT * (MT + 1) / 2;
    const int64_t work = ntiles * NB;
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < work;
         w += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = w / NB;
        const int c = (int)(w % NB);
        int n = 0;
        while (t >= MT - n) { t -= MT - n; n++; }
        const int m = n + (int)t;
        const int rows = tile_rows(N, NB, MT, m), cols = tile_rows(N, NB, MT, n);
        if (c >= cols) continue;
        const double *col = A + tile_index(MT, m, n) * (int64_t)NB * NB + (int64_t)c * NB;
        double cs = 0.0;
        for (int r = (m == n ? c : 0); r < rows; r++) {
            const double v = fabs(col[r]);
            if (!(m == n && r == c)) atomicAdd(&rs[m * NB + r], v);
            cs += v;
        }
        atomicAdd(&rs[n * NB + c], cs);
    }
}

__global__ __launch_bounds__(256) void k_ge_rowsum(const double *__restrict__ A, int M, int Nc, int lda,
                                                  double *__restrict__ rs)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M;
         i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < Nc; j++) s += fabs(A[(int64_t)j * lda + i]);
        rs[i] = s;
    }
}

// max over n values (NaN propagates), single workgroup
__global__ __launch_bounds__(256) void k_max(const double *__restrict__ v, int64_t n, double *__restrict__ out)
{
    __shared__ double sm[256];
    double mx = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double a = v[i];
        if (a > mx || a != a) mx = a;
    }
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double a = sm[threadIdx.x + s];
            if (a > sm[threadIdx.x] || a != a) sm[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}

__global__ __launch_bounds__(256) void k_dcopy(const double2 *__restrict__ s, double2 *__restrict__ d,
                                              int64_t n2)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// symmetric matrix in packed lower tiles -> full N x N column-major (ld N)
__global__ __launch_bounds__(256) void k_sym_to_full(const double *__restrict__ T, int N, int NB, int MT,
                                                    double *__restrict__ F)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)N * N;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / N), i = (int)(e - (int64_t)j * N);
        const int r = i >= j ? i : j, c = i >= j ? j : i;      // lower-triangle element
        const int m = r / NB, n = c / NB;
        F[e] = T[tile_index(MT, m, n) * (int64_t)NB * NB + (int64_t)(c - n * NB) * NB + (r - m * NB)];
    }
}

// column sums of |A| (M x Nc, ld lda)
__global__ __launch_bounds__(256) void k_ge_colsum(const double *__restrict__ A, int M, int Nc, int64_t lda,
                                                  double *__restrict__ cs)
{
    for (int j = blockIdx.x; j < Nc; j += gridDim.x) {
        __shared__ double sm[256];
        double a = 0.0;
        for (int i = threadIdx.x; i < M; i += 256) a += fabs(A[(int64_t)j * lda + i]);
        sm[threadIdx.x] = a;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) cs[j] = sm[0];
        __syncthreads();
    }
}

__global__ void k_add_diag(double *__restrict__ A, int n, int lda, double alpha)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) A[i * lda + i] += alpha;
}

__global__ void k_zero(double *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

inline unsigned grid_for(int64_t work, int per = 256)
{
    int64_t g = (work + per - 1) / per;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

int dmax_of(Ctx *c, const double *d, int64_t n, double *out)
{
    double *o = c->d_scratch;
    k_max<<<1, 256, 0, c->s_main>>>(d, n, o);
    MI_CHECK(hipGetLastError());
    MI_CHECK(hipMemcpyAsync(out, o, sizeof(double), hipMemcpyDeviceToHost, c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

}  // namespace

extern "C" {

int mi_dplgsy_tiles(mi_ctx_t *c, double bump, int N, int NB, double *dT, uint64_t seed)
{
    if (N <= 0) return 0;
    const int MT = (N + NB - 1) / NB;
    const int64_t work = (int64_t)MT * (MT + 1) / 2 * NB * ((NB + SEG - 1) / SEG);
    MI_LAUNCH(c, c->s_main, "dplgsy", 0.0, (double)MT * (MT + 1) / 2 * NB * NB * 8.0,
              k_dplgsy<<<grid_for(work), 256, 0, c->s_main>>>(dT, N, NB, MT, bump, seed));
    return mi_synchronize(c);
}

int mi_dplrnt(mi_ctx_t *c, int M, int Nc, double *dA, int lda, int gM, int m0, int n0, uint64_t seed)
{
    if (M <= 0 || Nc <= 0) return 0;
    const int64_t work = (int64_t)Nc * ((M + SEG - 1) / SEG);
    MI_LAUNCH(c, c->s_main, "dplrnt", 0.0, (double)M * Nc * 8.0,
              k_dplrnt<<<grid_for(work), 256, 0, c->s_main>>>(dA, M, Nc, lda, gM, m0, n0, seed));
    return mi_synchronize(c);
}

int mi_dcopy(mi_ctx_t *c, size_t n, const double *s, double *d)
{
    if (n == 0) return 0;
    if ((n & 1) || ((uintptr_t)s & 15) || ((uintptr_t)d & 15)) { set_error("mi_dcopy: even n, 16B alignment"); return -2; }
    const int64_t n2 = (int64_t)(n / 2);
    MI_LAUNCH(c, c->s_main, "dcopy", 0.0, 16.0 * n, k_dcopy<<<2048, 256, 0, c->s_main>>>(
                  (const double2 *)s, (double2 *)d, n2));
    return mi_synchronize(c);
}

int mi_dadd_diag(mi_ctx_t *c, int n, double *dA, int lda, double alpha)
{
    if (n <= 0) return 0;
    MI_LAUNCH(c, c->s_main, "add_diag", 0.0, 0.0,
              k_add_diag<<<grid_for(n), 256, 0, c->s_main>>>(dA, n, lda, alpha));
    return mi_synchronize(c);
}

int mi_dsym_upper_to_lower(mi_ctx_t *c, int N, int NB, const double *dU, double *dL)
{
    if (N <= 0) return 0;
    const int MT = (N + NB - 1) / NB;
    const int nb32 = (NB + 31) / 32;
    const int64_t work = (int64_t)MT * (MT + 1) / 2 * nb32 * nb32;
    MI_LAUNCH(c, c->s_main, "sym_transpose", 0.0, 0.0,
              k_sym_transpose<<<grid_for(work, 1), 256, 0, c->s_main>>>(dU, dL, N, NB, MT, 1));
    return mi_synchronize(c);
}

int mi_dsym_lower_to_upper(mi_ctx_t *c, int N, int NB, const double *dL, double *dU)
{
    if (N <= 0) return 0;
    const int MT = (N + NB - 1) / NB;
    const int nb32 = (NB + 31) / 32;
    const int64_t work = (int64_t)MT * (MT + 1) / 2 * nb32 * nb32;
    MI_LAUNCH(c, c->s_main, "sym_transpose", 0.0, 0.0,
              k_sym_transpose<<<grid_for(work, 1), 256, 0, c->s_main>>>(dL, dU, N, NB, MT, 0));
    return mi_synchronize(c);
}

int mi_dtiles_to_colmajor(mi_ctx_t *c, int M, int N, int MB, int NB, const double *dT, double *dA, int lda)
{
    if (M <= 0 || N <= 0) return 0;
    MI_LAUNCH(c, c->s_main, "tiles_colmajor", 0.0, 0.0,
              k_tiles_colmajor<<<grid_for((int64_t)M * N), 256, 0, c->s_main>>>(dT, dA, M, N, MB, NB, lda, 1));
    return mi_synchronize(c);
}

int mi_dcolmajor_to_tiles(mi_ctx_t *c, int M, int N, int MB, int NB, const double *dA, int lda, double *dT)
{
    if (M <= 0 || N <= 0) return 0;
    MI_LAUNCH(c, c->s_main, "tiles_colmajor", 0.0, 0.0,
              k_tiles_colmajor<<<grid_for((int64_t)M * N), 256, 0, c->s_main>>>(dA, dT, M, N, MB, NB, lda, 0));
    return mi_synchronize(c);
}

int mi_dcheck_potrf(mi_ctx_t *c, int N, int NB, const double *dL, const double *dA0, double *Anorm,
                    double *Rnorm)
{
    if (N <= 0) { *Anorm = *Rnorm = 0.0; return 0; }
    const int MT = (N + NB - 1) / NB;
    const int64_t npad = (int64_t)MT * NB;
    double *rs = nullptr;
    MI_CHECK(hipMalloc(&rs, 2 * npad * sizeof(double)));
    double *rsA = rs + npad;
    hipStream_t s = c->s_main;
    k_zero<<<grid_for(2 * npad), 256, 0, s>>>(rs, 2 * npad);
    const int nsub = (NB + 63) / 64;
    const int64_t grid = (int64_t)MT * (MT + 1) / 2 * nsub * nsub;
    MI_LAUNCH(c, s, "check_llt", (double)N * N * N / 3.0, 0.0,
              k_check_llt<<<(unsigned)grid, 256, 0, s>>>(dL, dA0, N, NB, MT, rs));
    const int64_t work = (int64_t)MT * (MT + 1) / 2 * NB;
    MI_LAUNCH(c, s, "sym_rowsum", 0.0, 0.0, k_sym_rowsum<<<grid_for(work), 256, 0, s>>>(dA0, N, NB, MT, rsA));
    int rc;
    if ((rc = dmax_of(c, rs, npad, Rnorm))) { (void)hipFree(rs); return rc; }
    if ((rc = dmax_of(c, rsA, npad, Anorm))) { (void)hipFree(rs); return rc; }
    MI_CHECK(hipFree(rs));
    return mi_synchronize(c);
}

int mi_dcheck_axmb(mi_ctx_t *c, int N, int NB, const double *dA0, int NRHS, const double *db, int ldb,
                   const double *dx, int ldx, double *Anorm, double *Bnorm, double *Xnorm, double *Rnorm)
{
    if (N <= 0) { *Anorm = *Bnorm = *Xnorm = *Rnorm = 0.0; return 0; }
    const int MT = (N + NB - 1) / NB;
    const int64_t npad = (int64_t)MT * NB;
    double *buf = nullptr;
    // r (npad x NRHS, ld npad) + row sums
    MI_CHECK(hipMalloc(&buf, (npad * NRHS + npad) * sizeof(double)));
    double *r = buf, *rs = buf + npad * NRHS;
    hipStream_t s = c->s_main;
    k_zero<<<grid_for(npad * NRHS + npad), 256, 0, s>>>(buf, npad * NRHS + npad);
    int rc;
    if ((rc = dsymm_resid(c, s, N, NB, dA0, NRHS, dx, ldx, db, ldb, r, (int)npad, rs))) goto out;
    if ((rc = dmax_of(c, rs, N, Anorm))) goto out;       // row sums of |A0| from the residual pass
    k_ge_rowsum<<<grid_for(N), 256, 0, s>>>(db, N, NRHS, ldb, rs);
    if ((rc = dmax_of(c, rs, N, Bnorm))) goto out;
    k_ge_rowsum<<<grid_for(N), 256, 0, s>>>(dx, N, NRHS, ldx, rs);
    if ((rc = dmax_of(c, rs, N, Xnorm))) goto out;
    k_ge_rowsum<<<grid_for(N), 256, 0, s>>>(r, N, NRHS, (int)npad, rs);
    if ((rc = dmax_of(c, rs, N, Rnorm))) goto out;
out:
    (void)hipFree(buf);
    if (rc) return rc;
    return mi_synchronize(c);
}

// Norms behind check_dpoinv (src/dplasma_zcheck.c:264-314), on the GPU:
// ||A0||_1, ||Ainv||_1 (both symmetric, packed lower tiles) and
// ||I - Ainv A0||_1, with Ainv A0 formed by the symmetric residual kernel.
int mi_dcheck_poinv(mi_ctx_t *c, int N, int NB, const double *dA0, const double *dAinv, double *Anorm,
                    double *Ainvnorm, double *Rnorm)
{
    if (N <= 0) { *Anorm = *Ainvnorm = *Rnorm = 0.0; return 0; }
    const int MT = (N + NB - 1) / NB;
    const int64_t npad = (int64_t)MT * NB;
    const int64_t nn = (int64_t)N * N;
    double *buf = nullptr;
    // A0 full (N x N), I (N x N), R (npad x N), one vector
    MI_CHECK(hipMalloc(&buf, (2 * nn + npad * N + npad) * sizeof(double)));
    double *F = buf, *I = buf + nn, *R = I + nn, *v = R + npad * N;
    hipStream_t s = c->s_main;
    int rc;
    k_sym_to_full<<<grid_for(nn), 256, 0, s>>>(dA0, N, NB, MT, F);
    k_zero<<<grid_for(nn + npad * N + npad), 256, 0, s>>>(I, nn + npad * N + npad);
    k_add_diag<<<grid_for(N), 256, 0, s>>>(I, N, N, 1.0);
    if ((rc = dsymm_resid(c, s, N, NB, dAinv, N, F, N, I, N, R, (int)npad, v))) goto out;
    if ((rc = dmax_of(c, v, N, Ainvnorm))) goto out;     // row sums of |Ainv| (symmetric: one-norm)
    k_zero<<<grid_for(npad), 256, 0, s>>>(v, npad);
    k_sym_rowsum<<<grid_for((int64_t)MT * (MT + 1) / 2 * NB), 256, 0, s>>>(dA0, N, NB, MT, v);
    if ((rc = dmax_of(c, v, N, Anorm))) goto out;
    k_ge_colsum<<<(unsigned)std::min(N, 65535), 256, 0, s>>>(R, N, N, npad, v);
    if ((rc = dmax_of(c, v, N, Rnorm))) goto out;
out:
    (void)hipFree(buf);
    if (rc) return rc;
    return mi_synchronize(c);
}

}  // extern "C"
