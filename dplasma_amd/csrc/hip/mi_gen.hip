// mi_gen.hip -- SPD test matrices of prescribed condition number for the
// mixed-precision solve: A = Q diag(lambda) Q^T with
//   Q       the orthogonal factor of the Householder QR of an N x N Gaussian
//           matrix (the product of its N reflectors, i.e. a Haar-distributed
//           orthogonal matrix), formed explicitly;
//   lambda  geometric from 1 down to 1/kappa: lambda_j = kappa^(-j/(N-1)).
// Gaussian entries: element t (column-major index) takes two 64-bit words of
// the splitmix64 sequence started at seed (words 2t and 2t+1, Steele, Lea &
// Flood 2014) and maps them by Box-Muller: u1 = (a>>11 + 1) 2^-53 in (0, 1],
// u2 = (b>>11) 2^-53, g = sqrt(-2 ln u1) cos(2 pi u2).
// The QR and the products run in rocSOLVER / rocBLAS (dgeqrf, dorgqr, dsyrk),
// opened with dlopen on first use: test-matrix generation, off the
// factor-and-solve path, which never loads them.  The reference's
// generator of this family (src/zlatms_wrapper.c) uses a linear spectrum and
// a uniform matrix; this one keeps the configuration's geometric spectrum.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {

__device__ __host__ inline uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// G[t] = element t0 + t of the Gaussian stream of seed
__global__ __launch_bounds__(256) void k_gauss(double *__restrict__ G, int64_t n, uint64_t seed, int64_t t0)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = t0 + i;
        // word w of the sequence started at seed is splitmix64(seed + w * golden)
        const uint64_t a = splitmix64(seed + (uint64_t)(2 * t) * 0x9E3779B97F4A7C15ULL);
        const uint64_t b = splitmix64(seed + (uint64_t)(2 * t + 1) * 0x9E3779B97F4A7C15ULL);
        const double u1 = (double)((a >> 11) + 1) * 0x1p-53;
        const double u2 = (double)(b >> 11) * 0x1p-53;
        G[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
    }
}

// W(:, j) = Q(:, j) * sqrt(lambda_j), lambda_j = kappa^(-j / (N - 1))
__global__ __launch_bounds__(256) void k_scale_cols(const double *__restrict__ Q, double *__restrict__ W, int N,
                                                   double log_kappa)
{
    const int64_t n = (int64_t)N * N;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / N);
        const double e = N > 1 ? -(double)j / (double)(N - 1) : 0.0;
        W[t] = Q[t] * exp(0.5 * e * log_kappa);
    }
}

// lower triangle of the column-major C (ldc) -> packed lower tiles; the
// diagonal tiles are filled symmetric, padding zero
__global__ __launch_bounds__(256) void k_lower_to_tiles(const double *__restrict__ C, int ldc, int N, int NB,
                                                       int MT, double *__restrict__ T)
{
    const int64_t ntiles = (int64_t)MT * (MT + 1) / 2;
    const int64_t n = ntiles * NB * NB;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t tile = t / ((int64_t)NB * NB);
        const int64_t e = t - tile * NB * NB;
        const int j = (int)(e / NB), i = (int)(e % NB);
        int tn = 0;
        while (tile >= MT - tn) { tile -= MT - tn; tn++; }
        const int tm = tn + (int)tile;
        const int64_t r = (int64_t)tm * NB + i, c = (int64_t)tn * NB + j;
        double v = 0.0;
        if (r < N && c < N) v = (r >= c) ? C[c * ldc + r] : C[r * ldc + c];
        T[t] = v;
    }
}

// A_i(j, j) += 1 for every matrix of a batch (column-major, lda = n_i)
__global__ __launch_bounds__(256) void k_batch_add_identity(double *__restrict__ A, const int *__restrict__ n,
                                                           const int64_t *__restrict__ off, int count)
{
    const int i = blockIdx.x;
    if (i >= count) return;
    const int ni = n[i];
    for (int j = threadIdx.x; j < ni; j += blockDim.x) A[off[i] + (int64_t)j * (ni + 1)] += 1.0;
}

inline unsigned grid_n(int64_t work)
{
    int64_t g = (work + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}

// rocBLAS / rocSOLVER entry points, resolved at run time
struct RocLib {
    bool tried = false, ok = false;
    decltype(&rocblas_create_handle) create = nullptr;
    decltype(&rocblas_destroy_handle) destroy = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocblas_dsyrk) dsyrk = nullptr;
    decltype(&rocsolver_dgeqrf) dgeqrf = nullptr;
    decltype(&rocsolver_dorgqr) dorgqr = nullptr;
};
RocLib g_roc;

bool roc_load()
{
    if (g_roc.tried) return g_roc.ok;
    g_roc.tried = true;
    void *b = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
    if (!b) b = dlopen("librocblas.so", RTLD_NOW | RTLD_LOCAL);
    void *s = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!s) s = dlopen("librocsolver.so", RTLD_NOW | RTLD_LOCAL);
    if (!b || !s) { set_error("rocBLAS / rocSOLVER not found (needed by the generator)"); return false; }
    g_roc.create = (decltype(g_roc.create))dlsym(b, "rocblas_create_handle");
    g_roc.destroy = (decltype(g_roc.destroy))dlsym(b, "rocblas_destroy_handle");
    g_roc.set_stream = (decltype(g_roc.set_stream))dlsym(b, "rocblas_set_stream");
    g_roc.dsyrk = (decltype(g_roc.dsyrk))dlsym(b, "rocblas_dsyrk");
    g_roc.dgeqrf = (decltype(g_roc.dgeqrf))dlsym(s, "rocsolver_dgeqrf");
    g_roc.dorgqr = (decltype(g_roc.dorgqr))dlsym(s, "rocsolver_dorgqr");
    g_roc.ok = g_roc.create && g_roc.destroy && g_roc.set_stream && g_roc.dsyrk && g_roc.dgeqrf && g_roc.dorgqr;
    if (!g_roc.ok) set_error("rocBLAS / rocSOLVER symbols missing");
    return g_roc.ok;
}

#define RB_CHECK(call)                                                              \
    do {                                                                            \
        rocblas_status st_ = (call);                                                \
        if (st_ != rocblas_status_success) {                                        \
            set_error("%s:%d %s: rocblas status %d", __FILE__, __LINE__, #call, (int)st_); \
            rc = -2000 - (int)st_;                                                  \
            goto out;                                                               \
        }                                                                           \
    } while (0)

}  // namespace

extern "C" {

int mi_dgauss(mi_ctx_t *c, int64_t n, double *dG, uint64_t seed)
{
    if (n <= 0) return 0;
    MI_LAUNCH(c, c->s_main, "gauss", 0.0, 8.0 * n, k_gauss<<<grid_n(n), 256, 0, c->s_main>>>(dG, n, seed, 0));
    return mi_synchronize(c);
}

int mi_dorth_haar(mi_ctx_t *c, int N, uint64_t seed, double *dQ)
{
    if (N <= 0) return 0;
    if ((int64_t)N * N > 0x7fffffffLL) { set_error("N=%d too large for 32-bit LAPACK indices", N); return -2; }
    if (!roc_load()) return -5;
    int rc = 0;
    rocblas_handle h = nullptr;
    double *tau = nullptr;
    MI_CHECK(hipMalloc(&tau, (size_t)N * sizeof(double)));
    k_gauss<<<grid_n((int64_t)N * N), 256, 0, c->s_main>>>(dQ, (int64_t)N * N, seed, 0);
    RB_CHECK(g_roc.create(&h));
    RB_CHECK(g_roc.set_stream(h, c->s_main));
    RB_CHECK(g_roc.dgeqrf(h, N, N, dQ, N, tau));
    RB_CHECK(g_roc.dorgqr(h, N, N, N, dQ, N, tau));
out:
    if (h) (void)g_roc.destroy(h);
    if (rc == 0) rc = mi_synchronize(c);
    (void)hipFree(tau);
    return rc;
}

int mi_dspd_from_orth_tiles(mi_ctx_t *c, int N, int NB, const double *dQ, double kappa, double *dT)
{
    if (N <= 0) return 0;
    if (NB <= 0 || !(kappa >= 1.0)) { set_error("bad arguments"); return -2; }
    if (!roc_load()) return -5;
    const int MT = (N + NB - 1) / NB;
    int rc = 0;
    rocblas_handle h = nullptr;
    double *W = nullptr, *C = nullptr;
    const double one = 1.0, zero = 0.0;
    MI_CHECK(hipMalloc(&W, (size_t)N * N * sizeof(double)));
    if (hipMalloc(&C, (size_t)N * N * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(W);
        set_error("out of device memory for the generator");
        return -3;
    }
    k_scale_cols<<<grid_n((int64_t)N * N), 256, 0, c->s_main>>>(dQ, W, N, std::log(kappa));
    RB_CHECK(g_roc.create(&h));
    RB_CHECK(g_roc.set_stream(h, c->s_main));
    RB_CHECK(g_roc.dsyrk(h, rocblas_fill_lower, rocblas_operation_none, N, N, &one, W, N, &zero, C, N));
    k_lower_to_tiles<<<grid_n((int64_t)MT * (MT + 1) / 2 * NB * NB), 256, 0, c->s_main>>>(C, N, N, NB, MT, dT);
out:
    if (h) (void)g_roc.destroy(h);
    if (rc == 0) rc = mi_synchronize(c);
    (void)hipFree(W);
    (void)hipFree(C);
    return rc;
}

// Batch of SPD matrices A_i = G_i^T G_i / n_i + I (lower triangle written),
// G_i n_i x n_i Gaussian: the elements of the Gaussian stream of seed taken
// in order, G_0 first (column-major each).  Layout as mi_batch_create.
int mi_dbatch_spd(mi_ctx_t *c, int count, const int *n, uint64_t seed, double *dA)
{
    if (count <= 0) return 0;
    if (!roc_load()) return -5;
    int rc = 0, nmax = 0;
    rocblas_handle h = nullptr;
    double *G = nullptr;
    int *dn = nullptr;
    int64_t *doff = nullptr;
    std::vector<int64_t> off(count);
    int64_t tot = 0;
    for (int i = 0; i < count; i++) {
        off[i] = tot;
        tot += (int64_t)n[i] * n[i];
        nmax = std::max(nmax, n[i]);
    }
    MI_CHECK(hipMalloc(&G, (size_t)nmax * nmax * sizeof(double)));
    if (hipMalloc(&dn, count * sizeof(int)) != hipSuccess || hipMalloc(&doff, count * sizeof(int64_t)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory for the batch generator");
        rc = -3;
        goto out;
    }
    if (hipMemcpy(dn, n, count * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(doff, off.data(), count * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("batch generator: copy failed");
        rc = -4;
        goto out;
    }
    RB_CHECK(g_roc.create(&h));
    RB_CHECK(g_roc.set_stream(h, c->s_main));
    for (int i = 0; i < count; i++) {
        const int64_t ne = (int64_t)n[i] * n[i];
        const double alpha = 1.0 / n[i], zero = 0.0;
        k_gauss<<<grid_n(ne), 256, 0, c->s_main>>>(G, ne, seed, off[i]);
        RB_CHECK(g_roc.dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, n[i], n[i], &alpha, G, n[i], &zero,
                             dA + off[i], n[i]));
    }
    k_batch_add_identity<<<count, 256, 0, c->s_main>>>(dA, dn, doff, count);
out:
    if (h) (void)g_roc.destroy(h);
    if (rc == 0) rc = mi_synchronize(c);
    (void)hipFree(G);
    (void)hipFree(dn);
    (void)hipFree(doff);
    return rc;
}

}  // extern "C"
