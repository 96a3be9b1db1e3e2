// mi_common.hpp -- shared definitions of the gfx950 kernel library.
//
// Device layout of a symmetric matrix (the "packed lower tile" layout):
// tile (m,n), m >= n, of an MT x MT tile grid lives at tile index
// n*MT - n*(n-1)/2 + (m-n), each tile NB*NB elements, column-major, ld = NB.
// It is the host layout of DPLASMA's lower symmetric block-cyclic descriptor
// on one rank (tiles of mb*nb with lda = mb, src/include/dplasma/types_lapack.h:112),
// so a lower descriptor stages in with one contiguous copy.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace mi {

__host__ __device__ inline int64_t tile_index(int MT, int m, int n)
{
    return (int64_t)n * MT - ((int64_t)n * (n - 1)) / 2 + (m - n);
}

__host__ __device__ inline int tile_rows(int N, int NB, int MT, int m)
{
    return (m == MT - 1) ? N - m * NB : NB;
}

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

// Element-type traits of the 16x16x4 MFMA shapes used here: FP64
// (v_mfma_f64_16x16x4_f64) and FP32 (v_mfma_f32_16x16x4_f32).  Both take one
// A and one B element per lane (A(i, k) and B(k, j) in lane 16k + i, resp.
// 16k + j) and return four elements per lane; row(q, l/16) is the output row
// of element q (column l % 16), as probed on gfx950.
template <class T> struct MF;
template <> struct MF<double> {
    typedef dx4 v4;
    typedef double2 v2;
    __device__ static inline v4 mfma(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int q, int fk) { return 4 * q + fk; }
};
template <> struct MF<float> {
    typedef fx4 v4;
    typedef float2 v2;
    __device__ static inline v4 mfma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int q, int fk) { return 4 * fk + q; }
};

// Wave-uniform copy of a value held by one lane (v_readlane_b32; two for a double).
__device__ inline double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double readlane_t(double v, int lane) { return readlane_d(v, lane); }
__device__ inline float readlane_t(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

struct ProfRec {
    const char *name;
    hipEvent_t ev0, ev1;
    double flops, bytes;
};

struct ProfAgg {
    std::string name;
    int launches = 0;
    double ms = 0.0, flops = 0.0, bytes = 0.0;
    // [start, end] of every launch, ms after the start event (t0) of the
    // blocking call it belonged to, and that call's sequence number
    std::vector<double> t_start, t_end;
    std::vector<int> call;
};

// Device workspaces owned by a context: grown on demand (ws_get), freed with
// the context, so that calls on different contexts never share memory.
enum WsId {
    WS_POTRF_WINV = 0,  // diagonal-block inverses handed from POTRF(k) to TRSM(k)
    WS_TRSM,            // solve: 32x32 and 128x128 block inverses, link matrices
    WS_MX_L32,          // mixed solve: FP32 tiles
    WS_MX_R32,          //   FP32 right-hand sides / corrections
    WS_MX_R64,          //   FP64 residual, norms, row sums
    WS_MX_W32,          //   FP32 solve workspace
    WS_MX_PART,         //   per-tile partials of the symmetric residual
    WS_MX_AUX,          //   FP64 vectors of the preconditioned iteration
    WS_ADJ_W,           // adjoint: inverses and tile scratch
    WS_ADJ_PART,        //   K-split partials
    WS_ADJ_JOBS,        //   job table
    WS_ADJ_SEGS,        //   segment table
    WS_POTRI,           // inverse: the seed / result image
    WS_DTD,             // task-insertion front-end: per-task diagonal-block inverses
    WS_POTRF_REC,       // recursive diagonal factor: packed image of one tile's sub-tiles
    WS_BF16_PANEL,      // BF16-split FP32 factor: hi/lo images of one panel column
    WS_COUNT
};
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t s_main = nullptr;   // trailing updates, solves, copies
    hipStream_t s_panel = nullptr;  // diagonal factorization + panel solve (look-ahead)
    std::vector<hipEvent_t> ev_pool;   // cross-stream dependencies
    size_t ev_next = 0;
    std::vector<hipEvent_t> ev_ring;   // fixed ring (created with the context)
    int *d_info = nullptr;          // device-side first-failure info
    double *d_scratch = nullptr;    // small reductions (norms)
    size_t scratch_bytes = 0;
    DevBuf ws[WS_COUNT];
    // hand-off memory of the wavefront solve (uncached; an uncached buffer
    // that a kernel has read is never freed, see ws_uncached)
    DevBuf wv_ticket, wv_stage;
    bool prof = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> prof_free;
    std::vector<ProfAgg> agg;
    hipEvent_t t0 = nullptr, t1 = nullptr;  // span of the last blocking call
    double *h_stat = nullptr;       // page-locked host copy of small device results
    size_t h_stat_n = 0;            //   (doubles), so that their copy is asynchronous
    double last_ms = 0.0;
    int prof_calls = 0;             // blocking calls whose launches were collected
};

void set_error(const char *fmt, ...);
hipEvent_t next_event(Ctx *c);
// Workspace `id` of the context with at least `bytes` bytes (nullptr and an
// error message when the device is out of memory).  Growing waits for the
// context's streams first, so no queued kernel still uses the old buffer.
void *ws_get(Ctx *c, int id, size_t bytes);
// Uncached buffer of at least `bytes` bytes in `b` (0 or a negative code); a
// grown buffer replaces the old one, which stays allocated until the process
// ends (uncached memory is not freed once a kernel has read it).
int ws_uncached(Ctx *c, DevBuf &b, size_t bytes);
void prof_begin(Ctx *c, hipStream_t s, const char *name, double flops, double bytes, ProfRec *r);
void prof_end(Ctx *c, hipStream_t s, ProfRec *r);

// Tile triangular solves (mi_solve.hip), enqueue only: trsm_prepare writes
// the 32x32 diagonal-block inverses of L (trsm_ws_elems(N, NB) elements),
// trsm_enqueue applies op(L)^{-1} (trans 111 / 112) to NRHS columns.
size_t trsm_ws_elems(int N, int NB);
template <class T> int trsm_prepare(Ctx *c, hipStream_t s, int unit, int N, int NB, const T *dT, T *W);
// first_reset_done: the hand-off memory of the first chunk of 16 columns
// was already reset on s by trsm_reset (after the last sweep that used it).
template <class T>
int trsm_enqueue(Ctx *c, hipStream_t s, int trans, int N, int NB, const T *dT, const T *W, int NRHS, T *dB,
                 int ldb, bool first_reset_done = false);
template <class T> int trsm_reset(Ctx *c, hipStream_t s, int N, int NB);
// Before the first trsm_enqueue of a blocking call: clears the abort word of
// the wavefront solve on s; after synchronizing: -5 if a hand-off wait gave
// up (results invalid), else 0.
int trsm_begin(Ctx *c, hipStream_t s);
int trsm_status(Ctx *c);
// the abort word itself (device memory, valid after trsm_begin), and the
// error that trsm_status reports when it is set
const int *trsm_abort_word(Ctx *c);
int trsm_abort_error();
// FP32 tile factorization (mi_spotrf_tiles) with the trailing-update engine
// given per call: bf16_split = BF16 MFMA on the two-term split.
// wait = false: enqueue only (both streams joined into s_main); *info is
// then 0 and the factorization's info is left in d_info[0].
int spotrf_tiles_split(Ctx *c, int N, int NB, float *dT, int *info, bool bf16_split, bool wait = true);
// Reads, once per device and with that device current, the device address of
// the fixed scratch of mi_dpotrf_tiles_noalloc (mi_potrf.hip).
void noalloc_setup(int device);
// Single-tile tasks of the tile Cholesky on the packed lower FP64 image,
// enqueued on s (mi_potrf.hip): POTRF(k) writes the inverses of its 32 x 32
// diagonal blocks to Winv (task_winv_elems(NB) elements), which TRSM(m, k)
// reads; update: A(m, n) -= A(m, k) A(n, k)^T, SYRK when m == n.
int task_prepare_fp64(int NB);
int task_potrf_fp64(Ctx *c, hipStream_t s, double *dT, int N, int NB, int k, double *Winv);
int task_trsm_fp64(Ctx *c, hipStream_t s, double *dT, int N, int NB, int m, int k, const double *Winv);
int task_update_fp64(Ctx *c, hipStream_t s, double *dT, int N, int NB, int m, int n, int k);
size_t task_winv_elems(int NB);
// R = B - A X for the symmetric A in lower tiles, FP64, enqueued on s
// (mi_mixed.hip); rowabs (N values, may be null) receives the row sums of |A|.
int dsymm_resid(Ctx *c, hipStream_t s, int N, int NB, const double *dA, int NRHS, const double *dX, int ldx,
                const double *dB, int ldb, double *dR, int ldr, double *rowabs);

}  // namespace mi

// the opaque handle of the C ABI is the context itself
struct mi_ctx_s : public mi::Ctx {};

// A failed call also resets HIP's per-thread last error, so that a
// recoverable failure (out of memory) is not reported again by the next
// launch's check; a sticky device error is reported again anyway.
#define MI_CHECK(call)                                                            \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            (void)hipGetLastError();                                              \
            mi::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                          hipGetErrorString(e_));                                 \
            return -1000 - (int)e_;                                               \
        }                                                                         \
    } while (0)

// Launch with optional event bracketing for the profiler.
// The last error is cleared right before the launch, so that the check
// after it reports this launch and not an earlier ignored failure.
#define MI_LAUNCH(ctx, stream, name, flops, bytes, ...)                           \
    do {                                                                          \
        mi::ProfRec r_{};                                                         \
        if ((ctx)->prof) mi::prof_begin((ctx), (stream), (name), (flops), (bytes), &r_); \
        (void)hipGetLastError();                                                  \
        __VA_ARGS__;                                                              \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            mi::set_error("launch %s: %s", (name), hipGetErrorString(e_));        \
            return -1000 - (int)e_;                                               \
        }                                                                         \
        if ((ctx)->prof) mi::prof_end((ctx), (stream), &r_);                      \
    } while (0)
