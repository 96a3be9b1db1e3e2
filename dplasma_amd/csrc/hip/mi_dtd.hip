// mi_dtd.hip -- task-insertion engine for the tile Cholesky (the execution
// side of DPLASMA's DTD front-end, src/dtd_wrappers/z{potrf,trsm,herk,gemm}.c
// as used by tests/testing_zpotrf_dtd.c:91-302).
//
// A caller inserts POTRF / TRSM / SYRK / GEMM tile tasks in a sequentially
// consistent order; each task is enqueued at once on one of four streams and
// made to wait, with events, only for the tasks it depends on through the
// tiles it reads and writes:
//   read  tile t  -> after the last writer of t              (RAW)
//   write tile t  -> after the last writer and every reader
//                    of t since that write                   (WAW, WAR)
// POTRF(k) also writes the inverses of its 32 x 32 diagonal blocks to a
// slot of its own, which TRSM(m, k) reads (tracked like a tile).  The tile
// kernels are the ones of the fused factorization (mi_potrf.hip), one tile
// per launch.  Tasks of one stream run in insertion order, so same-stream
// dependencies need no event.
#include <vector>
#include "mi_common.hpp"
#include "dplasma_mi.h"

using namespace mi;

namespace {
constexpr int DTD_STREAMS = 4;

struct Access {
    hipEvent_t ev = nullptr;
    hipStream_t st = nullptr;
};
struct TileState {
    Access writer;
    std::vector<Access> readers;       // since the last write
};
}  // namespace

struct mi_dtd_s {
    Ctx *c = nullptr;
    int N = 0, NB = 0, MT = 0;
    double *dT = nullptr;
    double *winv = nullptr;            // MT slots of task_winv_elems(NB)
    size_t wsz = 0;
    hipStream_t st[DTD_STREAMS] = {};
    int rr = 0;
    std::vector<TileState> tiles;      // packed lower tiles, then the MT inverse slots
    std::vector<hipEvent_t> events;    // every event recorded (destroyed at wait)
    long ntask = 0;
};

namespace {

int tile_id(const mi_dtd_s *d, int m, int n) { return (int)tile_index(d->MT, m, n); }
int winv_id(const mi_dtd_s *d, int k) { return d->MT * (d->MT + 1) / 2 + k; }

void wait_on(hipStream_t s, const Access &a, int &rc)
{
    if (rc || !a.ev || a.st == s) return;
    const hipError_t e = hipStreamWaitEvent(s, a.ev, 0);
    if (e != hipSuccess) {
        set_error("dtd: hipStreamWaitEvent: %s", hipGetErrorString(e));
        rc = -1000 - (int)e;
    }
}

// Enqueues one task: dependencies from the tiles it reads (r) and writes (w),
// the body, then the bookkeeping.
template <class F>
int run_task(mi_dtd_s *d, hipStream_t s, std::initializer_list<int> r, std::initializer_list<int> w, F body)
{
    int rc = 0;
    for (int t : r) wait_on(s, d->tiles[t].writer, rc);
    for (int t : w) {
        wait_on(s, d->tiles[t].writer, rc);
        for (const Access &a : d->tiles[t].readers) wait_on(s, a, rc);
    }
    if (rc) return rc;
    if ((rc = body(s))) return rc;
    hipEvent_t e = nullptr;
    MI_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    d->events.push_back(e);
    MI_CHECK(hipEventRecord(e, s));
    const Access a{e, s};
    for (int t : r) d->tiles[t].readers.push_back(a);
    for (int t : w) {
        d->tiles[t].writer = a;
        d->tiles[t].readers.clear();
    }
    d->ntask++;
    return 0;
}

hipStream_t next_stream(mi_dtd_s *d)
{
    // POTRF keeps stream 0 (the critical path); the other tasks rotate
    // over the remaining streams
    d->rr = (d->rr % (DTD_STREAMS - 1)) + 1;
    return d->st[d->rr];
}

int bad(const char *what)
{
    set_error("dtd: %s", what);
    return -2;
}

}  // namespace

extern "C" {

int mi_dtd_create(mi_ctx_t *c, int N, int NB, double *dT, mi_dtd_t **out)
{
    *out = nullptr;
    if (N <= 0 || NB <= 0 || !dT) return bad("bad sizes or image");
    int rc = task_prepare_fp64(NB);
    if (rc) return rc;
    mi_dtd_s *d = new mi_dtd_s();
    d->c = c;
    d->N = N;
    d->NB = NB;
    d->MT = (N + NB - 1) / NB;
    d->dT = dT;
    d->wsz = task_winv_elems(NB);
    d->tiles.resize((size_t)d->MT * (d->MT + 1) / 2 + d->MT);
    if (hipMalloc(&d->winv, d->wsz * d->MT * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        delete d;
        set_error("dtd: out of device memory for the block inverses");
        return -3;
    }
    for (auto &s : d->st) {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            mi_dtd_destroy(d);
            set_error("dtd: stream creation failed");
            return -1;
        }
    }
    // the context's info word starts at 0 for this task pool
    const int zero = 0;
    if (hipMemcpyAsync(c->d_info, &zero, sizeof(int), hipMemcpyHostToDevice, d->st[0]) != hipSuccess ||
        hipStreamSynchronize(d->st[0]) != hipSuccess) {
        mi_dtd_destroy(d);
        set_error("dtd: info reset failed");
        return -1;
    }
    *out = d;
    return 0;
}

int mi_dtd_potrf(mi_dtd_t *d, int k)
{
    if (k < 0 || k >= d->MT) return bad("POTRF tile out of range");
    double *w = d->winv + (size_t)k * d->wsz;
    return run_task(d, d->st[0], {}, {tile_id(d, k, k), winv_id(d, k)}, [&](hipStream_t s) {
        return task_potrf_fp64(d->c, s, d->dT, d->N, d->NB, k, w);
    });
}

int mi_dtd_trsm(mi_dtd_t *d, int m, int k)
{
    if (k < 0 || m <= k || m >= d->MT) return bad("TRSM tiles out of range");
    const double *w = d->winv + (size_t)k * d->wsz;
    return run_task(d, next_stream(d), {tile_id(d, k, k), winv_id(d, k)}, {tile_id(d, m, k)},
                    [&](hipStream_t s) { return task_trsm_fp64(d->c, s, d->dT, d->N, d->NB, m, k, w); });
}

int mi_dtd_syrk(mi_dtd_t *d, int k, int m)
{
    if (k < 0 || m <= k || m >= d->MT) return bad("SYRK tiles out of range");
    return run_task(d, next_stream(d), {tile_id(d, m, k)}, {tile_id(d, m, m)},
                    [&](hipStream_t s) { return task_update_fp64(d->c, s, d->dT, d->N, d->NB, m, m, k); });
}

int mi_dtd_gemm(mi_dtd_t *d, int m, int n, int k)
{
    if (k < 0 || n <= k || m <= n || m >= d->MT) return bad("GEMM tiles out of range");
    return run_task(d, next_stream(d), {tile_id(d, m, k), tile_id(d, n, k)}, {tile_id(d, m, n)},
                    [&](hipStream_t s) { return task_update_fp64(d->c, s, d->dT, d->N, d->NB, m, n, k); });
}

int mi_dtd_wait(mi_dtd_t *d, int *info, long *ntasks)
{
    for (auto s : d->st) MI_CHECK(hipStreamSynchronize(s));
    for (auto e : d->events) (void)hipEventDestroy(e);
    d->events.clear();
    for (auto &t : d->tiles) {
        t.writer = Access();
        t.readers.clear();
    }
    if (ntasks) *ntasks = d->ntask;
    if (info) MI_CHECK(hipMemcpy(info, d->c->d_info, sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

void mi_dtd_destroy(mi_dtd_t *d)
{
    if (!d) return;
    for (auto s : d->st)
        if (s) (void)hipStreamSynchronize(s);
    for (auto e : d->events) (void)hipEventDestroy(e);
    for (auto s : d->st)
        if (s) (void)hipStreamDestroy(s);
    if (d->winv) (void)hipFree(d->winv);
    delete d;
}

}  // extern "C"
