// mi_runtime.hip -- context, memory, copies and the event profiler of the
// gfx950 kernel library (C ABI in include/dplasma_mi.h).
#include <cstdarg>
#include <algorithm>
#include <cstring>
#include <mutex>
#include "mi_common.hpp"
#include "dplasma_mi.h"

namespace mi {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

void *ws_get(Ctx *c, int id, size_t bytes)
{
    DevBuf &b = c->ws[id];
    if (b.p && b.bytes >= bytes) return b.p;
    if (b.p) {
        // the old buffer may still be read by work queued on this context
        if (hipStreamSynchronize(c->s_panel) != hipSuccess || hipStreamSynchronize(c->s_main) != hipSuccess) {
            set_error("workspace %d: stream synchronization failed", id);
            return nullptr;
        }
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory: workspace %d of %zu bytes (%s)", id, bytes, hipGetErrorString(e));
        return nullptr;
    }
    b.p = p;
    b.bytes = bytes;
    return p;
}

// Uncached buffers replaced by larger ones or left by destroyed contexts:
// kept allocated until the process ends.
static std::mutex g_retired_mu;
static std::vector<void *> g_retired;

int ws_uncached(Ctx *c, DevBuf &b, size_t bytes)
{
    (void)c;
    if (b.p && b.bytes >= bytes) return 0;
    void *p = nullptr;
    MI_CHECK(hipExtMallocWithFlags(&p, bytes, hipDeviceMallocUncached));
    if (b.p) {
        std::lock_guard<std::mutex> g(g_retired_mu);
        g_retired.push_back(b.p);
    }
    b.p = p;
    b.bytes = bytes;
    return 0;
}

hipEvent_t next_event(Ctx *c)
{
    if (c->ev_next >= c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_next++];
}

static hipEvent_t prof_event(Ctx *c)
{
    if (!c->prof_free.empty()) {
        hipEvent_t e = c->prof_free.back();
        c->prof_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void prof_begin(Ctx *c, hipStream_t s, const char *name, double flops, double bytes, ProfRec *r)
{
    r->name = name;
    r->flops = flops;
    r->bytes = bytes;
    r->ev0 = prof_event(c);
    r->ev1 = prof_event(c);
    (void)hipEventRecord(r->ev0, s);
}

void prof_end(Ctx *c, hipStream_t s, ProfRec *r)
{
    (void)hipEventRecord(r->ev1, s);
    c->recs.push_back(*r);
}

// Resolve recorded launches into per-name totals (after a synchronize).
static void prof_collect(Ctx *c)
{
    if (!c->recs.empty()) c->prof_calls++;
    for (auto &r : c->recs) {
        float ms = 0.f, t_a = 0.f, t_b = 0.f;
        if (hipEventElapsedTime(&ms, r.ev0, r.ev1) != hipSuccess) ms = 0.f;
        // launch interval relative to the call's start event (t0 is recorded
        // first on the main stream by every blocking hot-path call; the
        // single-step entries do not record it, and then the interval is
        // the launch's own)
        if (hipEventElapsedTime(&t_a, c->t0, r.ev0) != hipSuccess) t_a = 0.f;
        if (hipEventElapsedTime(&t_b, c->t0, r.ev1) != hipSuccess) t_b = t_a + ms;
        (void)hipGetLastError();                   // failures above are handled
        ProfAgg *a = nullptr;
        for (auto &x : c->agg)
            if (x.name == r.name) { a = &x; break; }
        if (!a) {
            c->agg.push_back(ProfAgg());
            a = &c->agg.back();
            a->name = r.name;
        }
        a->launches++;
        a->ms += ms;
        a->t_start.push_back(t_a);
        a->t_end.push_back(t_b);
        a->call.push_back(c->prof_calls);
        a->flops += r.flops;
        a->bytes += r.bytes;
        c->prof_free.push_back(r.ev0);
        c->prof_free.push_back(r.ev1);
    }
    c->recs.clear();
}

}  // namespace mi

using namespace mi;

extern "C" {

const char *mi_last_error(void) { return mi::g_err; }

int mi_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; set_error("hipGetDeviceCount: %s", hipGetErrorString(e)); return -1; }
    *count = n;
    return 0;
}

int mi_ctx_create(int device, mi_ctx_t **out)
{
    *out = nullptr;
    int n = 0;
    if (mi_device_count(&n) != 0 || n <= 0) { set_error("no HIP device visible"); return -1; }
    if (device < 0 || device >= n) { set_error("device %d out of range (%d devices)", device, n); return -2; }
    MI_CHECK(hipSetDevice(device));
    mi_ctx_t *c = new mi_ctx_t();
    c->device = device;
    // default priority on both streams: the look-ahead relies on ordinary
    // queue arbitration, not on stream priorities
    MI_CHECK(hipStreamCreateWithFlags(&c->s_main, hipStreamNonBlocking));
    MI_CHECK(hipStreamCreateWithFlags(&c->s_panel, hipStreamNonBlocking));
    MI_CHECK(hipMalloc(&c->d_info, 64 * sizeof(int)));
    c->scratch_bytes = 1 << 20;
    MI_CHECK(hipMalloc(&c->d_scratch, c->scratch_bytes));
    MI_CHECK(hipEventCreate(&c->t0));
    MI_CHECK(hipEventCreate(&c->t1));
    for (int i = 0; i < 16; i++) {
        hipEvent_t e;
        MI_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_ring.push_back(e);
    }
    mi::noalloc_setup(device);
    *out = c;
    return 0;
}

void mi_ctx_destroy(mi_ctx_t *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->s_main);
    (void)hipStreamSynchronize(c->s_panel);
    mi::prof_collect(c);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto e : c->prof_free) (void)hipEventDestroy(e);
    for (auto e : c->ev_ring) (void)hipEventDestroy(e);
    (void)hipEventDestroy(c->t0);
    (void)hipEventDestroy(c->t1);
    (void)hipFree(c->d_info);
    (void)hipFree(c->d_scratch);
    if (c->h_stat) (void)hipHostFree(c->h_stat);
    for (auto &b : c->ws)
        if (b.p) (void)hipFree(b.p);
    {
        std::lock_guard<std::mutex> g(mi::g_retired_mu);
        if (c->wv_ticket.p) mi::g_retired.push_back(c->wv_ticket.p);
        if (c->wv_stage.p) mi::g_retired.push_back(c->wv_stage.p);
    }
    (void)hipStreamDestroy(c->s_main);
    (void)hipStreamDestroy(c->s_panel);
    delete c;
}

int mi_synchronize(mi_ctx_t *c)
{
    MI_CHECK(hipStreamSynchronize(c->s_panel));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    c->ev_next = 0;
    if (c->prof) mi::prof_collect(c);
    return 0;
}

int mi_malloc(mi_ctx_t *c, void **p, size_t bytes)
{
    (void)c;
    *p = nullptr;
    if (bytes == 0) return 0;
    MI_CHECK(hipMalloc(p, bytes));
    return 0;
}

int mi_free(mi_ctx_t *c, void *p)
{
    (void)c;
    if (p) MI_CHECK(hipFree(p));
    return 0;
}

int mi_host_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return 0;
    MI_CHECK(hipHostMalloc(p, bytes, hipHostMallocDefault));
    return 0;
}

int mi_host_free(void *p)
{
    if (p) MI_CHECK(hipHostFree(p));
    return 0;
}

int mi_memcpy_h2d(mi_ctx_t *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    MI_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

int mi_memcpy_d2h(mi_ctx_t *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    MI_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

int mi_memcpy_d2d(mi_ctx_t *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return 0;
    MI_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

int mi_memcpy2d_h2d(mi_ctx_t *c, void *dst, size_t dpitch, const void *src, size_t spitch,
                    size_t width, size_t height)
{
    if (!width || !height) return 0;
    MI_CHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyHostToDevice,
                              c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

int mi_memcpy2d_d2h(mi_ctx_t *c, void *dst, size_t dpitch, const void *src, size_t spitch,
                    size_t width, size_t height)
{
    if (!width || !height) return 0;
    MI_CHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost,
                              c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

int mi_memset(mi_ctx_t *c, void *p, int v, size_t bytes)
{
    if (!bytes) return 0;
    MI_CHECK(hipMemsetAsync(p, v, bytes, c->s_main));
    MI_CHECK(hipStreamSynchronize(c->s_main));
    return 0;
}

int mi_prof_enable(mi_ctx_t *c, int on)
{
    c->prof = (on != 0);
    return 0;
}

int mi_prof_reset(mi_ctx_t *c)
{
    mi_synchronize(c);
    c->agg.clear();
    c->prof_calls = 0;
    return 0;
}

int mi_prof_busy_ms(mi_ctx_t *c, const char *kernel, double *busy_ms)
{
    mi::prof_collect(c);
    *busy_ms = 0.0;
    // union of the launch intervals of the kernel (all kernels when kernel
    // is NULL), per blocking call (intervals of different calls share no
    // time origin), summed over the calls
    std::vector<std::pair<std::pair<int, double>, double>> iv;
    // a kernel name ending in '*' selects every name with that prefix
    const size_t kl = kernel ? strlen(kernel) : 0;
    const bool prefix = kl > 0 && kernel[kl - 1] == '*';
    for (auto &a : c->agg) {
        if (kernel && !(prefix ? a.name.compare(0, kl - 1, kernel, kl - 1) == 0 : a.name == kernel)) continue;
        for (size_t i = 0; i < a.t_start.size(); i++) iv.push_back({{a.call[i], a.t_start[i]}, a.t_end[i]});
    }
    if (iv.empty()) return kernel ? -1 : 0;
    std::sort(iv.begin(), iv.end());
    double tot = 0.0;
    int cur_call = iv[0].first.first;
    double s0 = iv[0].first.second, e0 = iv[0].second;
    for (size_t i = 1; i < iv.size(); i++) {
        const int call = iv[i].first.first;
        const double s = iv[i].first.second, e = iv[i].second;
        if (call == cur_call && s <= e0) {
            if (e > e0) e0 = e;
            continue;
        }
        tot += e0 - s0;
        cur_call = call;
        s0 = s;
        e0 = e;
    }
    tot += e0 - s0;
    *busy_ms = tot;
    return 0;
}

int mi_prof_get(mi_ctx_t *c, const char *kernel, int *launches, double *ms, double *flops,
                double *bytes)
{
    mi::prof_collect(c);
    for (auto &a : c->agg)
        if (a.name == kernel) {
            if (launches) *launches = a.launches;
            if (ms) *ms = a.ms;
            if (flops) *flops = a.flops;
            if (bytes) *bytes = a.bytes;
            return 0;
        }
    if (launches) *launches = 0;
    if (ms) *ms = 0;
    if (flops) *flops = 0;
    if (bytes) *bytes = 0;
    return -1;
}

int mi_prof_count(mi_ctx_t *c)
{
    mi::prof_collect(c);
    return (int)c->agg.size();
}

const char *mi_prof_name(mi_ctx_t *c, int i)
{
    if (i < 0 || i >= (int)c->agg.size()) return nullptr;
    return c->agg[i].name.c_str();
}

int mi_last_span_ms(mi_ctx_t *c, double *ms)
{
    *ms = c->last_ms;
    return 0;
}

}  // extern "C"
