// CPU stand-in for the HIP runtime calls libfvm2d_gpu's host code makes, for
// an AddressSanitizer build of that host code (no GPU): device memory is host
// memory, kernels do nothing, asynchronous copies, memsets and stream-ordered
// frees are queued per stream and run only when the stream is synchronised
// (so a host buffer that dies before its copy runs is reported), captured
// streams record into graphs that launches replay.  Misuse the runtime would
// not report at once (a synchronise or launch on a capturing stream, a
// destroyed stream, event or graph used, a free of memory it did not
// allocate) aborts with a message.
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <vector>

namespace {
constexpr unsigned kStreamMagic = 0x5354524d, kEventMagic = 0x45564e54, kGraphMagic = 0x47524150,
                   kExecMagic = 0x45584543;
struct Op {
    std::function<void()> f;
    const char *what;
};
std::recursive_mutex mu;
int g_query_flip = 0;
long g_launches = 0, g_graph_launches = 0, g_copies = 0;
void die(const char *m, const void *p)
{
    fprintf(stderr, "hipstub: %s %p\n", m, p);
    abort();
}
}  // namespace

struct ihipGraph {
    unsigned magic = kGraphMagic;
    std::vector<Op> ops;
};
struct hipGraphExec {
    unsigned magic = kExecMagic;
    std::vector<Op> ops;
};
struct ihipStream_t {
    unsigned magic = kStreamMagic;
    std::deque<Op> q;
    ihipGraph *cap = nullptr;
    unsigned flags = 0;
};
struct ihipEvent_t {
    unsigned magic = kEventMagic;
    ihipStream_t *last = nullptr;
    bool recorded = false;
};
struct ihipMemPoolHandle_t {
    int x;
};

namespace {
ihipStream_t g_null;                        // the null stream
std::set<ihipStream_t *> g_streams;
std::set<void *> g_dev;                      // live device allocations
ihipMemPoolHandle_t g_pool;

ihipStream_t *S(hipStream_t s)
{
    if (!s) return &g_null;
    if (s->magic != kStreamMagic) die("bad stream", s);
    if (!g_streams.count(s)) die("stream not live", s);
    return s;
}
void flush(ihipStream_t *s)
{
    if (s->cap) die("synchronising a capturing stream", s);
    while (!s->q.empty()) {
        Op o = std::move(s->q.front());
        s->q.pop_front();
        o.f();
    }
}
void flush_all()
{
    for (auto *s : g_streams)
        if (!s->cap) flush(s);
    flush(&g_null);
}
void enqueue(hipStream_t st, std::function<void()> f, const char *what)
{
    ihipStream_t *s = S(st);
    if (s->cap) {
        s->cap->ops.push_back({std::move(f), what});
        return;
    }
    if (s == &g_null) {             // legacy null stream: after every blocking stream's work
        for (auto *o : g_streams)
            if (!(o->flags & hipStreamNonBlocking) && !o->cap) flush(o);
        f();
        return;
    }
    s->q.push_back({std::move(f), what});
}
}  // namespace

extern "C" {

struct StubDim3 {
    unsigned x, y, z;
};
thread_local StubDim3 t_grid, t_block;
thread_local size_t t_shm;
thread_local hipStream_t t_stream;

hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t shm, hipStream_t st)
{
    t_grid = {g.x, g.y, g.z};
    t_block = {b.x, b.y, b.z};
    t_shm = shm;
    t_stream = st;
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3 *g, dim3 *b, size_t *shm, hipStream_t *st)
{
    *g = dim3(t_grid.x, t_grid.y, t_grid.z);
    *b = dim3(t_block.x, t_block.y, t_block.z);
    *shm = t_shm;
    *st = t_stream;
    return hipSuccess;
}
void **__hipRegisterFatBinary(const void *)
{
    static void *h[4];
    return h;
}
void __hipUnregisterFatBinary(void **) {}
void __hipRegisterFunction(void **, const void *, char *, const char *, unsigned int, void *, void *,
                           void *, void *, int *)
{
}
void __hipRegisterVar(void **, void *, char *, char *, int, size_t, int, int) {}

hipError_t hipLaunchKernel(const void *f, dim3 g, dim3 b, void **args, size_t shm, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!f) die("null kernel", f);
    if (g.x == 0 || g.y == 0 || g.z == 0 || b.x == 0 || b.x * b.y * b.z > 1024) die("bad launch shape", f);
    if (shm > 160 * 1024) die("too much LDS", f);
    (void)args;
    ++g_launches;
    enqueue(st, [] {}, "kernel");
    return hipSuccess;
}

hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
    switch (a) {
    case hipDeviceAttributeMultiprocessorCount: *v = 256; break;
    case hipDeviceAttributeMaxSharedMemoryPerBlock: *v = 160 * 1024; break;
    case hipDeviceAttributeMaxSharedMemoryPerMultiprocessor: *v = 160 * 1024; break;
    default: *v = 0; break;
    }
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "hipstub error"; }

hipError_t hipMalloc(void **p, size_t n)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    *p = calloc(1, n ? n : 1);
    g_dev.insert(*p);
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!p) return hipSuccess;
    flush_all();                    // hipFree waits for the device
    if (!g_dev.erase(p)) die("hipFree of a pointer not allocated by hipMalloc", p);
    free(p);
    return hipSuccess;
}
hipError_t hipMallocAsync(void **p, size_t n, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (S(st)->cap) die("hipMallocAsync in capture", st);
    *p = calloc(1, n ? n : 1);
    g_dev.insert(*p);
    return hipSuccess;
}
hipError_t hipFreeAsync(void *p, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!p) return hipSuccess;
    if (S(st)->cap) die("hipFreeAsync in capture", st);
    if (!g_dev.count(p)) die("hipFreeAsync of a pointer not allocated", p);
    enqueue(st, [p] {
        if (!g_dev.erase(p)) die("double stream-ordered free", p);
        free(p);
    }, "free");
    return hipSuccess;
}
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t *m, int)
{
    *m = &g_pool;
    return hipSuccess;
}
hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void *) { return hipSuccess; }
hipError_t hipHostMalloc(void **p, size_t n, unsigned int)
{
    *p = calloc(1, n ? n : 1);
    return hipSuccess;
}
hipError_t hipHostFree(void *p)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    flush_all();
    free(p);
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **d, const void *sym)
{
    *d = const_cast<void *>(sym);
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    ++g_copies;
    enqueue(st, [dst, src, n] { memmove(dst, src, n); }, "copy");
    return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind)
{
    // the null stream: after the blocking streams' work only
    std::lock_guard<std::recursive_mutex> l(mu);
    ++g_copies;
    for (auto *o : g_streams)
        if (!(o->flags & hipStreamNonBlocking) && !o->cap) flush(o);
    flush(&g_null);
    memmove(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemset(void *p, int v, size_t n)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    flush_all();
    memset(p, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    enqueue(st, [p, v, n] { memset(p, v, n); }, "memset");
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int flags)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    auto *x = new ihipStream_t;
    x->flags = flags;
    g_streams.insert(x);
    *s = x;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    ihipStream_t *s = S(st);
    if (s == &g_null) die("destroying the null stream", st);
    if (s->cap) die("destroying a capturing stream", st);
    flush(s);
    g_streams.erase(s);
    s->magic = 0;
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    ihipStream_t *s = S(st);
    if (s == &g_null) flush_all();
    else flush(s);
    return hipSuccess;
}
hipError_t hipStreamQuery(hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    ihipStream_t *s = S(st);
    if (s->cap) die("query of a capturing stream", st);
    // every other query reports the work as still running
    if (!s->q.empty() && (++g_query_flip & 1)) return hipErrorNotReady;
    flush(s);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    ihipStream_t *s = S(st);
    if (s == &g_null) return hipErrorInvalidValue;
    if (s->cap) return hipErrorIllegalState;
    s->cap = new ihipGraph;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t *g)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    ihipStream_t *s = S(st);
    if (!s->cap) return hipErrorIllegalState;
    *g = s->cap;
    s->cap = nullptr;
    return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t st, hipStreamCaptureStatus *c)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    *c = S(st)->cap ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!g || g->magic != kGraphMagic) die("bad graph", g);
    auto *x = new hipGraphExec;
    x->ops = g->ops;
    *e = x;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!e || e->magic != kExecMagic) die("bad graph exec", e);
    ihipStream_t *s = S(st);
    if (s->cap) die("graph launch into a capturing stream", st);
    ++g_graph_launches;
    for (auto &o : e->ops) enqueue(st, o.f, o.what);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!e || e->magic != kExecMagic) die("bad graph exec destroy", e);
    e->magic = 0;
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!g || g->magic != kGraphMagic) die("bad graph destroy", g);
    g->magic = 0;
    delete g;
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new ihipEvent_t;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    if (!e || e->magic != kEventMagic) die("bad event destroy", e);
    e->magic = 0;
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!e || e->magic != kEventMagic) die("bad event", e);
    ihipStream_t *s = S(st);
    e->last = s;
    enqueue(st, [e] {
        if (e->magic != kEventMagic) die("event destroyed before it ran", e);
        e->recorded = true;
    }, "event");
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    if (!e || e->magic != kEventMagic) die("bad event sync", e);
    if (e->last) flush(S(e->last == &g_null ? nullptr : e->last));
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    if (!a->recorded || !b->recorded) return hipErrorNotReady;
    *ms = 1.0f;
    return hipSuccess;
}

void hipstub_report(void)
{
    std::lock_guard<std::recursive_mutex> l(mu);
    size_t pending = g_null.q.size();
    for (auto *s : g_streams) pending += s->q.size();
    fprintf(stderr, "hipstub: %ld launches, %ld graph launches, %ld copies; %zu live streams, %zu live "
                    "device allocations, %zu queued ops\n",
            g_launches, g_graph_launches, g_copies, g_streams.size(), g_dev.size(), pending);
}
}
