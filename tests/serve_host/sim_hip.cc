// sim_hip.cc -- the HIP runtime calls of the request server, simulated on the
// host (deferred stream queues, events, graph capture and replay, device and
// pinned allocations), with AddressSanitizer poisoning of pinned ranges that
// a queued copy still owns.
#include "sim.hh"

#include <sanitizer/asan_interface.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <vector>

struct ihipGraph
{
    std::vector<std::function<void()>> ops;
};
struct hipGraphExec
{
    std::vector<std::function<void()>> ops;
};
struct ihipStream_t
{
    std::deque<std::function<void()>> q;
    ihipGraph * cap = nullptr;   // the graph a capture records into
};
struct EventState
{
    uint64_t want = 0, reached = 0;
    ihipStream_t * st = nullptr;
};
struct ihipEvent_t
{
    std::shared_ptr<EventState> e = std::make_shared<EventState>();
};

namespace
{
    enum Kind
    {
        DEVICE,
        PINNED
    };
    struct Alloc
    {
        size_t bytes;
        Kind kind;
    };

    std::map<uintptr_t, Alloc> g_allocs;
    std::set<ihipStream_t *> g_streams;
    ihipStream_t g_null;
    std::mt19937_64 g_rng(1);
    uint64_t g_device_bytes = 0;
    uint64_t g_marks = 0;

    ihipStream_t * S(hipStream_t st) { return st ? st : &g_null; }

    bool any_capture()
    {
        for (ihipStream_t * s : g_streams)
            if (s->cap)
                return true;
        return false;
    }

    void no_capture(const char * what)
    {
        if (any_capture())
            sim::fail("%s while a stream captures (forbidden in thread-local capture mode)", what);
    }

    // the allocation of kind k holding [p, p + n), or null
    const Alloc * find(const void * p, size_t n, Kind k)
    {
        const uintptr_t a = (uintptr_t)p;
        auto it = g_allocs.upper_bound(a);
        if (it == g_allocs.begin())
            return nullptr;
        --it;
        if (it->second.kind != k || a + n > it->first + it->second.bytes)
            return nullptr;
        return &it->second;
    }

    bool run_one(ihipStream_t * s)
    {
        if (s->q.empty())
            return false;
        std::function<void()> op = std::move(s->q.front());
        s->q.pop_front();
        op();
        return true;
    }

    void drain(ihipStream_t * s)
    {
        if (s->cap)
            sim::fail("synchronizing a stream that captures");
        while (run_one(s))
        {
        }
    }

    void drain_all()
    {
        for (ihipStream_t * s : g_streams)
            drain(s);
        drain(&g_null);
    }

    // the device works on its own: now and then one queued operation of a
    // random stream runs
    void progress()
    {
        if (g_streams.empty() || g_rng() % 4 != 0)
            return;
        auto it = g_streams.begin();
        std::advance(it, g_rng() % g_streams.size());
        if (! (*it)->cap)
            run_one(*it);
    }

    void * take(size_t bytes, Kind k)
    {
        const size_t n = (bytes + 255) & ~(size_t)255;
        void * p = std::aligned_alloc(256, n ? n : 256);
        if (! p)
            return nullptr;
        g_allocs[(uintptr_t)p] = Alloc{bytes, k};
        if (k == DEVICE)
            g_device_bytes += bytes;
        return p;
    }

    void give(void * p, Kind k, const char * what)
    {
        auto it = g_allocs.find((uintptr_t)p);
        if (it == g_allocs.end() || it->second.kind != k)
            sim::fail("%s of %p, which is not an allocation of that kind", what, p);
        drain_all();   // both calls wait for the device
        if (k == DEVICE)
            g_device_bytes -= it->second.bytes;
        g_allocs.erase(it);
        std::free(p);
    }
}

namespace sim
{
    void enqueue(hipStream_t st, std::function<void()> op)
    {
        ihipStream_t * s = S(st);
        if (s->cap)
            s->cap->ops.push_back(std::move(op));
        else if (s == &g_null)
            op();
        else
            s->q.push_back(std::move(op));
    }

    void seed(uint64_t v) { g_rng.seed(v); }

    void fail(const char * fmt, ...)
    {
        va_list ap;
        va_start(ap, fmt);
        std::fprintf(stderr, "SIM CONTRACT: ");
        std::vfprintf(stderr, fmt, ap);
        std::fprintf(stderr, "\n");
        va_end(ap);
        std::fflush(stderr);
        std::abort();
    }

    uint64_t device_bytes() { return g_device_bytes; }
}

extern "C" {

const char * hipGetErrorString(hipError_t e)
{
    switch (e)
    {
        case hipSuccess: return "hipSuccess";
        case hipErrorNotReady: return "hipErrorNotReady";
        case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
        default: return "hipError (simulated)";
    }
}

hipError_t hipMalloc(void ** ptr, size_t size)
{
    no_capture("hipMalloc");
    progress();
    *ptr = size ? take(size, DEVICE) : nullptr;
    return (size && ! *ptr) ? hipErrorOutOfMemory : hipSuccess;
}

hipError_t hipFree(void * p)
{
    no_capture("hipFree");
    if (p)
        give(p, DEVICE, "hipFree");
    return hipSuccess;
}

hipError_t hipHostMalloc(void ** ptr, size_t size, unsigned int)
{
    no_capture("hipHostMalloc");
    *ptr = take(size, PINNED);
    return *ptr ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void * p)
{
    no_capture("hipHostFree");
    if (p)
        give(p, PINNED, "hipHostFree");
    return hipSuccess;
}

hipError_t hipMemcpy(void * dst, const void * src, size_t n, hipMemcpyKind kind)
{
    no_capture("hipMemcpy");
    const bool ok = kind == hipMemcpyHostToDevice   ? find(dst, n, DEVICE) != nullptr
                    : kind == hipMemcpyDeviceToHost ? find(src, n, DEVICE) != nullptr
                    : kind == hipMemcpyDeviceToDevice ? (find(dst, n, DEVICE) && find(src, n, DEVICE))
                                                      : false;
    if (! ok)
        sim::fail("hipMemcpy of %zu bytes with pointers that do not match its kind %d", n, (int)kind);
    drain(&g_null);
    std::memcpy(dst, src, n);
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void * dst, const void * src, size_t n, hipMemcpyKind kind, hipStream_t st)
{
    progress();
    ihipStream_t * s = S(st);
    if (kind == hipMemcpyHostToDevice)
    {
        if (! find(dst, n, DEVICE) || find(src, 1, DEVICE))
            sim::fail("hipMemcpyAsync H2D of %zu bytes: the destination is not device memory", n);
        if (find(src, n, PINNED))
        {
            // the host must not touch the source until the copy has run
            ASAN_POISON_MEMORY_REGION(src, n);
            sim::enqueue(st, [dst, src, n] {
                ASAN_UNPOISON_MEMORY_REGION(src, n);
                std::memcpy(dst, src, n);
            });
        }
        else
        {
            // pageable host memory: HIP copies it synchronously
            if (s->cap)
                sim::fail("a pageable host copy inside a capture");
            drain(s);
            std::memcpy(dst, src, n);
        }
    }
    else if (kind == hipMemcpyDeviceToHost)
    {
        if (! find(src, n, DEVICE) || find(dst, 1, DEVICE))
            sim::fail("hipMemcpyAsync D2H of %zu bytes: the source is not device memory", n);
        if (find(dst, n, PINNED))
        {
            // the host must not read the destination before the copy has run
            ASAN_POISON_MEMORY_REGION(dst, n);
            sim::enqueue(st, [dst, src, n] {
                ASAN_UNPOISON_MEMORY_REGION(dst, n);
                std::memcpy(dst, src, n);
            });
        }
        else
        {
            if (s->cap)
                sim::fail("a pageable host copy inside a capture");
            drain(s);
            std::memcpy(dst, src, n);
        }
    }
    else if (kind == hipMemcpyDeviceToDevice)
    {
        if (! find(dst, n, DEVICE) || ! find(src, n, DEVICE))
            sim::fail("hipMemcpyAsync D2D of %zu bytes outside device allocations", n);
        sim::enqueue(st, [dst, src, n] { std::memcpy(dst, src, n); });
    }
    else
        sim::fail("hipMemcpyAsync kind %d is not simulated", (int)kind);
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t * st, unsigned int)
{
    *st = new ihipStream_t();
    g_streams.insert(*st);
    return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t st)
{
    if (! g_streams.count(st))
        sim::fail("hipStreamDestroy of an unknown stream");
    drain(st);
    g_streams.erase(st);
    delete st;
    return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t st)
{
    drain(S(st));
    return hipSuccess;
}

hipError_t hipDeviceSynchronize()
{
    no_capture("hipDeviceSynchronize");
    drain_all();
    return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t * ev, unsigned)
{
    *ev = new ihipEvent_t();
    return hipSuccess;
}

hipError_t hipEventDestroy(hipEvent_t ev)
{
    delete ev;
    return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t ev, hipStream_t st)
{
    ihipStream_t * s = S(st);
    if (s->cap)
        sim::fail("hipEventRecord inside a capture is not simulated");
    std::shared_ptr<EventState> e = ev->e;
    e->want = ++g_marks;
    e->st = s;
    const uint64_t mark = e->want;
    sim::enqueue(st, [e, mark] { e->reached = mark; });
    return hipSuccess;
}

hipError_t hipEventQuery(hipEvent_t ev)
{
    progress();
    EventState & e = *ev->e;
    if (e.reached == e.want)
        return hipSuccess;
    // the device gets some way through the stream: nothing, part, or all of it
    const uint64_t steps = g_rng() % 3 == 0 ? 0 : g_rng() % 4 + 1;
    for (uint64_t i = 0; i < steps && e.reached != e.want; ++i)
        run_one(e.st);
    return e.reached == e.want ? hipSuccess : hipErrorNotReady;
}

hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode)
{
    ihipStream_t * s = S(st);
    if (s->cap || s == &g_null)
        sim::fail("hipStreamBeginCapture on a stream that captures or on the null stream");
    s->cap = new ihipGraph();
    return hipSuccess;
}

hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t * g)
{
    ihipStream_t * s = S(st);
    if (! s->cap)
        sim::fail("hipStreamEndCapture without a capture");
    *g = s->cap;
    s->cap = nullptr;
    return hipSuccess;
}

hipError_t hipGraphInstantiate(hipGraphExec_t * x, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    *x = new hipGraphExec{g->ops};
    return hipSuccess;
}

hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}

hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
    delete x;
    return hipSuccess;
}

hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t st)
{
    progress();
    for (const std::function<void()> & op : x->ops)
        sim::enqueue(st, op);
    return hipSuccess;
}

}  // extern "C"
