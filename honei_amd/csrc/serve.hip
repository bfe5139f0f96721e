// serve.hip -- request server over the FP32-MG / FP64-IR engine: Poisson Q1
// solves of mixed grid sizes with continuous batching.
//
// Layout: one hierarchy (FillMatrix<POISSON, DIRICHLET_0> bands and inverted
// diagonals for levels min_level..max_served_level, FP64 bands for every
// served level) shared by all requests, and one *bucket* per served level: a
// device table of up to max_slots slot entries, a stream, a pinned report
// array and (optionally) the bucket's tick captured as a graph.
//
// Scheduling: each step, for every bucket whose previous tick has finished
// (event query, no blocking) the server
//   1. reads the slots' reports and retires finished requests,
//   2. admits queued requests of that level into free table entries while
//      the memory budget allows,
//   3. launches the next tick of the level.
// Buckets tick independently, so small grids are not held back by large
// ones, and their work overlaps on the device.  Memory: a request holds one
// slot (slot_bytes of its level); released slots stay cached per level and
// are freed when another level needs the memory.  Admission is first come,
// first served on memory: when the oldest waiting request does not fit, no
// younger request is admitted ahead of it.
#include "mg_engine.hh"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

using namespace hsolve;

namespace
{
    struct Request
    {
        uint64_t id = 0, tag = 0, level = 0, seed = 0;
        const double * b = nullptr;
        double * x = nullptr;
        double t_submit = 0.0, t_admit = 0.0;
    };

    struct Bucket
    {
        uint64_t L = 0;
        uint64_t slot_bytes = 0;
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        bool inflight = false;
        mg::Slot * table_host = nullptr;     // pinned image, max_slots entries
        mg::Slot * table = nullptr;          // device
        mg::Report * rep_host = nullptr;     // pinned
        mg::Report * rep = nullptr;          // device
        std::vector<void *> mem;             // slot allocation per table entry
        std::vector<Request> req;            // request per table entry
        std::vector<char> used;
        unsigned extent = 0;                 // entries covered by the last launched tick
        unsigned cap = 0;                    // table entries in use at most for this level
        unsigned running = 0;
        bool dirty = false;
        hipGraphExec_t graphs[16] = {};      // captured tick per launch extent 2^k
        std::deque<Request> queue;
        std::vector<void *> cache;           // released slot allocations
    };

    double seconds_since(std::chrono::steady_clock::time_point t0)
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

struct hsolve_serve_s
{
    hsolve_serve_params prm;
    mg::Hierarchy h;
    std::vector<Bucket> buckets;             // index L - min_served_level
    std::deque<hsolve_serve_result> done;
    std::chrono::steady_clock::time_point t0;
    uint64_t next_id = 0;
    uint64_t in_use = 0, cached = 0;
    uint64_t submitted = 0, admitted = 0, completed = 0, queued = 0, running = 0, peak_running = 0;
    uint64_t ticks = 0, captures = 0;

    ~hsolve_serve_s()
    {
        for (Bucket & b : buckets)
        {
            if (b.st)
                (void)hipStreamSynchronize(b.st);
            for (hipGraphExec_t g : b.graphs)
                if (g)
                    (void)hipGraphExecDestroy(g);
            for (void * p : b.mem)
                if (p)
                    (void)hipFree(p);
            for (void * p : b.cache)
                (void)hipFree(p);
            if (b.table_host)
                (void)hipHostFree(b.table_host);
            if (b.rep_host)
                (void)hipHostFree(b.rep_host);
            if (b.ev)
                (void)hipEventDestroy(b.ev);
            if (b.st)
                (void)hipStreamDestroy(b.st);
        }
    }

    Bucket & bucket(uint64_t L) { return buckets[L - prm.min_served_level]; }

    // frees cached slots of other levels until `need` more bytes fit
    bool make_room(uint64_t need, uint64_t keep_level)
    {
        if (in_use + need > prm.memory_budget)
            return false;
        for (Bucket & b : buckets)
        {
            if (in_use + cached + need <= prm.memory_budget)
                break;
            if (b.L == keep_level)
                continue;
            while (! b.cache.empty() && in_use + cached + need > prm.memory_budget)
            {
                HIP_CHECK(hipFree(b.cache.back()));
                b.cache.pop_back();
                cached -= b.slot_bytes;
            }
        }
        return in_use + cached + need <= prm.memory_budget;
    }

    // does the head of level L's queue fit in the budget now?
    bool head_fits(const Bucket & b) const { return in_use + b.slot_bytes <= prm.memory_budget; }

    void admit(Bucket & b, uint64_t barrier)
    {
        while (! b.queue.empty())
        {
            Request & r = b.queue.front();
            if (r.id > barrier)
                return;
            unsigned idx = 0;
            while (idx < b.cap && b.used[idx])
                ++idx;
            if (idx == b.cap)
                return;   // every table entry of this level is taken
            void * mem = nullptr;
            if (! b.cache.empty())
            {
                if (in_use + b.slot_bytes > prm.memory_budget)
                    return;
                mem = b.cache.back();
                b.cache.pop_back();
                cached -= b.slot_bytes;
            }
            else
            {
                if (! make_room(b.slot_bytes, b.L))
                    return;
                HIP_CHECK(hipMalloc(&mem, b.slot_bytes));
            }
            in_use += b.slot_bytes;
            mg::Slot sl = mg::slot_carve(h, b.L, (char *)mem, true, true);
            double * own_b = const_cast<double *>(sl.b);
            const uint64_t n = mg::size_of(b.L);
            if (r.b)
                HIP_CHECK(hipMemcpyAsync(own_b, r.b, sizeof(double) * n, hipMemcpyDeviceToDevice, b.st));
            else
                mg::enqueue_rhs_uniform(own_b, b.L, r.seed, 1, b.st);
            if (r.x)
                sl.result = r.x;
            const mg::Scalars s0 = mg::start_scalars(true);
            HIP_CHECK(hipMemcpyAsync(sl.s, &s0, sizeof(s0), hipMemcpyHostToDevice, b.st));
            b.table_host[idx] = sl;
            b.mem[idx] = mem;
            b.used[idx] = 1;
            r.t_admit = seconds_since(t0);
            b.req[idx] = r;
            b.queue.pop_front();
            b.dirty = true;
            ++b.running;
            ++admitted;
            --queued;
            ++running;
            peak_running = std::max(peak_running, running);
        }
    }

    void collect(Bucket & b)
    {
        const double now = seconds_since(t0);
        for (unsigned i = 0; i < b.extent; ++i)
        {
            if (! b.used[i] || b.rep_host[i].phase != mg::PH_DONE)
                continue;
            const mg::Report & rp = b.rep_host[i];
            const Request & r = b.req[i];
            hsolve_serve_result res;
            res.id = r.id;
            res.tag = r.tag;
            res.level = r.level;
            res.status = rp.status ? HSOLVE_ERR_NOT_CONVERGED : HSOLVE_OK;
            res.outer_iters = rp.outer_iters;
            res.inner_iters = rp.inner_total;
            res.ticks = rp.ticks;
            res.initial_defect = rp.initial_defect;
            res.final_defect = rp.def_norm;
            res.t_submit = r.t_submit;
            res.t_admit = r.t_admit;
            res.t_done = now;
            done.push_back(res);
            b.cache.push_back(b.mem[i]);
            cached += b.slot_bytes;
            in_use -= b.slot_bytes;
            b.mem[i] = nullptr;
            b.used[i] = 0;
            std::memset(&b.table_host[i], 0, sizeof(mg::Slot));
            b.dirty = true;
            --b.running;
            ++completed;
            --running;
        }
    }

    void launch(Bucket & b)
    {
        unsigned hw = 0;
        for (unsigned i = 0; i < b.cap; ++i)
            if (b.used[i])
                hw = i + 1;
        if (hw == 0)
            return;
        // launch extent: the high-water entry rounded up to a power of two
        // (entries past it are empty; kernels of empty entries return at once)
        unsigned ext = 1, lg = 0;
        while (ext < hw)
        {
            ext <<= 1;
            ++lg;
        }
        if (b.dirty || ext != b.extent)
        {
            HIP_CHECK(hipMemcpyAsync(b.table, b.table_host, sizeof(mg::Slot) * ext, hipMemcpyHostToDevice, b.st));
            b.dirty = false;
        }
        b.extent = ext;
        if (prm.use_graphs)
        {
            hipGraphExec_t & gx = b.graphs[lg];
            if (gx == nullptr)
            {
                hipGraph_t g = nullptr;
                HIP_CHECK(hipStreamBeginCapture(b.st, hipStreamCaptureModeThreadLocal));
                mg::enqueue_tick(h, b.L, b.table, ext, b.st);
                mg::enqueue_report(b.table, ext, b.rep, b.st);
                HIP_CHECK(hipStreamEndCapture(b.st, &g));
                const hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_CHECK(e);
                ++captures;
            }
            HIP_CHECK(hipGraphLaunch(gx, b.st));
        }
        else
        {
            mg::enqueue_tick(h, b.L, b.table, ext, b.st);
            mg::enqueue_report(b.table, ext, b.rep, b.st);
        }
        HIP_CHECK(hipMemcpyAsync(b.rep_host, b.rep, sizeof(mg::Report) * ext, hipMemcpyDeviceToHost, b.st));
        HIP_CHECK(hipEventRecord(b.ev, b.st));
        b.inflight = true;
        ++ticks;
    }

    uint64_t step()
    {
        // collect every level whose tick is over
        for (Bucket & b : buckets)
        {
            if (! b.inflight)
                continue;
            const hipError_t e = hipEventQuery(b.ev);
            if (e == hipErrorNotReady)
                continue;
            HIP_CHECK(e);
            b.inflight = false;
            collect(b);
        }
        // first come, first served on memory: the oldest queue head that
        // does not fit bars every younger request
        uint64_t barrier = std::numeric_limits<uint64_t>::max();
        for (Bucket & b : buckets)
            if (! b.queue.empty() && ! head_fits(b))
                barrier = std::min(barrier, b.queue.front().id);
        uint64_t busy = 0;
        for (Bucket & b : buckets)
        {
            if (! b.inflight)
            {
                admit(b, barrier);
                if (b.running)
                    launch(b);
            }
            busy += b.inflight ? 1 : 0;
        }
        return busy;
    }

    uint64_t submit(uint64_t level, const double * b, uint64_t seed, double * x, uint64_t tag, double t_submit)
    {
        require(level >= prm.min_served_level && level <= prm.max_served_level, HSOLVE_ERR_INVALID,
                "serve_submit: level outside the served range");
        Bucket & bk = bucket(level);
        if (bk.slot_bytes > prm.memory_budget)
            throw Error(HSOLVE_ERR_BUDGET, "serve_submit: one request of this level exceeds the memory budget");
        Request r;
        r.id = next_id++;
        r.tag = tag;
        r.level = level;
        r.seed = seed;
        r.b = b;
        r.x = x;
        r.t_submit = t_submit < 0.0 ? seconds_since(t0) : t_submit;
        bk.queue.push_back(r);
        ++submitted;
        ++queued;
        return r.id;
    }
};

extern "C" {

int hsolve_serve_create(const hsolve_serve_params * params, hsolve_serve * out)
{
    HSOLVE_API_BEGIN
    require(params && out, HSOLVE_ERR_INVALID, "serve_create: null pointer");
    *out = nullptr;
    const hsolve_serve_params & p = *params;
    require(p.min_served_level > p.min_level && p.min_served_level <= p.max_served_level &&
                p.max_served_level < (uint64_t)mg::kMaxLevels,
            HSOLVE_ERR_INVALID, "serve_create: need min_level < min_served_level <= max_served_level <= 14");
    require(p.max_slots >= 1 && p.max_slots <= 4096, HSOLVE_ERR_INVALID, "serve_create: max_slots in [1, 4096]");
    std::unique_ptr<hsolve_serve_s> srv(new hsolve_serve_s());
    srv->prm = p;
    hsolve_mg_params mp;
    std::memset(&mp, 0, sizeof(mp));
    mp.min_level = p.min_level;
    mp.max_level = p.max_served_level;
    mp.n_max_iter = p.n_max_iter;
    mp.initial_zero = p.initial_zero;
    mp.convergence_check = p.convergence_check;
    mp.tolerance = p.tolerance;
    mp.n_pre_smooth = p.n_pre_smooth;
    mp.adapt_correction_factor = p.adapt_correction_factor;
    for (int i = 0; i < 8; ++i)
        mp.macro_border_mask[i] = p.macro_border_mask[i];
    mp.max_outer = p.max_outer;
    mg::hierarchy_alloc(srv->h, mg::make_params(mp), p.max_served_level);
    for (uint64_t L = p.min_served_level; L <= p.max_served_level; ++L)
        mg::hierarchy_alloc64(srv->h, L);
    mg::hierarchy_fill_poisson(srv->h, 0);
    HIP_CHECK(hipDeviceSynchronize());
    srv->buckets.resize(p.max_served_level - p.min_served_level + 1);
    for (uint64_t L = p.min_served_level; L <= p.max_served_level; ++L)
    {
        Bucket & b = srv->bucket(L);
        b.L = L;
        b.slot_bytes = mg::slot_bytes(srv->h, L, true, true);
        HIP_CHECK(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
        // a tick covers the table entries up to the high-water entry rounded
        // up to a power of two, so every per-entry array holds that many
        uint64_t entries = 1;
        while (entries < p.max_slots)
            entries <<= 1;
        HIP_CHECK(hipHostMalloc((void **)&b.table_host, sizeof(mg::Slot) * entries, hipHostMallocDefault));
        HIP_CHECK(hipHostMalloc((void **)&b.rep_host, sizeof(mg::Report) * entries, hipHostMallocDefault));
        std::memset(b.table_host, 0, sizeof(mg::Slot) * entries);
        std::memset(b.rep_host, 0, sizeof(mg::Report) * entries);
        b.table = (mg::Slot *)srv->h.alloc(sizeof(mg::Slot) * entries);
        b.rep = (mg::Report *)srv->h.alloc(sizeof(mg::Report) * entries);
        HIP_CHECK(hipMemcpy(b.table, b.table_host, sizeof(mg::Slot) * entries, hipMemcpyHostToDevice));
        b.mem.assign(entries, nullptr);
        b.req.resize(entries);
        b.used.assign(entries, 0);
        b.cap = (unsigned)p.max_slots;
        if (p.rows_per_tick)
            b.cap = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(p.max_slots, p.rows_per_tick / mg::size_of(L)));
    }
    srv->t0 = std::chrono::steady_clock::now();
    *out = srv.release();
    HSOLVE_API_END
}

int hsolve_serve_destroy(hsolve_serve srv)
{
    HSOLVE_API_BEGIN
    delete srv;
    HSOLVE_API_END
}

int hsolve_serve_request_bytes(hsolve_serve srv, uint64_t level, uint64_t * bytes)
{
    HSOLVE_API_BEGIN
    require(srv && bytes, HSOLVE_ERR_INVALID, "serve_request_bytes: null pointer");
    require(level >= srv->prm.min_served_level && level <= srv->prm.max_served_level, HSOLVE_ERR_INVALID,
            "serve_request_bytes: level outside the served range");
    *bytes = srv->bucket(level).slot_bytes;
    HSOLVE_API_END
}

int hsolve_serve_submit(hsolve_serve srv, uint64_t level, const double * b, uint64_t rhs_seed, double * x,
                        uint64_t tag, double t_submit, uint64_t * id)
{
    HSOLVE_API_BEGIN
    require(srv != nullptr, HSOLVE_ERR_INVALID, "serve_submit: null server");
    const uint64_t i = srv->submit(level, b, rhs_seed, x, tag, t_submit);
    if (id)
        *id = i;
    HSOLVE_API_END
}

int hsolve_serve_step(hsolve_serve srv, uint64_t * busy)
{
    HSOLVE_API_BEGIN
    require(srv != nullptr, HSOLVE_ERR_INVALID, "serve_step: null server");
    const uint64_t k = srv->step();
    if (busy)
        *busy = k;
    HSOLVE_API_END
}

int hsolve_serve_poll(hsolve_serve srv, hsolve_serve_result * out, uint64_t max, uint64_t * count)
{
    HSOLVE_API_BEGIN
    require(srv && count && (out || max == 0), HSOLVE_ERR_INVALID, "serve_poll: null pointer");
    uint64_t k = 0;
    while (k < max && ! srv->done.empty())
    {
        out[k++] = srv->done.front();
        srv->done.pop_front();
    }
    *count = k;
    HSOLVE_API_END
}

int hsolve_serve_now(hsolve_serve srv, double * seconds)
{
    HSOLVE_API_BEGIN
    require(srv && seconds, HSOLVE_ERR_INVALID, "serve_now: null pointer");
    *seconds = seconds_since(srv->t0);
    HSOLVE_API_END
}

int hsolve_serve_get_stats(hsolve_serve srv, hsolve_serve_stats * out)
{
    HSOLVE_API_BEGIN
    require(srv && out, HSOLVE_ERR_INVALID, "serve_get_stats: null pointer");
    out->bytes_in_use = srv->in_use;
    out->bytes_held = srv->in_use + srv->cached;
    out->bytes_shared = srv->h.bytes;
    out->budget = srv->prm.memory_budget;
    out->submitted = srv->submitted;
    out->admitted = srv->admitted;
    out->completed = srv->completed;
    out->queued = srv->queued;
    out->running = srv->running;
    out->peak_running = srv->peak_running;
    out->ticks = srv->ticks;
    out->graph_captures = srv->captures;
    HSOLVE_API_END
}

int hsolve_serve_run(hsolve_serve srv, uint64_t count, const double * arrival_s, const uint64_t * levels,
                     const uint64_t * seeds, double * const * x, hsolve_serve_result * results)
{
    HSOLVE_API_BEGIN
    require(srv && (count == 0 || (arrival_s && levels && seeds && results)), HSOLVE_ERR_INVALID,
            "serve_run: null pointer");
    for (uint64_t i = 1; i < count; ++i)
        require(arrival_s[i] >= arrival_s[i - 1], HSOLVE_ERR_INVALID, "serve_run: arrivals must be sorted");
    require(srv->running == 0 && srv->queued == 0 && srv->done.empty(), HSOLVE_ERR_INVALID,
            "serve_run: the server must be idle");
    const double start = seconds_since(srv->t0);
    const uint64_t first_id = srv->next_id;
    uint64_t next = 0, finished = 0;
    while (finished < count)
    {
        const double now = seconds_since(srv->t0) - start;
        while (next < count && arrival_s[next] <= now)
        {
            srv->submit(levels[next], nullptr, seeds[next], x ? x[next] : nullptr, next, start + arrival_s[next]);
            ++next;
        }
        const uint64_t busy = srv->step();
        while (! srv->done.empty())
        {
            const hsolve_serve_result r = srv->done.front();
            srv->done.pop_front();
            require(r.id >= first_id && r.id - first_id < count, HSOLVE_ERR_INVALID,
                    "serve_run: a request from outside this trace finished");
            const uint64_t i = r.tag;
            results[i] = r;
            results[i].t_submit -= start;
            results[i].t_admit -= start;
            results[i].t_done -= start;
            ++finished;
        }
        if (busy == 0 && srv->queued == 0 && next < count)
        {
            // idle until the next arrival
            const double wait = arrival_s[next] - (seconds_since(srv->t0) - start);
            if (wait > 2e-4)
                std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
        }
    }
    HSOLVE_API_END
}

}  // extern "C"
