// sim_mg.cc -- the engine calls of the request server (mg_engine.hh),
// simulated on the host over sim_hip.cc's streams.  A slot's solve runs
// ticks_for(b) ticks; each tick reads the slot's right-hand side and writes
// its whole result array through the device table; the report copies each
// table entry's scalars.  Compiled as host-only HIP (the engine header holds
// device code that is not emitted here).
#include "mg_engine.hh"

#include "sim.hh"

#include <cmath>
#include <cstring>
#include <string>

namespace hsolve
{
    namespace
    {
        thread_local std::string g_err;
    }
    void set_last_error(const std::string & msg) { g_err = msg; }
}

extern "C" const char * sim_last_error() { return hsolve::g_err.c_str(); }

namespace sim
{
    double rhs_value(uint64_t seed, uint64_t i)
    {
        // 24-bit fractions: every sum result_value() forms is exact
        return (double)(hsolve::mix64(seed * 0x9E3779B97F4A7C15ull + i) >> 40) / 16777216.0 - 0.5;
    }

    uint64_t ticks_for(const double * b) { return 1 + (uint64_t)(std::fabs(b[1]) * 997.0) % 6; }
}

namespace hsolve
{
namespace mg
{
    namespace
    {
        constexpr uint64_t kHead = 512;   // scalars and partials
        uint64_t r256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }
    }

    void * Hierarchy::alloc(size_t b)
    {
        void * p = nullptr;
        HIP_CHECK(hipMalloc(&p, b));
        allocations.push_back(p);
        bytes += b;
        return p;
    }

    Hierarchy::~Hierarchy()
    {
        for (void * p : allocations)
            (void)hipFree(p);
    }

    Params make_params(const hsolve_mg_params & p)
    {
        Params q;
        std::memset(&q, 0, sizeof(q));
        q.min_level = p.min_level;
        q.n_max_iter = p.n_max_iter;
        q.max_outer = p.max_outer;
        return q;
    }

    void hierarchy_alloc(Hierarchy & h, const Params & p, uint64_t max_level)
    {
        h.p = p;
        h.max_level = max_level;
        for (uint64_t L = p.min_level; L <= max_level; ++L)
        {
            h.lv[L].n = size_of(L);
            h.lv[L].bands = (float *)h.alloc(sizeof(float) * 64);
        }
    }

    void hierarchy_alloc64(Hierarchy & h, uint64_t L) { h.bands64[L] = (double *)h.alloc(sizeof(double) * 64); }

    void hierarchy_fill_poisson(Hierarchy & h, hipStream_t st)
    {
        for (uint64_t L = h.p.min_level; L <= h.max_level; ++L)
        {
            float * b = h.lv[L].bands;
            sim::enqueue(st, [b] {
                for (int i = 0; i < 64; ++i)
                    b[i] = 1.0f;
            });
        }
    }

    uint64_t slot_bytes(const Hierarchy &, uint64_t L, bool own_b, bool own_result)
    {
        const uint64_t v = r256(sizeof(double) * size_of(L));
        return kHead + (own_b ? v : 0) + (own_result ? v : 0);
    }

    Slot slot_carve(const Hierarchy &, uint64_t L, char * base, bool own_b, bool own_result)
    {
        Slot s;
        std::memset(&s, 0, sizeof(s));
        s.s = (Scalars *)base;
        s.partials = (double *)(base + 256);
        uint64_t off = kHead;
        const uint64_t v = r256(sizeof(double) * size_of(L));
        if (own_b)
        {
            s.b = (const double *)(base + off);
            off += v;
        }
        if (own_result)
            s.result = (double *)(base + off);
        return s;
    }

    Scalars start_scalars(bool zero_guess)
    {
        Scalars s;
        std::memset(&s, 0, sizeof(s));
        s.phase = PH_START;
        s.zero_guess = zero_guess ? 1 : 0;
        return s;
    }

    void enqueue_rhs_uniform(double * b, uint64_t L, uint64_t seed, uint64_t, hipStream_t st)
    {
        sim::enqueue(st, [b, L, seed] {
            for (uint64_t i = 0; i < size_of(L); ++i)
                b[i] = sim::rhs_value(seed, i);
        });
    }

    void enqueue_tick(const Hierarchy &, uint64_t L, const Slot * table, unsigned nslots, hipStream_t st)
    {
        sim::enqueue(st, [L, table, nslots] {
            for (unsigned i = 0; i < nslots; ++i)
            {
                const Slot sl = table[i];
                if (! sl.s)
                    continue;
                Scalars & s = *sl.s;
                if (s.phase == PH_START)
                {
                    s.phase = PH_SOLVE;
                    s.ticks = 0;
                    s.initial_defect = sl.b[1];
                }
                if (s.phase != PH_SOLVE)
                    sim::fail("a tick of level %llu found table entry %u in phase %d (a finished or stale slot)",
                              (unsigned long long)L, i, s.phase);
                ++s.ticks;
                for (uint64_t j = 0; j < size_of(L); ++j)
                    sl.result[j] = sim::result_value(sl.b[j], s.ticks);
                if (s.ticks >= sim::ticks_for(sl.b))
                {
                    s.phase = PH_DONE;
                    s.outer_iters = s.ticks;
                    s.inner_total = 2 * s.ticks;
                    s.def_norm = 1e-9;
                    s.status = 0;
                }
            }
        });
    }

    void enqueue_report(const Slot * table, unsigned nslots, Report * reports, hipStream_t st)
    {
        sim::enqueue(st, [table, nslots, reports] {
            for (unsigned i = 0; i < nslots; ++i)
            {
                Report r;
                std::memset(&r, 0, sizeof(r));
                const Slot sl = table[i];
                if (sl.s)
                {
                    r.phase = sl.s->phase;
                    r.status = sl.s->status;
                    r.outer_iters = sl.s->outer_iters;
                    r.inner_total = sl.s->inner_total;
                    r.ticks = sl.s->ticks;
                    r.initial_defect = sl.s->initial_defect;
                    r.def_norm = sl.s->def_norm;
                }
                reports[i] = r;
            }
        });
    }
}  // namespace mg
}  // namespace hsolve
