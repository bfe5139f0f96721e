// driver.cc -- drives the request server's C-ABI (hsolve_serve_*) over the
// simulated device of sim_hip.cc / sim_mg.cc with seeded random workloads:
// mixed levels, seeded and caller-owned right-hand sides, caller-owned and
// discarded results, memory budgets from one largest request up to
// unlimited, 1..9 slots per level, tick row caps, graph replay on and off,
// the trace replayer, and destruction with requests in flight.
//
// Checked per request: it finishes exactly once, at its level, with the
// tick count its right-hand side fixes, and a caller-owned result holds its
// own values.  Checked after every scheduling pass: the held bytes stay
// within the budget and the counters add up.  Checked per level: requests
// are admitted in submission order.  At the end every device byte is freed.
// Usage: driver SEED [SEED ...]; prints "ok SEED ..." per seed.
#include "hsolve.h"
#include "sim.hh"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#define CHECK(c, ...)                                                                                  \
    do                                                                                                 \
    {                                                                                                  \
        if (! (c))                                                                                     \
        {                                                                                              \
            std::fprintf(stderr, "FAIL seed %llu, %s:%d: ", (unsigned long long)g_seed, __FILE__,      \
                         __LINE__);                                                                    \
            std::fprintf(stderr, __VA_ARGS__);                                                         \
            std::fprintf(stderr, "\n");                                                                \
            std::exit(1);                                                                              \
        }                                                                                              \
    } while (0)
#define OK(expr)                                                                                       \
    do                                                                                                 \
    {                                                                                                  \
        const int rc_ = (expr);                                                                        \
        CHECK(rc_ == HSOLVE_OK, "%s returned %d: %s", #expr, rc_, sim_last_error());                   \
    } while (0)

namespace
{
    uint64_t g_seed = 0;

    uint64_t side(uint64_t L) { return (1ull << L) + 1; }
    uint64_t size_of(uint64_t L) { return side(L) * side(L); }

    struct Req
    {
        uint64_t level = 0, seed = 0, id = 0;
        double * b = nullptr;         // caller-owned right-hand side (device) or null
        double * x = nullptr;         // caller-owned result (device) or null
        std::vector<double> bh;       // the right-hand side's values
        bool done = false;
        double t_admit = 0.0;
    };

    hsolve_serve_params make_params(uint64_t min_served, uint64_t max_served, uint64_t budget, uint64_t slots,
                                    uint64_t rows_per_tick, int graphs)
    {
        hsolve_serve_params p = {};
        p.min_level = 1;
        p.min_served_level = min_served;
        p.max_served_level = max_served;
        p.n_max_iter = 16;
        p.initial_zero = 1;
        p.convergence_check = 1;
        p.tolerance = 1e-8;
        p.n_pre_smooth = 2;
        p.adapt_correction_factor = 1.0;
        for (int i = 0; i < 8; ++i)
            p.macro_border_mask[i] = 2;
        p.max_outer = 50;
        p.memory_budget = budget;
        p.max_slots = slots;
        p.rows_per_tick = rows_per_tick;
        p.use_graphs = graphs;
        return p;
    }

    void check_result(std::vector<Req> & reqs, const std::map<uint64_t, size_t> & by_id,
                      const hsolve_serve_result & r)
    {
        auto it = by_id.find(r.id);
        CHECK(it != by_id.end(), "result for unknown id %llu", (unsigned long long)r.id);
        Req & q = reqs[it->second];
        CHECK(! q.done, "request %llu finished twice", (unsigned long long)r.id);
        q.done = true;
        q.t_admit = r.t_admit;
        CHECK(r.level == q.level, "request %llu: level %llu, submitted at %llu", (unsigned long long)r.id,
              (unsigned long long)r.level, (unsigned long long)q.level);
        CHECK(r.status == HSOLVE_OK, "request %llu: status %d", (unsigned long long)r.id, r.status);
        const uint64_t want = sim::ticks_for(q.bh.data());
        CHECK(r.outer_iters == want && r.ticks == want, "request %llu: %llu / %llu ticks, want %llu",
              (unsigned long long)r.id, (unsigned long long)r.outer_iters, (unsigned long long)r.ticks,
              (unsigned long long)want);
        CHECK(r.initial_defect == q.bh[1], "request %llu: solved another right-hand side",
              (unsigned long long)r.id);
        CHECK(r.t_submit <= r.t_admit && r.t_admit <= r.t_done, "request %llu: times out of order",
              (unsigned long long)r.id);
        if (q.x)
        {
            std::vector<double> xh(q.bh.size());
            OK(hipMemcpy(xh.data(), q.x, sizeof(double) * xh.size(), hipMemcpyDeviceToHost) == hipSuccess
                   ? HSOLVE_OK
                   : HSOLVE_ERR_HIP);
            for (size_t j = 0; j < xh.size(); ++j)
                CHECK(xh[j] == sim::result_value(q.bh[j], want), "request %llu: x[%zu] = %g, want %g",
                      (unsigned long long)r.id, j, xh[j], sim::result_value(q.bh[j], want));
            (void)hipFree(q.x);
            q.x = nullptr;
        }
        if (q.b)
        {
            (void)hipFree(q.b);
            q.b = nullptr;
        }
    }

    void check_stats(hsolve_serve srv, uint64_t budget, uint64_t cap_total)
    {
        hsolve_serve_stats s;
        OK(hsolve_serve_get_stats(srv, &s));
        CHECK(s.bytes_held <= budget, "held %llu bytes over the budget %llu", (unsigned long long)s.bytes_held,
              (unsigned long long)budget);
        CHECK(s.bytes_in_use <= s.bytes_held, "in use above held");
        CHECK(s.submitted == s.admitted + s.queued, "submitted %llu != admitted %llu + queued %llu",
              (unsigned long long)s.submitted, (unsigned long long)s.admitted, (unsigned long long)s.queued);
        CHECK(s.admitted == s.completed + s.running, "admitted != completed + running");
        CHECK(s.running <= cap_total, "running %llu above the slot caps %llu", (unsigned long long)s.running,
              (unsigned long long)cap_total);
    }

    void run_case(uint64_t seed)
    {
        g_seed = seed;
        std::mt19937_64 rng(seed);
        sim::seed(seed ^ 0x5EEDull);
        const uint64_t max_served = 2 + rng() % 4;
        const uint64_t min_served = 2 + rng() % (max_served - 1);
        const uint64_t slots = 1 + rng() % 9;
        const uint64_t rows_per_tick = rng() % 3 == 0 ? 2 * size_of(max_served) : 0;
        const int graphs = (int)(rng() % 2);

        // request sizes, from a server with no budget limit
        std::vector<uint64_t> rb(max_served + 1, 0);
        {
            hsolve_serve probe = nullptr;
            const hsolve_serve_params p = make_params(min_served, max_served, 1ull << 40, slots, 0, graphs);
            OK(hsolve_serve_create(&p, &probe));
            for (uint64_t L = min_served; L <= max_served; ++L)
                OK(hsolve_serve_request_bytes(probe, L, &rb[L]));
            OK(hsolve_serve_destroy(probe));
        }
        const uint64_t mode = rng() % 3;
        const uint64_t budget = mode == 0   ? rb[max_served]
                                : mode == 1 ? 2 * rb[max_served] + 3 * rb[min_served]
                                            : 1ull << 40;
        hsolve_serve srv = nullptr;
        const hsolve_serve_params p = make_params(min_served, max_served, budget, slots, rows_per_tick, graphs);
        OK(hsolve_serve_create(&p, &srv));
        uint64_t cap_total = 0;
        for (uint64_t L = min_served; L <= max_served; ++L)
            cap_total += rows_per_tick ? std::max<uint64_t>(1, std::min<uint64_t>(slots, rows_per_tick / size_of(L)))
                                       : slots;

        // rejected requests
        uint64_t id = 0;
        CHECK(hsolve_serve_submit(srv, max_served + 1, nullptr, 1, nullptr, 0, -1.0, &id) == HSOLVE_ERR_INVALID,
              "a level above the served range was accepted");
        CHECK(hsolve_serve_submit(srv, min_served - 1, nullptr, 1, nullptr, 0, -1.0, &id) == HSOLVE_ERR_INVALID,
              "a level below the served range was accepted");
        if (mode == 0 && min_served < max_served)
        {
            // a server whose budget holds one request of the smallest level only
            hsolve_serve small = nullptr;
            const hsolve_serve_params ps = make_params(min_served, max_served, rb[min_served], slots, 0, graphs);
            OK(hsolve_serve_create(&ps, &small));
            CHECK(hsolve_serve_submit(small, max_served, nullptr, 1, nullptr, 0, -1.0, &id) == HSOLVE_ERR_BUDGET,
                  "a request above the budget was accepted");
            OK(hsolve_serve_destroy(small));
        }

        const uint64_t count = 150 + rng() % 150;
        const bool use_run = seed % 4 == 3;
        const bool destroy_early = ! use_run && seed % 5 == 4;
        std::vector<Req> reqs(count);
        for (uint64_t i = 0; i < count; ++i)
        {
            Req & q = reqs[i];
            q.level = min_served + rng() % (max_served - min_served + 1);
            q.seed = rng();
            const uint64_t n = size_of(q.level);
            q.bh.resize(n);
            const bool own_b = ! use_run && rng() % 3 == 0;
            for (uint64_t j = 0; j < n; ++j)
                q.bh[j] = own_b ? sim::rhs_value(q.seed ^ 0xABCDull, j) : sim::rhs_value(q.seed, j);
            if (own_b)
            {
                CHECK(hipMalloc((void **)&q.b, sizeof(double) * n) == hipSuccess, "hipMalloc");
                CHECK(hipMemcpy(q.b, q.bh.data(), sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess,
                      "hipMemcpy");
            }
            if (rng() % 2)
                CHECK(hipMalloc((void **)&q.x, sizeof(double) * n) == hipSuccess, "hipMalloc");
        }

        std::map<uint64_t, size_t> by_id;
        if (use_run)
        {
            std::vector<double> arrival(count);
            std::vector<uint64_t> levels(count), seeds(count);
            std::vector<double *> xs(count);
            for (uint64_t i = 0; i < count; ++i)
            {
                arrival[i] = 2e-6 * (double)i;
                levels[i] = reqs[i].level;
                seeds[i] = reqs[i].seed;
                xs[i] = reqs[i].x;
            }
            std::vector<hsolve_serve_result> res(count);
            hsolve_serve_stats s0;
            OK(hsolve_serve_get_stats(srv, &s0));
            for (uint64_t i = 0; i < count; ++i)
                by_id[s0.submitted + i] = i;
            OK(hsolve_serve_run(srv, count, arrival.data(), levels.data(), seeds.data(), xs.data(), res.data()));
            for (uint64_t i = 0; i < count; ++i)
            {
                CHECK(res[i].tag == i, "trace result %llu carries tag %llu", (unsigned long long)i,
                      (unsigned long long)res[i].tag);
                check_result(reqs, by_id, res[i]);
            }
            check_stats(srv, budget, cap_total);
        }
        else
        {
            uint64_t next = 0, finished = 0, passes = 0;
            std::vector<hsolve_serve_result> out(8);
            while (finished < count)
            {
                CHECK(++passes < 2000000, "no progress: %llu of %llu finished", (unsigned long long)finished,
                      (unsigned long long)count);
                for (uint64_t k = rng() % 4; k > 0 && next < count; --k, ++next)
                {
                    Req & q = reqs[next];
                    OK(hsolve_serve_submit(srv, q.level, q.b, q.seed, q.x, next, -1.0, &q.id));
                    by_id[q.id] = next;
                }
                uint64_t busy = 0;
                OK(hsolve_serve_step(srv, &busy));
                check_stats(srv, budget, cap_total);
                uint64_t got = 0;
                do
                {
                    const uint64_t want = 1 + rng() % out.size();
                    OK(hsolve_serve_poll(srv, out.data(), want, &got));
                    for (uint64_t i = 0; i < got; ++i)
                        check_result(reqs, by_id, out[i]);
                    finished += got;
                } while (got > 0);
                if (destroy_early && finished >= count / 2)
                    break;
            }
        }

        if (! destroy_early)
        {
            hsolve_serve_stats s;
            OK(hsolve_serve_get_stats(srv, &s));
            CHECK(s.running == 0 && s.queued == 0 && s.bytes_in_use == 0, "server not idle at the end");
            CHECK(s.completed == count, "completed %llu of %llu", (unsigned long long)s.completed,
                  (unsigned long long)count);
            // requests of one level are admitted in the order they came
            for (uint64_t L = min_served; L <= max_served; ++L)
            {
                double last = -1.0;
                for (const Req & q : reqs)
                    if (q.level == L)
                    {
                        CHECK(q.t_admit >= last, "level %llu admitted out of order", (unsigned long long)L);
                        last = q.t_admit;
                    }
            }
        }
        OK(hsolve_serve_destroy(srv));
        for (Req & q : reqs)
        {
            (void)hipFree(q.b);
            (void)hipFree(q.x);
        }
        CHECK(sim::device_bytes() == 0, "%llu device bytes left after destroy",
              (unsigned long long)sim::device_bytes());
        std::printf("ok %llu (levels %llu..%llu, budget mode %llu, slots %llu, row cap %llu, graphs %d, %s)\n",
                    (unsigned long long)seed, (unsigned long long)min_served, (unsigned long long)max_served,
                    (unsigned long long)mode, (unsigned long long)slots, (unsigned long long)rows_per_tick, graphs,
                    use_run ? "trace" : destroy_early ? "destroyed in flight" : "submit/step/poll");
        std::fflush(stdout);
    }
}

int main(int argc, char ** argv)
{
    for (int i = 1; i < argc; ++i)
        run_case(std::strtoull(argv[i], nullptr, 10));
    return 0;
}
