// mg_solve.hip -- C-ABI of the multigrid handle (hsolve_mg_*): one hierarchy
// of the FP32-MG / FP64-IR engine (Q1 bands or ELL matrices) with a fixed
// number of slots.  A solve of `count` right-hand sides runs them in the
// same ticks (one launch per step for all of them), enqueues ticks in
// chunks (the first sized by the previous solve's tick count) and reads the
// slots' reports once per chunk.
#include "mg_engine.hh"

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

using namespace hsolve;

struct hsolve_mg_s
{
    mg::Hierarchy h;
    uint64_t max_level = 0;
    unsigned slots = 1;
    std::vector<mg::Slot> slot;         // host image of the table
    mg::Slot * table = nullptr;         // device, `slots` entries
    mg::Report * report = nullptr;      // device, `slots` entries
    mg::Report * report_host = nullptr; // pinned
    uint64_t last_ticks = 0;

    ~hsolve_mg_s()
    {
        if (report_host)
            (void)hipHostFree(report_host);
    }
};

namespace
{
    // the slot tables of a hierarchy whose levels are allocated
    void mg_slots(hsolve_mg_s * mg, uint64_t slots)
    {
        require(slots >= 1 && slots <= 4096, HSOLVE_ERR_INVALID, "mg: slots in [1, 4096]");
        mg->slots = (unsigned)slots;
        mg->slot.resize(slots);
        for (uint64_t i = 0; i < slots; ++i)
        {
            char * base = (char *)mg->h.alloc(mg::slot_bytes(mg->h, mg->max_level, false, false));
            mg->slot[i] = mg::slot_carve(mg->h, mg->max_level, base, false, false);
        }
        mg->table = (mg::Slot *)mg->h.alloc(sizeof(mg::Slot) * slots);
        mg->report = (mg::Report *)mg->h.alloc(sizeof(mg::Report) * slots);
        HIP_CHECK(hipHostMalloc((void **)&mg->report_host, sizeof(mg::Report) * slots, hipHostMallocDefault));
    }

    hsolve_mg_s * mg_new(const hsolve_mg_params * params)
    {
        require(params != nullptr, HSOLVE_ERR_INVALID, "mg: null params");
        std::unique_ptr<hsolve_mg_s> mg(new hsolve_mg_s());
        const mg::Params p = mg::make_params(*params);
        mg::hierarchy_alloc(mg->h, p, params->max_level);
        mg::hierarchy_alloc64(mg->h, params->max_level);
        mg->max_level = params->max_level;
        mg_slots(mg.get(), 1);
        return mg.release();
    }

    // `count` solves in the same ticks; slot i solves b[i] into x[i]
    int mg_solve_batch(hsolve_mg mg, uint64_t count, const void * const * b, void * const * x, uint64_t * outer,
                       uint64_t * inner, double * final_defect, double * initial_defect, hipStream_t st)
    {
        require(count >= 1 && count <= mg->slots, HSOLVE_ERR_INVALID, "mg_solve: count above the handle's slots");
        std::vector<mg::Slot> tab(count);
        const mg::Scalars s0 = mg::start_scalars(false);
        for (uint64_t i = 0; i < count; ++i)
        {
            require(b[i] && x[i], HSOLVE_ERR_INVALID, "mg_solve: null vector");
            tab[i] = mg->slot[i];
            tab[i].b = (const double *)b[i];
            tab[i].result = (double *)x[i];    // in: initial guess, out: solution
            HIP_CHECK(hipMemcpyAsync(tab[i].s, &s0, sizeof(s0), hipMemcpyHostToDevice, st));
        }
        HIP_CHECK(hipMemcpyAsync(mg->table, tab.data(), sizeof(mg::Slot) * count, hipMemcpyHostToDevice, st));
        const unsigned S = (unsigned)count;
        uint64_t chunk = mg->last_ticks > 2 ? mg->last_ticks : 2;
        while (true)
        {
            for (uint64_t c = 0; c < chunk; ++c)
                mg::enqueue_tick(mg->h, mg->max_level, mg->table, S, st);
            mg::enqueue_report(mg->table, S, mg->report, st);
            HIP_CHECK(hipMemcpyAsync(mg->report_host, mg->report, sizeof(mg::Report) * S, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            bool done = true;
            for (unsigned i = 0; i < S; ++i)
                done = done && mg->report_host[i].phase == mg::PH_DONE;
            if (done)
                break;
            chunk = chunk < 2 ? 2 : (chunk >= 8 ? 8 : chunk * 2);
        }
        uint64_t ticks = 0;
        int status = HSOLVE_OK;
        for (unsigned i = 0; i < S; ++i)
        {
            const mg::Report & r = mg->report_host[i];
            ticks = std::max<uint64_t>(ticks, r.ticks);
            if (outer) outer[i] = r.outer_iters;
            if (inner) inner[i] = r.inner_total;
            if (final_defect) final_defect[i] = r.def_norm;
            if (initial_defect) initial_defect[i] = r.initial_defect;
            if (r.status != 0)
                status = HSOLVE_ERR_NOT_CONVERGED;
        }
        mg->last_ticks = ticks;
        if (status != HSOLVE_OK)
            set_last_error("mg_solve: max_outer reached");
        return status;
    }
}

extern "C" {

int hsolve_mg_create(const hsolve_mg_params * params, const double * const * bands64_host,
                     const float * const * bands32_host, const float * const * dinv32_host, hsolve_mg * out)
{
    HSOLVE_API_BEGIN
    require(out && bands64_host && bands32_host, HSOLVE_ERR_INVALID, "mg_create: null pointer");
    *out = nullptr;
    std::unique_ptr<hsolve_mg_s> mg(mg_new(params));
    mg::Hierarchy & h = mg->h;
    const uint64_t L = mg->max_level;
    const uint64_t n = mg::size_of(L);
    for (int b = 0; b < 9; ++b)
    {
        require(bands64_host[b] != nullptr, HSOLVE_ERR_INVALID, "mg_create: null FP64 band");
        HIP_CHECK(hipMemcpy(h.bands64[L] + (uint64_t)b * h.bs64[L], bands64_host[b], sizeof(double) * n,
                            hipMemcpyHostToDevice));
    }
    for (uint64_t l = params->min_level; l <= L; ++l)
    {
        mg::HierLevel & lv = h.lv[l];
        for (int b = 0; b < 9; ++b)
        {
            const float * src = bands32_host[l * 9 + b];
            require(src != nullptr, HSOLVE_ERR_INVALID, "mg_create: null FP32 band");
            HIP_CHECK(hipMemcpy(lv.bands + (uint64_t)b * lv.bs, src, sizeof(float) * lv.n, hipMemcpyHostToDevice));
        }
        if (dinv32_host && dinv32_host[l])
            HIP_CHECK(hipMemcpy(lv.dinv, dinv32_host[l], sizeof(float) * lv.n, hipMemcpyHostToDevice));
        else
            mg::hierarchy_dinv(h, l, 0);
    }
    mg::hierarchy_codes(h);
    HIP_CHECK(hipDeviceSynchronize());
    *out = mg.release();
    HSOLVE_API_END
}

int hsolve_mg_create_poisson(const hsolve_mg_params * params, hsolve_mg * out)
{
    HSOLVE_API_BEGIN
    require(out != nullptr, HSOLVE_ERR_INVALID, "mg_create_poisson: null out");
    *out = nullptr;
    std::unique_ptr<hsolve_mg_s> mg(mg_new(params));
    mg::hierarchy_fill_poisson(mg->h, 0);
    HIP_CHECK(hipDeviceSynchronize());
    *out = mg.release();
    HSOLVE_API_END
}

int hsolve_mg_destroy(hsolve_mg mg)
{
    HSOLVE_API_BEGIN
    delete mg;
    HSOLVE_API_END
}

int hsolve_mg_set_value_codes(hsolve_mg mg, int enable)
{
    HSOLVE_API_BEGIN
    require(mg != nullptr, HSOLVE_ERR_INVALID, "mg_set_value_codes: null handle");
    mg->h.use_codes = enable != 0;
    HSOLVE_API_END
}

int hsolve_mg_device_bytes(hsolve_mg mg, uint64_t * bytes)
{
    HSOLVE_API_BEGIN
    require(mg && bytes, HSOLVE_ERR_INVALID, "mg_device_bytes: null pointer");
    *bytes = mg->h.bytes;
    HSOLVE_API_END
}

int hsolve_mg_solve(hsolve_mg mg, const void * b_dev, void * x_dev, uint64_t * outer_iters, uint64_t * inner_iters,
                    double * final_defect, double * initial_defect, void * stream)
{
    HSOLVE_API_BEGIN
    require(mg && b_dev && x_dev, HSOLVE_ERR_INVALID, "mg_solve: null pointer");
    const void * b[1] = {b_dev};
    void * x[1] = {x_dev};
    return mg_solve_batch(mg, 1, b, x, outer_iters, inner_iters, final_defect, initial_defect, as_stream(stream));
    HSOLVE_API_END
}

int hsolve_mg_solve_batch(hsolve_mg mg, uint64_t count, const void * const * b, void * const * x,
                          uint64_t * outer_iters, uint64_t * inner_iters, double * final_defect,
                          double * initial_defect, void * stream)
{
    HSOLVE_API_BEGIN
    require(mg && b && x, HSOLVE_ERR_INVALID, "mg_solve_batch: null pointer");
    return mg_solve_batch(mg, count, b, x, outer_iters, inner_iters, final_defect, initial_defect, as_stream(stream));
    HSOLVE_API_END
}

int hsolve_mg_create_ell(const hsolve_mg_params * params, const hsolve_ell_host * A64, const hsolve_ell_host * a32,
                         const float * const * dinv32, const hsolve_ell_host * prol32, const hsolve_ell_host * res32,
                         uint64_t slots, hsolve_mg * out)
{
    HSOLVE_API_BEGIN
    require(params && A64 && a32 && out, HSOLVE_ERR_INVALID, "mg_create_ell: null pointer");
    *out = nullptr;
    std::unique_ptr<hsolve_mg_s> mg(new hsolve_mg_s());
    static_assert(sizeof(hsolve_ell_host) == sizeof(mg::EllHost), "host ELL layout");
    mg::hierarchy_alloc_ell(mg->h, mg::make_params(*params), params->max_level, (const mg::EllHost *)a32,
                            *(const mg::EllHost *)A64, dinv32, (const mg::EllHost *)prol32,
                            (const mg::EllHost *)res32);
    mg->max_level = params->max_level;
    mg_slots(mg.get(), slots ? slots : 1);
    HIP_CHECK(hipDeviceSynchronize());
    *out = mg.release();
    HSOLVE_API_END
}

int hsolve_mg_slots(hsolve_mg mg, uint64_t * slots)
{
    HSOLVE_API_BEGIN
    require(mg && slots, HSOLVE_ERR_INVALID, "mg_slots: null pointer");
    *slots = mg->slots;
    HSOLVE_API_END
}

}  // extern "C"
