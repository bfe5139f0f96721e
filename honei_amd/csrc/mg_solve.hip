// mg_solve.hip -- C-ABI of the multigrid handle (hsolve_mg_*): one hierarchy
// of the FP32-MG / FP64-IR engine with one slot.  A solve enqueues ticks in
// chunks (first chunk sized by the previous solve's tick count) and reads
// the slot's report once per chunk.
#include "mg_engine.hh"

#include <cstring>
#include <memory>

using namespace hsolve;

struct hsolve_mg_s
{
    mg::Hierarchy h;
    uint64_t max_level = 0;
    mg::Slot slot;                      // host image of the table entry
    mg::Slot * table = nullptr;         // device, one entry
    mg::Report * report = nullptr;      // device, one entry
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
    hsolve_mg_s * mg_new(const hsolve_mg_params * params)
    {
        require(params != nullptr, HSOLVE_ERR_INVALID, "mg: null params");
        std::unique_ptr<hsolve_mg_s> mg(new hsolve_mg_s());
        const mg::Params p = mg::make_params(*params);
        mg::hierarchy_alloc(mg->h, p, params->max_level);
        mg::hierarchy_alloc64(mg->h, params->max_level);
        mg->max_level = params->max_level;
        char * base = (char *)mg->h.alloc(mg::slot_bytes(mg->h, mg->max_level, false, false));
        mg->slot = mg::slot_carve(mg->h, mg->max_level, base, false, false);
        mg->table = (mg::Slot *)mg->h.alloc(sizeof(mg::Slot));
        mg->report = (mg::Report *)mg->h.alloc(sizeof(mg::Report));
        HIP_CHECK(hipHostMalloc((void **)&mg->report_host, sizeof(mg::Report), hipHostMallocDefault));
        return mg.release();
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
    hipStream_t st = as_stream(stream);
    mg->slot.b = (const double *)b_dev;
    mg->slot.result = (double *)x_dev;     // in: initial guess, out: solution
    const mg::Slot slot = mg->slot;
    const mg::Scalars s0 = mg::start_scalars(false);
    HIP_CHECK(hipMemcpyAsync(mg->table, &slot, sizeof(slot), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(slot.s, &s0, sizeof(s0), hipMemcpyHostToDevice, st));

    mg::Report & r = *mg->report_host;
    uint64_t chunk = mg->last_ticks > 2 ? mg->last_ticks : 2;
    while (true)
    {
        for (uint64_t c = 0; c < chunk; ++c)
            mg::enqueue_tick(mg->h, mg->max_level, mg->table, 1, st);
        mg::enqueue_report(mg->table, 1, mg->report, st);
        HIP_CHECK(hipMemcpyAsync(mg->report_host, mg->report, sizeof(mg::Report), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (r.phase == mg::PH_DONE)
            break;
        chunk = chunk < 2 ? 2 : (chunk >= 8 ? 8 : chunk * 2);
    }
    mg->last_ticks = r.ticks;
    if (outer_iters) *outer_iters = r.outer_iters;
    if (inner_iters) *inner_iters = r.inner_total;
    if (final_defect) *final_defect = r.def_norm;
    if (initial_defect) *initial_defect = r.initial_defect;
    if (r.status != 0)
    {
        set_last_error("mg_solve: max_outer reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

}  // extern "C"
