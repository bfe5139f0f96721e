// ops.hip -- C-ABI of libhsolve: runtime helpers and the op launchers that
// mirror honei/backends/cuda/operations.hh, on gfx950 kernels.
#include "kernels_mg.hh"

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace hsolve
{
    static thread_local std::string g_last_error;

    void set_last_error(const std::string & msg) { g_last_error = msg; }

    // --------------------------------------------------------- scratch
    namespace
    {
        std::mutex g_scratch_mutex;
        std::map<std::pair<int, hipStream_t>, ReduceScratch> g_reduce;
    }

    ReduceScratch & reduce_scratch(hipStream_t stream)
    {
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        auto key = std::make_pair(dev, stream);
        auto it = g_reduce.find(key);
        if (it != g_reduce.end())
            return it->second;
        ReduceScratch s;
        void * p = nullptr;
        HIP_CHECK(hipMalloc(&p, sizeof(double) * kReduceBlocks));
        s.partials = (double *)p;
        return g_reduce.emplace(key, s).first->second;
    }

    // -------------------------------------------------------- launchers
    template <typename DT>
    void launch_dot(const DT * a, const DT * b, uint64_t n, DT * out, int root, hipStream_t st)
    {
        ReduceScratch & s = reduce_scratch(st);
        unsigned grid = grid_for(n, kBlock * 4);
        if (grid > (unsigned)kReduceBlocks)
            grid = kReduceBlocks;
        hipLaunchKernelGGL((k::dot_kernel<DT>), dim3(grid), dim3(kBlock), 0, st, a, b, n, s.partials);
        hipLaunchKernelGGL((k::dot_finish_kernel<DT>), dim3(1), dim3(kBlock), 0, st, (const double *)s.partials, grid,
                           out, root);
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT, int OP>
    void launch_elem(DT * r, const DT * x, const DT * y, const DT * z, DT b, uint64_t n, hipStream_t st)
    {
        if (n == 0)
            return;
        unsigned grid = grid_for(n, kBlock);
        if (grid > 16384)
            grid = 16384;
        hipLaunchKernelGGL((k::elem_kernel<DT, OP>), dim3(grid), dim3(kBlock), 0, st, r, x, y, z, b, n);
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT>
    void launch_q1(int mode, const k::Bands & B, const DT * x, DT * y, const DT * rhs, const DT * dinv,
                   uint64_t n, hipStream_t st)
    {
        if (n == 0)
            return;
        const int64_t m = (int64_t)std::sqrt((double)n);
        require((uint64_t)(m * m) == n, HSOLVE_ERR_SIZE, "Q1 size is not a square");
        if (mode == 0)
            k::q1_launch<DT, 0>(B, x, y, rhs, dinv, (int64_t)n, m, nullptr, st);
        else if (mode == 1)
            k::q1_launch<DT, 1>(B, x, y, rhs, dinv, (int64_t)n, m, nullptr, st);
        else
            k::q1_launch<DT, 2>(B, x, y, rhs, dinv, (int64_t)n, m, nullptr, st);
        HIP_CHECK(hipGetLastError());
    }

    // ELL launcher over any index / row-length type.  MAXK picks the
    // unrolled variant covering num_cols_per_row.
    template <typename DT, typename IT, typename LT, bool DEFECT>
    void launch_ell(const DT * x, DT * y, const IT * Aj, const DT * Ax, const LT * Arl, const DT * rhs,
                    uint64_t row_start, uint64_t row_end, uint64_t ncpr, uint64_t stride, uint64_t threads,
                    hipStream_t st)
    {
        if (row_end <= row_start)
            return;
        constexpr int MODE = DEFECT ? 1 : 0;
        const DT * dinv = nullptr;
        const unsigned grid = grid_for(row_end - row_start, kBlock);
        if (threads != 1)
        {
            hipLaunchKernelGGL((k::ell_spmv_tn<DT, IT, LT, MODE>), dim3(grid), dim3(kBlock), 0, st, x, y,
                               Aj, Ax, Arl, rhs, dinv, row_start, row_end, stride, (uint32_t)threads);
        }
#define HS_ELL(K)                                                                                  \
        else if (ncpr <= K)                                                                        \
            hipLaunchKernelGGL((k::ell_spmv_t1<DT, IT, LT, K, MODE>), dim3(grid), dim3(kBlock), 0, st,   \
                               x, y, Aj, Ax, Arl, rhs, dinv, row_start, row_end, stride);
        HS_ELL(4) HS_ELL(8) HS_ELL(9) HS_ELL(12) HS_ELL(16) HS_ELL(24) HS_ELL(32)
#undef HS_ELL
        else
            hipLaunchKernelGGL((k::ell_spmv_t1<DT, IT, LT, 0, MODE>), dim3(grid), dim3(kBlock), 0, st, x, y,
                               Aj, Ax, Arl, rhs, dinv, row_start, row_end, stride);
        HIP_CHECK(hipGetLastError());
    }

    static k::Bands bands_of(const void * ll, const void * ld, const void * lu, const void * dl, const void * dd,
                             const void * du, const void * ul, const void * ud, const void * uu)
    {
        k::Bands B;
        B.b[0] = ll; B.b[1] = ld; B.b[2] = lu; B.b[3] = dl; B.b[4] = dd;
        B.b[5] = du; B.b[6] = ul; B.b[7] = ud; B.b[8] = uu;
        for (int i = 0; i < 9; ++i)
            require(B.b[i] != nullptr, HSOLVE_ERR_INVALID, "null band pointer");
        return B;
    }

    static k::Mask mask_of(const uint64_t * m)
    {
        require(m != nullptr, HSOLVE_ERR_INVALID, "null macro border mask");
        k::Mask mk;
        for (int i = 0; i < 8; ++i)
            mk.dir[i] = (m[i] == 2) ? 1 : 0;
        return mk;
    }

    static int64_t grid_side(uint64_t n, const char * what)
    {
        const int64_t M = (int64_t)std::sqrt((double)n);
        require(M >= 2 && (uint64_t)(M * M) == n && ((M - 1) & (M - 2)) == 0, HSOLVE_ERR_SIZE,
                std::string(what) + ": size is not (2^L+1)^2");
        // the transfer kernels put grid rows on gridDim.y (at most 65535)
        require(M <= 32769, HSOLVE_ERR_SIZE, std::string(what) + ": grid side above 2^15+1");
        return M;
    }

    template <typename DT>
    void launch_restriction(DT * coarse, uint64_t nc, const DT * fine, uint64_t nf, const uint64_t * mask,
                            hipStream_t st)
    {
        const int64_t Mc = grid_side(nc, "restriction coarse"), Mf = grid_side(nf, "restriction fine");
        require(Mf == 2 * Mc - 1, HSOLVE_ERR_SIZE, "restriction: fine side != 2*coarse side - 1");
        hipLaunchKernelGGL((k::restriction_kernel<DT>), dim3(grid_for(Mc, kBlock), Mc), dim3(kBlock), 0, st, coarse,
                           fine, Mc, mask_of(mask), (const int *)nullptr);
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT, bool AXPY>
    void launch_prolongation(DT * fine, uint64_t nf, const DT * coarse, uint64_t nc, DT alpha,
                             const uint64_t * mask, hipStream_t st)
    {
        const int64_t Mc = grid_side(nc, "prolongation coarse"), Mf = grid_side(nf, "prolongation fine");
        require(Mf == 2 * Mc - 1, HSOLVE_ERR_SIZE, "prolongation: fine side != 2*coarse side - 1");
        hipLaunchKernelGGL((k::prolongation2_kernel<DT, AXPY>), dim3(grid_for(Mc, kBlock), Mf), dim3(kBlock), 0, st,
                           fine, coarse, (int)Mf, alpha, mask_of(mask), (const int *)nullptr);
        HIP_CHECK(hipGetLastError());
    }

    // per-stream grow-only workspace for entry points that need scratch
    namespace
    {
        struct Work
        {
            void * p = nullptr;
            size_t bytes = 0;
        };
        std::map<std::pair<int, hipStream_t>, Work> g_work;
    }

    void * stream_work(hipStream_t st, size_t bytes)
    {
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        Work & w = g_work[std::make_pair(dev, st)];
        if (w.bytes < bytes)
        {
            if (w.p)
            {
                HIP_CHECK(hipStreamSynchronize(st));
                HIP_CHECK(hipFree(w.p));
                w.p = nullptr;
                w.bytes = 0;
            }
            HIP_CHECK(hipMalloc(&w.p, bytes));
            w.bytes = bytes;
        }
        return w.p;
    }

    template <typename DT>
    void launch_cg_konv(const k::Bands & B, const DT * rhs, DT * x, DT * work, uint64_t n, double konv_rad,
                        uint64_t max_iters, uint64_t * iters_dev, const int * active, hipStream_t st)
    {
        const int64_t m = (int64_t)std::sqrt((double)n);
        require((uint64_t)(m * m) == n, HSOLVE_ERR_SIZE, "coarse CG: size is not a square");
        require(n <= 16641, HSOLVE_ERR_SIZE, "coarse CG: size above 16641 (129^2)");
        hipLaunchKernelGGL((k::cg_konv_q1_kernel<DT>), dim3(1), dim3(kBlock), 0, st, B, rhs, x, work, (int64_t)n, m,
                           konv_rad, max_iters, iters_dev, active);
        HIP_CHECK(hipGetLastError());
    }

    // one workgroup with sequential sums: up to 2^20 rows
    template <typename DT>
    void launch_cg_iters(const k::Bands & B, const DT * rhs, DT * x, DT * work, uint64_t n, uint64_t iters,
                         uint64_t * used_dev, hipStream_t st)
    {
        const int64_t m = (int64_t)std::sqrt((double)n);
        require((uint64_t)(m * m) == n && n > 0, HSOLVE_ERR_SIZE, "cg_iters_q1: size is not a square");
        require(n <= (1u << 20), HSOLVE_ERR_SIZE, "cg_iters_q1: more than 2^20 rows for one workgroup");
        hipLaunchKernelGGL((k::cg_iters_q1_kernel<DT>), dim3(1), dim3(kBlock), 0, st, B, rhs, x, work, (int64_t)n, m,
                           iters, used_dev);
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT>
    void fill_poisson_q1(void * const * bands9, uint64_t size, hipStream_t st)
    {
        require(bands9 != nullptr, HSOLVE_ERR_INVALID, "fill_poisson_q1: null band array");
        k::Bands B;
        for (int i = 0; i < 9; ++i)
        {
            require(bands9[i] != nullptr, HSOLVE_ERR_INVALID, "fill_poisson_q1: null band pointer");
            B.b[i] = bands9[i];
        }
        const int64_t m = (int64_t)std::sqrt((double)size);
        require((uint64_t)(m * m) == size, HSOLVE_ERR_SIZE, "fill_poisson_q1: size is not a square");
        if (size == 0)
            return;
        hipLaunchKernelGGL((k::fill_poisson_kernel<DT>), dim3(grid_for(size, kBlock)), dim3(kBlock), 0, st, B,
                           (int64_t)size, m);
        HIP_CHECK(hipGetLastError());
    }

    template void launch_dot<float>(const float *, const float *, uint64_t, float *, int, hipStream_t);
    template void launch_dot<double>(const double *, const double *, uint64_t, double *, int, hipStream_t);
}

using namespace hsolve;

#define ST as_stream(stream)

extern "C" {

const char * hsolve_last_error(void) { return g_last_error.c_str(); }
const char * hsolve_version(void) { return "hsolve 0.1 gfx950"; }

// ------------------------------------------------------------------ runtime
int hsolve_device_count(int * count)
{
    HSOLVE_API_BEGIN
    require(count != nullptr, HSOLVE_ERR_INVALID, "null count");
    HIP_CHECK(hipGetDeviceCount(count));
    HSOLVE_API_END
}

int hsolve_get_device(int * device)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipGetDevice(device));
    HSOLVE_API_END
}

int hsolve_set_device(int device)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipSetDevice(device));
    HSOLVE_API_END
}

int hsolve_malloc(void ** ptr, uint64_t bytes)
{
    HSOLVE_API_BEGIN
    require(ptr != nullptr, HSOLVE_ERR_INVALID, "null ptr");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess)
        throw Error(HSOLVE_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " +
                                          hipGetErrorString(e));
    HSOLVE_API_END
}

int hsolve_malloc_host(void ** ptr, uint64_t bytes)
{
    HSOLVE_API_BEGIN
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess)
        throw Error(HSOLVE_ERR_ALLOC, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    HSOLVE_API_END
}

int hsolve_free(void * ptr)
{
    HSOLVE_API_BEGIN
    if (ptr)
        HIP_CHECK(hipFree(ptr));
    HSOLVE_API_END
}

int hsolve_free_host(void * ptr)
{
    HSOLVE_API_BEGIN
    if (ptr)
        HIP_CHECK(hipHostFree(ptr));
    HSOLVE_API_END
}

int hsolve_upload(const void * src_host, void * dst, uint64_t bytes, void * stream)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipMemcpyAsync(dst, src_host, bytes, hipMemcpyHostToDevice, ST));
    HIP_CHECK(hipStreamSynchronize(ST));
    HSOLVE_API_END
}

int hsolve_download(const void * src, void * dst_host, uint64_t bytes, void * stream)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipMemcpyAsync(dst_host, src, bytes, hipMemcpyDeviceToHost, ST));
    HIP_CHECK(hipStreamSynchronize(ST));
    HSOLVE_API_END
}

int hsolve_copy(const void * src, void * dst, uint64_t bytes, void * stream)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ST));
    HSOLVE_API_END
}

int hsolve_fill_zero(void * dst, uint64_t bytes, void * stream)
{
    HSOLVE_API_BEGIN
    if (bytes)
    {
        unsigned grid = grid_for(bytes, kBlock);
        if (grid > 16384)
            grid = 16384;
        hipLaunchKernelGGL(k::zero_bytes_kernel, dim3(grid), dim3(kBlock), 0, ST, (uint8_t *)dst, bytes);
        HIP_CHECK(hipGetLastError());
    }
    HSOLVE_API_END
}

int hsolve_fill_float(void * dst, float value, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    launch_elem<float, k::E_FILL>((float *)dst, nullptr, nullptr, nullptr, value, size, ST);
    HSOLVE_API_END
}

int hsolve_fill_double(void * dst, double value, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    launch_elem<double, k::E_FILL>((double *)dst, nullptr, nullptr, nullptr, value, size, ST);
    HSOLVE_API_END
}

int hsolve_synchronize(void * stream)
{
    HSOLVE_API_BEGIN
    if (stream)
        HIP_CHECK(hipStreamSynchronize(ST));
    else
        HIP_CHECK(hipDeviceSynchronize());
    HSOLVE_API_END
}

int hsolve_stream_create(void ** stream)
{
    HSOLVE_API_BEGIN
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    HSOLVE_API_END
}

int hsolve_stream_destroy(void * stream)
{
    HSOLVE_API_BEGIN
    if (stream)
        HIP_CHECK(hipStreamDestroy(ST));
    HSOLVE_API_END
}

int hsolve_event_create(void ** event)
{
    HSOLVE_API_BEGIN
    require(event != nullptr, HSOLVE_ERR_INVALID, "null event");
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    *event = (void *)e;
    HSOLVE_API_END
}

int hsolve_event_destroy(void * event)
{
    HSOLVE_API_BEGIN
    if (event)
        HIP_CHECK(hipEventDestroy((hipEvent_t)event));
    HSOLVE_API_END
}

int hsolve_event_record(void * event, void * stream)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipEventRecord((hipEvent_t)event, ST));
    HSOLVE_API_END
}

int hsolve_event_elapsed_ms(void * start, void * stop, float * ms)
{
    HSOLVE_API_BEGIN
    HIP_CHECK(hipEventSynchronize((hipEvent_t)stop));
    HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    HSOLVE_API_END
}

int hsolve_copy_kernel(const void * src, void * dst, uint64_t bytes, void * stream)
{
    HSOLVE_API_BEGIN
    require(src && dst && (bytes % 16) == 0, HSOLVE_ERR_INVALID, "copy_kernel: null pointer or bytes % 16 != 0");
    const uint64_t n = bytes / 16;
    if (n)
    {
        unsigned grid = grid_for(n, kBlock);
        if (grid > 65536)
            grid = 65536;
        hipLaunchKernelGGL(k::copy16_kernel, dim3(grid), dim3(kBlock), 0, ST, (const uint4 *)src, (uint4 *)dst, n);
        HIP_CHECK(hipGetLastError());
    }
    HSOLVE_API_END
}

// ------------------------------------------------------------- ELL (raw)
#define HS_ELL_RAW(NAME, DT, DEFECT, RHS)                                                             \
    {                                                                                                 \
        HSOLVE_API_BEGIN                                                                              \
        require(x && y && Aj && Ax && Arl, HSOLVE_ERR_INVALID, NAME ": null pointer");                \
        require(threads >= 1 && stride >= 1, HSOLVE_ERR_INVALID, NAME ": bad layout");                \
        launch_ell<DT, uint64_t, uint64_t, DEFECT>((const DT *)x, (DT *)y, (const uint64_t *)Aj,      \
                (const DT *)Ax, (const uint64_t *)Arl, (const DT *)(RHS), row_start, row_end,         \
                num_cols_per_row, stride, threads, ST);                                               \
        HSOLVE_API_END                                                                                \
    }

int hsolve_product_smell_dv_float(const void * x, void * y, const void * Aj, const void * Ax, const void * Arl,
                                  uint64_t row_start, uint64_t row_end, uint64_t num_cols_per_row, uint64_t stride,
                                  uint64_t threads, void * stream)
HS_ELL_RAW("product_smell_dv_float", float, false, nullptr)

int hsolve_product_smell_dv_double(const void * x, void * y, const void * Aj, const void * Ax, const void * Arl,
                                   uint64_t row_start, uint64_t row_end, uint64_t num_cols_per_row,
                                   uint64_t stride, uint64_t threads, void * stream)
HS_ELL_RAW("product_smell_dv_double", double, false, nullptr)

int hsolve_defect_smell_dv_float(const void * rhs, void * result, const void * Aj, const void * Ax,
                                 const void * Arl, const void * b, uint64_t row_start, uint64_t row_end,
                                 uint64_t num_cols_per_row, uint64_t stride, uint64_t threads, void * stream)
{
    const void * x = b;
    void * y = result;
    HS_ELL_RAW("defect_smell_dv_float", float, true, rhs)
}

int hsolve_defect_smell_dv_double(const void * rhs, void * result, const void * Aj, const void * Ax,
                                  const void * Arl, const void * b, uint64_t row_start, uint64_t row_end,
                                  uint64_t num_cols_per_row, uint64_t stride, uint64_t threads, void * stream)
{
    const void * x = b;
    void * y = result;
    HS_ELL_RAW("defect_smell_dv_double", double, true, rhs)
}
#undef HS_ELL_RAW

// ------------------------------------------------------------------- Q1
#define HS_Q1(DT, MODE, RHS)                                                                          \
    HSOLVE_API_BEGIN                                                                                  \
    require(x && y && (MODE == 0 || (RHS) != nullptr), HSOLVE_ERR_INVALID, "q1: null vector");        \
    require(m * m == size, HSOLVE_ERR_SIZE, "q1: m*m != size");                                       \
    launch_q1<DT>(MODE, bands_of(ll, ld, lu, dl, dd, du, ul, ud, uu), (const DT *)x, (DT *)y,         \
                  (const DT *)(RHS), nullptr, size, ST);                                              \
    HSOLVE_API_END

int hsolve_product_bmdv_q1_float(const void * ll, const void * ld, const void * lu, const void * dl,
                                 const void * dd, const void * du, const void * ul, const void * ud,
                                 const void * uu, const void * x, void * y, uint64_t size, uint64_t m,
                                 void * stream)
{
    HS_Q1(float, 0, nullptr)
}

int hsolve_product_bmdv_q1_double(const void * ll, const void * ld, const void * lu, const void * dl,
                                  const void * dd, const void * du, const void * ul, const void * ud,
                                  const void * uu, const void * x, void * y, uint64_t size, uint64_t m,
                                  void * stream)
{
    HS_Q1(double, 0, nullptr)
}

int hsolve_defect_q1_float(const void * rhs, const void * ll, const void * ld, const void * lu, const void * dl,
                           const void * dd, const void * du, const void * ul, const void * ud, const void * uu,
                           const void * x, void * y, uint64_t size, uint64_t m, void * stream)
{
    HS_Q1(float, 1, rhs)
}

int hsolve_defect_q1_double(const void * rhs, const void * ll, const void * ld, const void * lu,
                            const void * dl, const void * dd, const void * du, const void * ul, const void * ud,
                            const void * uu, const void * x, void * y, uint64_t size, uint64_t m, void * stream)
{
    HS_Q1(double, 1, rhs)
}
#undef HS_Q1

// ------------------------------------------------------------ reductions
#define HS_DOT(DT, A, B, ROOT)                                                                        \
    HSOLVE_API_BEGIN                                                                                  \
    require((A) && result_dev, HSOLVE_ERR_INVALID, "reduction: null pointer");                        \
    launch_dot<DT>((const DT *)(A), (const DT *)(B), size, (DT *)result_dev, ROOT, ST);               \
    HSOLVE_API_END

int hsolve_dot_product_two_float(const void * a, const void * b, uint64_t size, void * result_dev, void * stream)
{
    HS_DOT(float, a, b, 0)
}
int hsolve_dot_product_two_double(const void * a, const void * b, uint64_t size, void * result_dev, void * stream)
{
    HS_DOT(double, a, b, 0)
}
int hsolve_norm_l2_one_float(const void * a, uint64_t size, int root, void * result_dev, void * stream)
{
    HS_DOT(float, a, nullptr, root)
}
int hsolve_norm_l2_one_double(const void * a, uint64_t size, int root, void * result_dev, void * stream)
{
    HS_DOT(double, a, nullptr, root)
}
#undef HS_DOT

#define HS_DOT_HOST(DT, A, B, ROOT)                                                                   \
    HSOLVE_API_BEGIN                                                                                  \
    require((A) && result, HSOLVE_ERR_INVALID, "reduction: null pointer");                            \
    DT * dev = (DT *)stream_work(ST, 256);                                                            \
    launch_dot<DT>((const DT *)(A), (const DT *)(B), size, dev, ROOT, ST);                            \
    HIP_CHECK(hipMemcpyAsync(result, dev, sizeof(DT), hipMemcpyDeviceToHost, ST));                    \
    HIP_CHECK(hipStreamSynchronize(ST));                                                              \
    HSOLVE_API_END

int hsolve_dot_product_two_float_host(const void * a, const void * b, uint64_t size, float * result,
                                      void * stream)
{
    HS_DOT_HOST(float, a, b, 0)
}
int hsolve_dot_product_two_double_host(const void * a, const void * b, uint64_t size, double * result,
                                       void * stream)
{
    HS_DOT_HOST(double, a, b, 0)
}
int hsolve_norm_l2_one_float_host(const void * a, uint64_t size, int root, float * result, void * stream)
{
    HS_DOT_HOST(float, a, nullptr, root)
}
int hsolve_norm_l2_one_double_host(const void * a, uint64_t size, int root, double * result, void * stream)
{
    HS_DOT_HOST(double, a, nullptr, root)
}
#undef HS_DOT_HOST

// -------------------------------------------------------------- BLAS-1
#define HS_ELEM(DT, OP, R, X, Y, Z, B)                                                                \
    HSOLVE_API_BEGIN                                                                                  \
    require((R) != nullptr, HSOLVE_ERR_INVALID, "elementwise: null pointer");                         \
    launch_elem<DT, OP>((DT *)(R), (const DT *)(X), (const DT *)(Y), (const DT *)(Z), (DT)(B), size, ST); \
    HSOLVE_API_END

int hsolve_scaled_sum_three_float_s(void * r, const void * x, const void * y, float b, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_SCALED_SUM_S, r, x, y, nullptr, b)
}
int hsolve_scaled_sum_three_double_s(void * r, const void * x, const void * y, double b, uint64_t size,
                                     void * stream)
{
    HS_ELEM(double, k::E_SCALED_SUM_S, r, x, y, nullptr, b)
}
int hsolve_scaled_sum_three_float(void * x, const void * y, const void * z, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_SCALED_SUM_V, x, nullptr, y, z, 0)
}
int hsolve_scaled_sum_three_double(void * x, const void * y, const void * z, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_SCALED_SUM_V, x, nullptr, y, z, 0)
}
int hsolve_sum_two_float(void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_SUM, a, nullptr, b, nullptr, 0)
}
int hsolve_sum_two_double(void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_SUM, a, nullptr, b, nullptr, 0)
}
int hsolve_scale_one_float(void * x, float a, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_SCALE, x, nullptr, nullptr, nullptr, a)
}
int hsolve_scale_one_double(void * x, double a, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_SCALE, x, nullptr, nullptr, nullptr, a)
}
int hsolve_difference_two_float(void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_DIFF2, a, nullptr, b, nullptr, 0)
}
int hsolve_difference_two_double(void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_DIFF2, a, nullptr, b, nullptr, 0)
}
int hsolve_difference_three_float(void * r, const void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_DIFF3, r, a, b, nullptr, 0)
}
int hsolve_difference_three_double(void * r, const void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_DIFF3, r, a, b, nullptr, 0)
}
int hsolve_element_product_three_float(void * r, const void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_EPROD3, r, a, b, nullptr, 0)
}
int hsolve_element_product_three_double(void * r, const void * a, const void * b, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_EPROD3, r, a, b, nullptr, 0)
}
int hsolve_element_inverse_one_float(void * x, uint64_t size, void * stream)
{
    HS_ELEM(float, k::E_EINV, x, nullptr, nullptr, nullptr, 0)
}
int hsolve_element_inverse_one_double(void * x, uint64_t size, void * stream)
{
    HS_ELEM(double, k::E_EINV, x, nullptr, nullptr, nullptr, 0)
}
#undef HS_ELEM

int hsolve_convert_float_double(const void * src_float, void * dst_double, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    if (size)
    {
        unsigned grid = grid_for(size, kBlock);
        if (grid > 16384) grid = 16384;
        hipLaunchKernelGGL((k::convert_kernel<float, double>), dim3(grid), dim3(kBlock), 0, ST,
                           (const float *)src_float, (double *)dst_double, size);
        HIP_CHECK(hipGetLastError());
    }
    HSOLVE_API_END
}

int hsolve_convert_double_float(const void * src_double, void * dst_float, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    if (size)
    {
        unsigned grid = grid_for(size, kBlock);
        if (grid > 16384) grid = 16384;
        hipLaunchKernelGGL((k::convert_kernel<double, float>), dim3(grid), dim3(kBlock), 0, ST,
                           (const double *)src_double, (float *)dst_float, size);
        HIP_CHECK(hipGetLastError());
    }
    HSOLVE_API_END
}

// ------------------------------------------------------------- transfers
int hsolve_prolongation_float(void * fine, uint64_t size_fine, const void * coarse, uint64_t size_coarse,
                              const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    launch_prolongation<float, false>((float *)fine, size_fine, (const float *)coarse, size_coarse, 0.f, mask, ST);
    HSOLVE_API_END
}
int hsolve_prolongation_double(void * fine, uint64_t size_fine, const void * coarse, uint64_t size_coarse,
                               const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    launch_prolongation<double, false>((double *)fine, size_fine, (const double *)coarse, size_coarse, 0.0, mask,
                                       ST);
    HSOLVE_API_END
}
int hsolve_prolongation_axpy_float(void * fine_x, uint64_t size_fine, const void * coarse, uint64_t size_coarse,
                                   float alpha, const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    launch_prolongation<float, true>((float *)fine_x, size_fine, (const float *)coarse, size_coarse, alpha, mask,
                                     ST);
    HSOLVE_API_END
}
int hsolve_prolongation_axpy_double(void * fine_x, uint64_t size_fine, const void * coarse, uint64_t size_coarse,
                                    double alpha, const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    launch_prolongation<double, true>((double *)fine_x, size_fine, (const double *)coarse, size_coarse, alpha,
                                      mask, ST);
    HSOLVE_API_END
}
int hsolve_restriction_float(void * coarse, uint64_t size_coarse, const void * fine, uint64_t size_fine,
                             const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    launch_restriction<float>((float *)coarse, size_coarse, (const float *)fine, size_fine, mask, ST);
    HSOLVE_API_END
}
int hsolve_restriction_double(void * coarse, uint64_t size_coarse, const void * fine, uint64_t size_fine,
                              const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    launch_restriction<double>((double *)coarse, size_coarse, (const double *)fine, size_fine, mask, ST);
    HSOLVE_API_END
}

int hsolve_apply_dirichlet_float(void * grid, uint64_t size, const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    require(grid != nullptr, HSOLVE_ERR_INVALID, "apply_dirichlet: null grid");
    const int64_t M = grid_side(size, "apply_dirichlet");
    hipLaunchKernelGGL((k::dirichlet_kernel<float>), dim3(grid_for(size, kBlock)), dim3(kBlock), 0, ST, (float *)grid,
                       M, mask_of(mask));
    HIP_CHECK(hipGetLastError());
    HSOLVE_API_END
}

int hsolve_apply_dirichlet_double(void * grid, uint64_t size, const uint64_t * mask, void * stream)
{
    HSOLVE_API_BEGIN
    require(grid != nullptr, HSOLVE_ERR_INVALID, "apply_dirichlet: null grid");
    const int64_t M = grid_side(size, "apply_dirichlet");
    hipLaunchKernelGGL((k::dirichlet_kernel<double>), dim3(grid_for(size, kBlock)), dim3(kBlock), 0, ST,
                       (double *)grid, M, mask_of(mask));
    HIP_CHECK(hipGetLastError());
    HSOLVE_API_END
}

// ------------------------------------------------------- fused smoothers
int hsolve_fill_poisson_q1_float(void * const * bands9, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    fill_poisson_q1<float>(bands9, size, ST);
    HSOLVE_API_END
}

int hsolve_fill_poisson_q1_double(void * const * bands9, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    fill_poisson_q1<double>(bands9, size, ST);
    HSOLVE_API_END
}

int hsolve_jacobi_q1_float(const void * rhs, const void * const * bands9, const void * dinv, const void * x_old,
                           void * x_new, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    require(bands9 && rhs && dinv && x_old && x_new, HSOLVE_ERR_INVALID, "jacobi_q1: null pointer");
    require(x_old != x_new, HSOLVE_ERR_INVALID, "jacobi_q1: x_new must not alias x_old");
    launch_q1<float>(2, bands_of(bands9[0], bands9[1], bands9[2], bands9[3], bands9[4], bands9[5], bands9[6],
                                 bands9[7], bands9[8]),
                     (const float *)x_old, (float *)x_new, (const float *)rhs, (const float *)dinv, size, ST);
    HSOLVE_API_END
}
int hsolve_jacobi_q1_double(const void * rhs, const void * const * bands9, const void * dinv, const void * x_old,
                            void * x_new, uint64_t size, void * stream)
{
    HSOLVE_API_BEGIN
    require(bands9 && rhs && dinv && x_old && x_new, HSOLVE_ERR_INVALID, "jacobi_q1: null pointer");
    require(x_old != x_new, HSOLVE_ERR_INVALID, "jacobi_q1: x_new must not alias x_old");
    launch_q1<double>(2, bands_of(bands9[0], bands9[1], bands9[2], bands9[3], bands9[4], bands9[5], bands9[6],
                                  bands9[7], bands9[8]),
                      (const double *)x_old, (double *)x_new, (const double *)rhs, (const double *)dinv, size, ST);
    HSOLVE_API_END
}

int hsolve_cg_konv_q1_float(const void * const * bands9, const void * rhs, void * x, uint64_t size, double konv_rad,
                            uint64_t max_iters, void * iters_dev, void * stream)
{
    HSOLVE_API_BEGIN
    require(bands9 && rhs && x, HSOLVE_ERR_INVALID, "cg_konv: null pointer");
    float * work = (float *)stream_work(ST, sizeof(float) * 4 * size + 256);
    launch_cg_konv<float>(bands_of(bands9[0], bands9[1], bands9[2], bands9[3], bands9[4], bands9[5], bands9[6],
                                   bands9[7], bands9[8]),
                          (const float *)rhs, (float *)x, work, size, konv_rad,
                          max_iters ? max_iters : 100 * size + 1000, (uint64_t *)iters_dev, nullptr, ST);
    HSOLVE_API_END
}
int hsolve_cg_konv_q1_double(const void * const * bands9, const void * rhs, void * x, uint64_t size,
                             double konv_rad, uint64_t max_iters, void * iters_dev, void * stream)
{
    HSOLVE_API_BEGIN
    require(bands9 && rhs && x, HSOLVE_ERR_INVALID, "cg_konv: null pointer");
    double * work = (double *)stream_work(ST, sizeof(double) * 4 * size + 256);
    launch_cg_konv<double>(bands_of(bands9[0], bands9[1], bands9[2], bands9[3], bands9[4], bands9[5], bands9[6],
                                    bands9[7], bands9[8]),
                           (const double *)rhs, (double *)x, work, size, konv_rad,
                           max_iters ? max_iters : 100 * size + 1000, (uint64_t *)iters_dev, nullptr, ST);
    HSOLVE_API_END
}

int hsolve_cg_iters_q1_float(const void * const * bands9, const void * rhs, void * x, uint64_t size,
                             uint64_t iters, void * used_dev, void * stream)
{
    HSOLVE_API_BEGIN
    require(bands9 && rhs && x, HSOLVE_ERR_INVALID, "cg_iters_q1: null pointer");
    float * work = (float *)stream_work(ST, sizeof(float) * 5 * size + 256);
    launch_cg_iters<float>(bands_of(bands9[0], bands9[1], bands9[2], bands9[3], bands9[4], bands9[5], bands9[6],
                                    bands9[7], bands9[8]),
                           (const float *)rhs, (float *)x, work, size, iters, (uint64_t *)used_dev, ST);
    HSOLVE_API_END
}
int hsolve_cg_iters_q1_double(const void * const * bands9, const void * rhs, void * x, uint64_t size,
                              uint64_t iters, void * used_dev, void * stream)
{
    HSOLVE_API_BEGIN
    require(bands9 && rhs && x, HSOLVE_ERR_INVALID, "cg_iters_q1: null pointer");
    double * work = (double *)stream_work(ST, sizeof(double) * 5 * size + 256);
    launch_cg_iters<double>(bands_of(bands9[0], bands9[1], bands9[2], bands9[3], bands9[4], bands9[5], bands9[6],
                                     bands9[7], bands9[8]),
                            (const double *)rhs, (double *)x, work, size, iters, (uint64_t *)used_dev, ST);
    HSOLVE_API_END
}

}  // extern "C"
