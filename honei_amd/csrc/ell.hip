// ell.hip -- device-resident SparseMatrixELL handles.
//
// The host layout (HONEI's SMELL: uint64 column indices, uint64 row step
// counts, column-major slots of `stride` rows) is converted once at upload
// into the backend layout: the same slot order with uint32 column indices
// and the row lengths in the narrowest unsigned type that holds the longest
// row.  That halves the index stream (12 instead of 16 bytes per FP64
// nonzero) and makes the row-length stream 1 byte per row for ordinary
// stencils, without changing which products are summed or their order.
#include "kernels_la.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hsolve
{
    void * stream_work(hipStream_t st, size_t bytes);
    template <typename DT>
    void launch_dot(const DT * a, const DT * b, uint64_t n, DT * out, int root, hipStream_t st);
}

struct hsolve_ell_s
{
    int dtype;                 // 0 float, 1 double
    uint64_t rows, cols, stride, ncpr, threads, nnz;
    int len_bytes;             // 1, 2 or 4
    uint32_t * Aj;             // device, ncpr * stride
    void * Ax;                 // device, ncpr * stride values
    void * Arl;                // device, rows lengths
    uint64_t device_bytes;
};

using namespace hsolve;

namespace
{
    template <typename DT, typename LT, bool DEFECT>
    void ell_launch(const hsolve_ell_s * m, const DT * x, DT * y, const DT * rhs, hipStream_t st)
    {
        const unsigned grid = grid_for(m->rows, kBlock);
        const uint32_t * Aj = m->Aj;
        const DT * Ax = (const DT *)m->Ax;
        const LT * Arl = (const LT *)m->Arl;
        if (m->threads != 1)
            hipLaunchKernelGGL((k::ell_spmv_tn<DT, uint32_t, LT, DEFECT>), dim3(grid), dim3(kBlock), 0, st, x, y,
                               Aj, Ax, Arl, rhs, (uint64_t)0, m->rows, m->stride, (uint32_t)m->threads);
#define HS_K(K)                                                                                               \
        else if (m->ncpr <= K)                                                                                \
            hipLaunchKernelGGL((k::ell_spmv_t1<DT, uint32_t, LT, K, DEFECT>), dim3(grid), dim3(kBlock), 0, st, x, \
                               y, Aj, Ax, Arl, rhs, (uint64_t)0, m->rows, m->stride);
        HS_K(4) HS_K(8) HS_K(9) HS_K(12) HS_K(13) HS_K(16) HS_K(24) HS_K(32)
#undef HS_K
        else
            hipLaunchKernelGGL((k::ell_spmv_t1<DT, uint32_t, LT, 0, DEFECT>), dim3(grid), dim3(kBlock), 0, st, x, y,
                               Aj, Ax, Arl, rhs, (uint64_t)0, m->rows, m->stride);
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT, typename LT>
    void ell_jacobi_launch(const hsolve_ell_s * m, const DT * rhs, const DT * dinv, const DT * x_old, DT * x_new,
                           hipStream_t st)
    {
        const unsigned grid = grid_for(m->rows, kBlock);
        const uint32_t * Aj = m->Aj;
        const DT * Ax = (const DT *)m->Ax;
        const LT * Arl = (const LT *)m->Arl;
        if (m->threads != 1)
        {
            // defect into scratch, then x_new = x_old + t * dinv
            DT * t = (DT *)stream_work(st, sizeof(DT) * m->rows + 256);
            ell_launch<DT, LT, true>(m, x_old, t, rhs, st);
            unsigned g2 = grid_for(m->rows, kBlock);
            if (g2 > 16384) g2 = 16384;
            hipLaunchKernelGGL((k::elem_kernel<DT, k::E_SCALED_SUM_V>), dim3(g2), dim3(kBlock), 0, st, x_new,
                               (const DT *)nullptr, (const DT *)t, dinv, DT(0), m->rows);
            // x_new currently holds garbage + t*dinv unless it started as x_old
            HIP_CHECK(hipGetLastError());
            return;
        }
#define HS_K(K)                                                                                               \
        if (m->ncpr <= K)                                                                                     \
        {                                                                                                     \
            hipLaunchKernelGGL((k::ell_jacobi_t1<DT, uint32_t, LT, K>), dim3(grid), dim3(kBlock), 0, st, x_old, \
                               x_new, Aj, Ax, Arl, rhs, dinv, m->rows, m->stride);                            \
            HIP_CHECK(hipGetLastError());                                                                     \
            return;                                                                                           \
        }
        HS_K(4) HS_K(8) HS_K(9) HS_K(12) HS_K(13) HS_K(16) HS_K(24) HS_K(32)
#undef HS_K
        hipLaunchKernelGGL((k::ell_jacobi_t1<DT, uint32_t, LT, 0>), dim3(grid), dim3(kBlock), 0, st, x_old, x_new,
                           Aj, Ax, Arl, rhs, dinv, m->rows, m->stride);
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT, bool DEFECT>
    void ell_dispatch_len(const hsolve_ell_s * m, const DT * x, DT * y, const DT * rhs, hipStream_t st)
    {
        if (m->len_bytes == 1)
            ell_launch<DT, uint8_t, DEFECT>(m, x, y, rhs, st);
        else if (m->len_bytes == 2)
            ell_launch<DT, uint16_t, DEFECT>(m, x, y, rhs, st);
        else
            ell_launch<DT, uint32_t, DEFECT>(m, x, y, rhs, st);
    }

    template <typename DT>
    void ell_jacobi_dispatch(const hsolve_ell_s * m, const DT * rhs, const DT * dinv, const DT * x_old, DT * x_new,
                             hipStream_t st)
    {
        if (m->len_bytes == 1)
            ell_jacobi_launch<DT, uint8_t>(m, rhs, dinv, x_old, x_new, st);
        else if (m->len_bytes == 2)
            ell_jacobi_launch<DT, uint16_t>(m, rhs, dinv, x_old, x_new, st);
        else
            ell_jacobi_launch<DT, uint32_t>(m, rhs, dinv, x_old, x_new, st);
    }

    template <typename DT>
    std::vector<uint64_t> row_lengths_by_scan(uint64_t rows, uint64_t stride, uint64_t threads, uint64_t size,
                                              const DT * Ax)
    {
        // sparse_matrix_ell-impl.hh:186-207
        std::vector<uint64_t> rl(rows, 0);
        for (uint64_t row = 0; row < rows; ++row)
        {
            uint64_t count = 0;
            bool stop = false;
            for (uint64_t i = row * threads; i < size && ! stop; i += stride)
            {
                for (uint64_t t = 0; t < threads; ++t)
                {
                    if (Ax[i + t] == DT(0))
                    {
                        stop = true;
                        break;
                    }
                    if (t == 0)
                        ++count;
                }
            }
            rl[row] = count;
        }
        return rl;
    }

    template <typename LT>
    std::vector<uint8_t> pack_lengths(const std::vector<uint64_t> & rl)
    {
        std::vector<uint8_t> out(rl.size() * sizeof(LT));
        LT * p = (LT *)out.data();
        for (size_t i = 0; i < rl.size(); ++i)
            p[i] = (LT)rl[i];
        return out;
    }
}

extern "C" {

int hsolve_ell_create(int dtype, uint64_t rows, uint64_t columns, uint64_t stride, uint64_t num_cols_per_row,
                      uint64_t threads, const uint64_t * Aj_host, const void * Ax_host, const uint64_t * Arl_host,
                      hsolve_ell * out)
{
    HSOLVE_API_BEGIN
    require(out && Aj_host && Ax_host, HSOLVE_ERR_INVALID, "ell_create: null pointer");
    require(dtype == 0 || dtype == 1, HSOLVE_ERR_INVALID, "ell_create: dtype must be 0 (float) or 1 (double)");
    require(threads >= 1, HSOLVE_ERR_INVALID, "ell_create: threads must be >= 1");
    require(stride >= rows * threads, HSOLVE_ERR_INVALID, "ell_create: stride < rows * threads");
    require(columns <= 0xFFFFFFFFull, HSOLVE_ERR_SIZE, "ell_create: more than 2^32 columns");
    *out = nullptr;
    const uint64_t size = num_cols_per_row * stride;
    const size_t vs = dtype ? sizeof(double) : sizeof(float);

    std::vector<uint64_t> rl;
    if (Arl_host)
        rl.assign(Arl_host, Arl_host + rows);
    else if (dtype)
        rl = row_lengths_by_scan<double>(rows, stride, threads, size, (const double *)Ax_host);
    else
        rl = row_lengths_by_scan<float>(rows, stride, threads, size, (const float *)Ax_host);
    uint64_t maxlen = 0, nnz = 0;
    for (uint64_t r = 0; r < rows; ++r)
    {
        require(rl[r] <= num_cols_per_row, HSOLVE_ERR_INVALID, "ell_create: row length > num_cols_per_row");
        maxlen = std::max(maxlen, rl[r]);
        nnz += rl[r] * threads;
    }
    std::vector<uint32_t> aj32(size);
    for (uint64_t i = 0; i < size; ++i)
    {
        require(Aj_host[i] < (columns ? columns : 1), HSOLVE_ERR_INVALID, "ell_create: column index out of range");
        aj32[i] = (uint32_t)Aj_host[i];
    }
    std::vector<uint8_t> lens;
    int lb;
    if (maxlen <= 0xFF) { lens = pack_lengths<uint8_t>(rl); lb = 1; }
    else if (maxlen <= 0xFFFF) { lens = pack_lengths<uint16_t>(rl); lb = 2; }
    else { lens = pack_lengths<uint32_t>(rl); lb = 4; }

    hsolve_ell_s * m = new hsolve_ell_s();
    m->dtype = dtype; m->rows = rows; m->cols = columns; m->stride = stride; m->ncpr = num_cols_per_row;
    m->threads = threads; m->nnz = nnz; m->len_bytes = lb;
    m->Aj = nullptr; m->Ax = nullptr; m->Arl = nullptr;
    try
    {
        hipError_t e1 = hipMalloc((void **)&m->Aj, sizeof(uint32_t) * (size ? size : 1));
        hipError_t e2 = hipMalloc(&m->Ax, vs * (size ? size : 1));
        hipError_t e3 = hipMalloc(&m->Arl, lens.size() ? lens.size() : 1);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
            throw Error(HSOLVE_ERR_ALLOC, "ell_create: device allocation failed");
        HIP_CHECK(hipMemcpy(m->Aj, aj32.data(), sizeof(uint32_t) * size, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(m->Ax, Ax_host, vs * size, hipMemcpyHostToDevice));
        if (! lens.empty())
            HIP_CHECK(hipMemcpy(m->Arl, lens.data(), lens.size(), hipMemcpyHostToDevice));
    }
    catch (...)
    {
        if (m->Aj) (void)hipFree(m->Aj);
        if (m->Ax) (void)hipFree(m->Ax);
        if (m->Arl) (void)hipFree(m->Arl);
        delete m;
        throw;
    }
    m->device_bytes = sizeof(uint32_t) * size + vs * size + lens.size();
    *out = m;
    HSOLVE_API_END
}

int hsolve_ell_destroy(hsolve_ell m)
{
    HSOLVE_API_BEGIN
    if (m)
    {
        HIP_CHECK(hipFree(m->Aj));
        HIP_CHECK(hipFree(m->Ax));
        HIP_CHECK(hipFree(m->Arl));
        delete m;
    }
    HSOLVE_API_END
}

int hsolve_ell_info(hsolve_ell m, uint64_t * rows, uint64_t * columns, uint64_t * stride, uint64_t * ncpr,
                    uint64_t * threads, uint64_t * nnz, uint64_t * device_bytes)
{
    HSOLVE_API_BEGIN
    require(m != nullptr, HSOLVE_ERR_INVALID, "ell_info: null handle");
    if (rows) *rows = m->rows;
    if (columns) *columns = m->cols;
    if (stride) *stride = m->stride;
    if (ncpr) *ncpr = m->ncpr;
    if (threads) *threads = m->threads;
    if (nnz) *nnz = m->nnz;
    if (device_bytes) *device_bytes = m->device_bytes;
    HSOLVE_API_END
}

int hsolve_ell_spmv(hsolve_ell m, const void * x, void * y, const void * rhs, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && x && y, HSOLVE_ERR_INVALID, "ell_spmv: null pointer");
    require(x != y, HSOLVE_ERR_INVALID, "ell_spmv: y must not alias x");
    hipStream_t st = as_stream(stream);
    if (m->rows == 0)
        return HSOLVE_OK;
    if (m->dtype)
    {
        if (rhs)
            ell_dispatch_len<double, true>(m, (const double *)x, (double *)y, (const double *)rhs, st);
        else
            ell_dispatch_len<double, false>(m, (const double *)x, (double *)y, nullptr, st);
    }
    else
    {
        if (rhs)
            ell_dispatch_len<float, true>(m, (const float *)x, (float *)y, (const float *)rhs, st);
        else
            ell_dispatch_len<float, false>(m, (const float *)x, (float *)y, nullptr, st);
    }
    HSOLVE_API_END
}

int hsolve_ell_jacobi(hsolve_ell m, const void * rhs, const void * dinv, const void * x_old, void * x_new,
                      void * stream)
{
    HSOLVE_API_BEGIN
    require(m && rhs && dinv && x_old && x_new, HSOLVE_ERR_INVALID, "ell_jacobi: null pointer");
    require(x_old != x_new, HSOLVE_ERR_INVALID, "ell_jacobi: x_new must not alias x_old");
    hipStream_t st = as_stream(stream);
    if (m->threads != 1)
    {
        // the two-pass form needs x_new = x_old first
        const size_t vs = m->dtype ? sizeof(double) : sizeof(float);
        HIP_CHECK(hipMemcpyAsync(x_new, x_old, vs * m->rows, hipMemcpyDeviceToDevice, st));
    }
    if (m->dtype)
        ell_jacobi_dispatch<double>(m, (const double *)rhs, (const double *)dinv, (const double *)x_old,
                                    (double *)x_new, st);
    else
        ell_jacobi_dispatch<float>(m, (const float *)rhs, (const float *)dinv, (const float *)x_old, (float *)x_new,
                                   st);
    HSOLVE_API_END
}

}  // extern "C"
