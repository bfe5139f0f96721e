// ell.hip -- device-resident SparseMatrixELL handles.
//
// The host layout (HONEI's SMELL: uint64 column indices, uint64 row step
// counts, column-major slots of `stride` rows) is converted once at upload
// into the backend layout: the same slots in the same order, with
//  * column indices as int16 offsets from the row when every stored column
//    lies within +-32767 of its row (any banded or stencil matrix), else as
//    uint32 columns, and
//  * row lengths in the narrowest unsigned type that holds the longest row.
// For an FP64 9-point stencil that is 10 instead of 16 bytes per nonzero and
// 1 instead of 8 bytes per row length.  Which products are summed, and in
// which order, does not change.  A matrix with rows of varying length whose
// columns stay near the diagonal also gets a copy with the rows of every
// LDS block (1024 rows FP64, 512 FP32) sorted by length, read by the LDS
// window kernel.
#include "ell_handle.hh"
#include "kernels_mg.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hsolve
{
    void * stream_work(hipStream_t st, size_t bytes);

    // ConjugateGradients<NONE>(ELL, rhs, x, konv_rad) in one workgroup
    template <typename DT, typename IT, typename LT>
    __global__ __launch_bounds__(256) void cg_konv_ell_kernel(EllRowOp<DT, IT, LT> A, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, double konv_rad, uint64_t max_iters,
            uint64_t * iters_out)
    {
        k::cg_konv_body_op<DT>(A, rhs, x, work, n, konv_rad, max_iters, iters_out);
    }
}

using namespace hsolve;

namespace
{
    // Kernel variant of a handle: explicit choice, or for the threads == 1
    // layout two rows per lane with unconditional slot loads when at least
    // 90% of the slots below num_cols_per_row are used (stencils), and with
    // XCD-contiguous blocks otherwise (measured on MI355X: the FP64 stencil
    // is fastest without the remap, scattered matrices gain from it).
    int ell_variant(const hsolve_ell_s * m)
    {
        if (m->variant & HSOLVE_ELL_SET)
            return m->variant;
        if (m->nnz * 10 >= m->rows * m->ncpr * 9)
            return HSOLVE_ELL_FULL;
        return HSOLVE_ELL_XCD | HSOLVE_ELL_LDS | (m->sAx ? HSOLVE_ELL_SORT : 0);
    }

    // LDS bytes of ell_spmv_lds's x window (16 more for its start rounded
    // down to a 16-byte boundary), or 0 when the handle does not qualify
    // (threads == 1, columns within +-max_offset of the row, a window of at
    // most 48 KiB so that 3 blocks fit on a CU)
    size_t lds_window_bytes(const hsolve_ell_s * m, size_t vs)
    {
        const uint64_t w = (uint64_t)(vs == 8 ? k::lds_rows<double>() : k::lds_rows<float>()) + 2 * m->max_offset;
        if (m->threads != 1 || w * vs > 48 * 1024)
            return 0;
        return (size_t)(w * vs + 16);
    }

    // rows [r0, r1) (r0 even: the rows pair up as in the whole product)
    template <typename DT, typename IT, typename LT, int MODE, int K>
    bool ell_launch_r2(const hsolve_ell_s * m, int v, const DT * x, DT * y, const IT * Aj, const DT * Ax,
                       const LT * Arl, const DT * rhs, const DT * dinv, hipStream_t st, uint64_t r0, uint64_t r1)
    {
        if (m->ncpr > (uint64_t)K || m->ncpr == 0 || m->cols == 0)
            return false;
        const unsigned grid = grid_for((r1 - r0 + 1) / 2, kBlock);
        const bool full = v & HSOLVE_ELL_FULL, xcd = v & HSOLVE_ELL_XCD;
        if constexpr (K == 9)
        {
            // 9-point structured (automatic choice only): columns implied
            if (m->q1_side != 0 && full && ! xcd && (m->variant & HSOLVE_ELL_SET) == 0)
            {
                hipLaunchKernelGGL((k::ell_spmv_r2<DT, IT, LT, 9, MODE, true, false, true>), dim3(grid), dim3(kBlock),
                                   0, st, x, y, Aj, Ax, Arl, rhs, dinv, r0, r1, m->stride, (uint32_t)m->ncpr,
                                   (int64_t)m->q1_side);
                return true;
            }
        }
#define HS_R2(F, X)                                                                                           \
        hipLaunchKernelGGL((k::ell_spmv_r2<DT, IT, LT, K, MODE, F, X>), dim3(grid), dim3(kBlock), 0, st, x, y, Aj, \
                           Ax, Arl, rhs, dinv, r0, r1, m->stride, (uint32_t)m->ncpr)
        if (full && xcd) HS_R2(true, true);
        else if (full) HS_R2(true, false);
        else if (xcd) HS_R2(false, true);
        else HS_R2(false, false);
#undef HS_R2
        return true;
    }

    // ell_spmv_lds_p on a grid of as many workgroups as are resident on the
    // device at once (at most one per row block)
    template <typename DT, typename IT, typename LT, int K, int MODE, bool SORTED>
    void launch_lds_p(hipStream_t st, size_t lds, uint32_t nblk, const DT * x, DT * y, const IT * Aj, const DT * Ax,
                      const LT * Arl, const DT * rhs, const DT * dinv, const hsolve_ell_s * m, const uint16_t * perm)
    {
        auto kern = k::ell_spmv_lds_p<DT, IT, LT, K, MODE, SORTED>;
        int dev = 0, cus = 0, per_cu = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, k::lds_rows<DT>() / 2, lds));
        uint32_t G = (uint32_t)std::max(1, cus * std::max(1, per_cu));
        if (G > nblk)
            G = nblk;
        hipLaunchKernelGGL(kern, dim3(G), dim3(k::lds_rows<DT>() / 2), lds, st, x, y, Aj, Ax, Arl, rhs, dinv, m->rows,
                           m->cols, m->stride, (uint32_t)m->ncpr, m->max_offset, perm, nblk);
    }

    // rows [r0, r1) of the product / defect / sweep (the whole matrix by
    // default); a row range other than the whole matrix takes the row
    // kernels (the LDS-window kernels work on whole blocks)
    template <typename DT, int MODE>
    void ell_launch(const hsolve_ell_s * m, const DT * x, DT * y, const DT * rhs, const DT * dinv, hipStream_t st,
                    uint64_t r0 = 0, uint64_t r1 = ~(uint64_t)0)
    {
        if (r1 > m->rows)
            r1 = m->rows;
        if (r0 >= r1)
            return;
        const bool whole = r0 == 0 && r1 == m->rows;
        const unsigned grid = grid_for(r1 - r0, kBlock);
        const DT * Ax = (const DT *)m->Ax;
        const int v = ell_variant(m);
        // two rows per lane reads the slot pair at row_end - 1: safe when the
        // stride leaves a padding row or the row count is even (a range that
        // ends before the last row reads a row of the matrix); pairs start on
        // even rows
        const bool pairs_ok = m->threads == 1 && ! (v & HSOLVE_ELL_ONE_ROW);
        const bool r2_ok = pairs_ok && (r0 % 2 == 0) && (r1 < m->rows || m->stride > m->rows || m->rows % 2 == 0);
        // the LDS-window kernel reads the slot pairs of whole blocks: the
        // last pair of the matrix needs a padding row or an even row count
        const bool lds_pairs_ok = pairs_ok && (m->stride > m->rows || m->rows % 2 == 0);
        with_ell_types(m, [&](auto Aj, auto Arl) {
            using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
            using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
            const size_t lds = lds_window_bytes(m, sizeof(DT));
            // (ell_spmv_lds loads its window in 16-byte chunks: x 16-byte aligned)
            if (lds_pairs_ok && (v & HSOLVE_ELL_LDS) && lds && m->cols != 0 && m->ncpr >= 1 && m->ncpr <= 16 &&
                (((uintptr_t)x % 16) == 0 || (v & HSOLVE_ELL_PERSIST)))
            {
                // a row range takes the blocks that hold its rows
                const uint64_t R = (uint64_t)k::lds_rows<DT>();
                const uint64_t b0 = r0 / R;
                const unsigned g = (unsigned)((r1 + R - 1) / R - b0);
                const bool sorted = (v & HSOLVE_ELL_SORT) && m->sAx;
                const IT * sAj = (const IT *)m->sAj;
                const LT * sArl = (const LT *)m->sArl;
                const DT * sAx = (const DT *)m->sAx;
                const uint16_t * perm = (const uint16_t *)m->sperm;
                if ((v & HSOLVE_ELL_PERSIST) && whole)
                {
                    // a fixed grid: as many workgroups as fit on the device at once
                    const uint32_t nblk = (uint32_t)g;
#define HS_LP(K)                                                                                              \
                    if (sorted)                                                                               \
                        launch_lds_p<DT, IT, LT, K, MODE, true>(st, lds, nblk, x, y, sAj, sAx, sArl, rhs, dinv, m, perm); \
                    else                                                                                      \
                        launch_lds_p<DT, IT, LT, K, MODE, false>(st, lds, nblk, x, y, Aj, Ax, Arl, rhs, dinv, m,  \
                                                                 (const uint16_t *)nullptr)
                    if (m->ncpr <= 9)
                    {
                        HS_LP(9);
                    }
                    else if (m->ncpr <= 13)
                    {
                        HS_LP(13);
                    }
                    else
                    {
                        HS_LP(16);
                    }
#undef HS_LP
                    HIP_CHECK(hipGetLastError());
                    return;
                }
#define HS_L(K)                                                                                               \
                if (sorted)                                                                                   \
                    hipLaunchKernelGGL((k::ell_spmv_lds<DT, IT, LT, K, MODE, true>), dim3(g), dim3(k::lds_rows<DT>() / 2), lds, \
                                       st, x, y, sAj, sAx, sArl, rhs, dinv, m->rows, m->cols, m->stride,        \
                                       (uint32_t)m->ncpr, m->max_offset, perm, b0, r0, r1);                   \
                else                                                                                          \
                    hipLaunchKernelGGL((k::ell_spmv_lds<DT, IT, LT, K, MODE, false>), dim3(g), dim3(k::lds_rows<DT>() / 2), lds, \
                                       st, x, y, Aj, Ax, Arl, rhs, dinv, m->rows, m->cols, m->stride,          \
                                       (uint32_t)m->ncpr, m->max_offset, (const uint16_t *)nullptr, b0, r0, r1)
                if (m->ncpr <= 9)
                {
                    HS_L(9);
                }
                else if (m->ncpr <= 13)
                {
                    HS_L(13);
                }
                else
                {
                    HS_L(16);
                }
#undef HS_L
                HIP_CHECK(hipGetLastError());
                return;
            }
            if (r2_ok && (ell_launch_r2<DT, IT, LT, MODE, 4>(m, v, x, y, Aj, Ax, Arl, rhs, dinv, st, r0, r1) ||
                          ell_launch_r2<DT, IT, LT, MODE, 9>(m, v, x, y, Aj, Ax, Arl, rhs, dinv, st, r0, r1) ||
                          ell_launch_r2<DT, IT, LT, MODE, 13>(m, v, x, y, Aj, Ax, Arl, rhs, dinv, st, r0, r1) ||
                          ell_launch_r2<DT, IT, LT, MODE, 16>(m, v, x, y, Aj, Ax, Arl, rhs, dinv, st, r0, r1)))
                return;
            if (m->threads != 1)
                hipLaunchKernelGGL((k::ell_spmv_tn<DT, IT, LT, MODE>), dim3(grid), dim3(kBlock), 0, st, x, y, Aj, Ax,
                                   Arl, rhs, dinv, r0, r1, m->stride, (uint32_t)m->threads);
#define HS_K(K)                                                                                               \
            else if (m->ncpr <= K)                                                                            \
                hipLaunchKernelGGL((k::ell_spmv_t1<DT, IT, LT, K, MODE>), dim3(grid), dim3(kBlock), 0, st, x, y, Aj, \
                                   Ax, Arl, rhs, dinv, r0, r1, m->stride);
            HS_K(4) HS_K(8) HS_K(9) HS_K(12) HS_K(13) HS_K(16) HS_K(24) HS_K(32)
#undef HS_K
            else
                hipLaunchKernelGGL((k::ell_spmv_t1<DT, IT, LT, 0, MODE>), dim3(grid), dim3(kBlock), 0, st, x, y, Aj,
                                   Ax, Arl, rhs, dinv, r0, r1, m->stride);
        });
        HIP_CHECK(hipGetLastError());
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
    require(columns <= 0xFFFFFFFFull && rows <= 0xFFFFFFFFull, HSOLVE_ERR_SIZE, "ell_create: more than 2^32 rows/columns");
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
    int64_t maxoff = 0;
    for (uint64_t r = 0; r < rows; ++r)
    {
        require(rl[r] <= num_cols_per_row, HSOLVE_ERR_INVALID, "ell_create: row length > num_cols_per_row");
        maxlen = std::max(maxlen, rl[r]);
        nnz += rl[r] * threads;
        for (uint64_t k = 0; k < rl[r]; ++k)
            for (uint64_t t = 0; t < threads; ++t)
            {
                const uint64_t c = Aj_host[r * threads + t + k * stride];
                require(c < columns, HSOLVE_ERR_INVALID, "ell_create: column index out of range");
                maxoff = std::max<int64_t>(maxoff, std::llabs((int64_t)c - (int64_t)r));
            }
    }
    const int idx_kind = (maxoff <= 32767) ? 1 : 0;
    // 9-point structure on a side-M grid (Q1 stencils stored as ELL)
    uint64_t q1_side = 0;
    {
        const uint64_t M = (uint64_t)std::llround(std::sqrt((double)rows));
        bool q1 = threads == 1 && rows == columns && M >= 3 && M * M == rows && num_cols_per_row <= 9;
        for (uint64_t r = 0; q1 && r < rows; ++r)
        {
            const int64_t y = (int64_t)(r / M), x = (int64_t)(r % M);
            int64_t prev = -1;
            for (uint64_t k = 0; q1 && k < rl[r]; ++k)
            {
                const int64_t c = (int64_t)Aj_host[r + k * stride];
                const int64_t dy = c / (int64_t)M - y, dx = c % (int64_t)M - x;
                q1 = c > prev && dy >= -1 && dy <= 1 && dx >= -1 && dx <= 1;
                prev = c;
            }
        }
        if (q1)
            q1_side = M;
    }
    std::vector<uint8_t> idx(size * (idx_kind ? sizeof(int16_t) : sizeof(uint32_t)));
    if (idx_kind)
    {
        int16_t * d = (int16_t *)idx.data();
        for (uint64_t j = 0; j < size; ++j)
            d[j] = 0;
        for (uint64_t r = 0; r < rows; ++r)
            for (uint64_t k = 0; k < rl[r]; ++k)
                for (uint64_t t = 0; t < threads; ++t)
                {
                    const uint64_t j = r * threads + t + k * stride;
                    d[j] = (int16_t)((int64_t)Aj_host[j] - (int64_t)r);
                }
    }
    else
    {
        uint32_t * d = (uint32_t *)idx.data();
        for (uint64_t j = 0; j < size; ++j)
            d[j] = (uint32_t)std::min<uint64_t>(Aj_host[j], columns ? columns - 1 : 0);
    }
    std::vector<uint8_t> lens;
    int lb;
    if (maxlen <= 0xFF) { lens = pack_lengths<uint8_t>(rl); lb = 1; }
    else if (maxlen <= 0xFFFF) { lens = pack_lengths<uint16_t>(rl); lb = 2; }
    else { lens = pack_lengths<uint32_t>(rl); lb = 4; }
    // one zero length past the last row: the two-rows-per-lane kernels read
    // the pair (r, r + 1) without a branch
    lens.resize(lens.size() + 4, 0);

    // the LDS window kernel's copy with rows sorted by length inside each
    // LDS block, when that kernel is the automatic choice
    const size_t isz = idx_kind ? sizeof(int16_t) : sizeof(uint32_t);
    const uint64_t lds_rows = dtype ? (uint64_t)k::lds_rows<double>() : (uint64_t)k::lds_rows<float>();
    const bool sort_copy = threads == 1 && num_cols_per_row <= 16 && nnz * 10 < rows * num_cols_per_row * 9 &&
                           (lds_rows + 2 * (uint64_t)maxoff) * vs <= 48 * 1024 &&
                           (stride > rows || rows % 2 == 0) && rows > 0;
    std::vector<uint16_t> perm;
    std::vector<uint8_t> s_idx, s_ax, s_lens;
    if (sort_copy)
    {
        perm.assign(rows + 2, 0);     // padding entries past the last row, as for the lengths
        std::vector<uint64_t> order;
        for (uint64_t b0 = 0; b0 < rows; b0 += lds_rows)
        {
            const uint64_t b1 = std::min<uint64_t>(rows, b0 + lds_rows);
            order.resize(b1 - b0);
            for (uint64_t i = 0; i < b1 - b0; ++i)
                order[i] = b0 + i;
            std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return rl[a] > rl[b]; });
            for (uint64_t i = 0; i < b1 - b0; ++i)
                perm[b0 + i] = (uint16_t)(order[i] - b0);
        }
        s_idx.assign(idx.size(), 0);
        s_ax.assign(vs * size, 0);
        s_lens.assign(lens.size(), 0);
        const uint8_t * ax = (const uint8_t *)Ax_host;
        for (uint64_t p = 0; p < rows; ++p)
        {
            const uint64_t o = (p / lds_rows) * lds_rows + perm[p];
            std::memcpy(&s_lens[p * lb], &lens[o * lb], lb);
            for (uint64_t kk = 0; kk < num_cols_per_row; ++kk)
            {
                std::memcpy(&s_idx[(kk * stride + p) * isz], &idx[(kk * stride + o) * isz], isz);
                std::memcpy(&s_ax[(kk * stride + p) * vs], &ax[(kk * stride + o) * vs], vs);
            }
        }
    }

    hsolve_ell_s * m = new hsolve_ell_s();
    m->sAj = nullptr; m->sAx = nullptr; m->sArl = nullptr; m->sperm = nullptr;
    m->dtype = dtype; m->rows = rows; m->cols = columns; m->stride = stride; m->ncpr = num_cols_per_row;
    m->threads = threads; m->nnz = nnz; m->idx_kind = idx_kind; m->len_bytes = lb; m->variant = 0;
    m->q1_side = q1_side;
    m->max_offset = (uint64_t)maxoff;
    m->Aj = nullptr; m->Ax = nullptr; m->Arl = nullptr;
    try
    {
        hipError_t e1 = hipMalloc(&m->Aj, idx.size() ? idx.size() : 1);
        hipError_t e2 = hipMalloc(&m->Ax, vs * (size ? size : 1));
        hipError_t e3 = hipMalloc(&m->Arl, lens.size() ? lens.size() : 1);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
            throw Error(HSOLVE_ERR_ALLOC, "ell_create: device allocation failed");
        HIP_CHECK(hipMemcpy(m->Aj, idx.data(), idx.size(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(m->Ax, Ax_host, vs * size, hipMemcpyHostToDevice));
        if (! lens.empty())
            HIP_CHECK(hipMemcpy(m->Arl, lens.data(), lens.size(), hipMemcpyHostToDevice));
        if (sort_copy)
        {
            hipError_t f1 = hipMalloc(&m->sAj, s_idx.size());
            hipError_t f2 = hipMalloc(&m->sAx, s_ax.size());
            hipError_t f3 = hipMalloc(&m->sArl, s_lens.size());
            hipError_t f4 = hipMalloc(&m->sperm, perm.size() * sizeof(uint16_t));
            if (f1 != hipSuccess || f2 != hipSuccess || f3 != hipSuccess || f4 != hipSuccess)
                throw Error(HSOLVE_ERR_ALLOC, "ell_create: device allocation of the sorted copy failed");
            HIP_CHECK(hipMemcpy(m->sAj, s_idx.data(), s_idx.size(), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(m->sAx, s_ax.data(), s_ax.size(), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(m->sArl, s_lens.data(), s_lens.size(), hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(m->sperm, perm.data(), perm.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        }
    }
    catch (...)
    {
        if (m->Aj) (void)hipFree(m->Aj);
        if (m->Ax) (void)hipFree(m->Ax);
        if (m->Arl) (void)hipFree(m->Arl);
        if (m->sAj) (void)hipFree(m->sAj);
        if (m->sAx) (void)hipFree(m->sAx);
        if (m->sArl) (void)hipFree(m->sArl);
        if (m->sperm) (void)hipFree(m->sperm);
        delete m;
        throw;
    }
    m->device_bytes = idx.size() + vs * size + lens.size() +
                      (sort_copy ? s_idx.size() + s_ax.size() + s_lens.size() + perm.size() * sizeof(uint16_t) : 0);
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
        if (m->sAx)
        {
            HIP_CHECK(hipFree(m->sAj));
            HIP_CHECK(hipFree(m->sAx));
            HIP_CHECK(hipFree(m->sArl));
            HIP_CHECK(hipFree(m->sperm));
        }
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

int hsolve_ell_set_variant(hsolve_ell m, int variant)
{
    HSOLVE_API_BEGIN
    require(m != nullptr, HSOLVE_ERR_INVALID, "ell_set_variant: null handle");
    require(variant >= 0 && variant < 128 && (variant & HSOLVE_ELL_SET) == 0, HSOLVE_ERR_INVALID,
            "ell_set_variant: variant bits out of range");
    m->variant = variant ? (variant | HSOLVE_ELL_SET) : 0;
    HSOLVE_API_END
}

int hsolve_ell_q1_side(hsolve_ell m, uint64_t * side)
{
    HSOLVE_API_BEGIN
    require(m && side, HSOLVE_ERR_INVALID, "ell_q1_side: null pointer");
    *side = m->q1_side;
    HSOLVE_API_END
}

int hsolve_ell_spmv_rows(hsolve_ell m, const void * x, void * y, const void * rhs, uint64_t row_start,
                         uint64_t row_end, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && x && y, HSOLVE_ERR_INVALID, "ell_spmv_rows: null pointer");
    require(x != y, HSOLVE_ERR_INVALID, "ell_spmv_rows: y must not alias x");
    require(row_end <= m->rows, HSOLVE_ERR_SIZE, "ell_spmv_rows: row_end beyond the matrix");
    hipStream_t st = as_stream(stream);
    if (row_start >= row_end)
        return HSOLVE_OK;
    if (m->dtype)
    {
        if (rhs)
            ell_launch<double, 1>(m, (const double *)x, (double *)y, (const double *)rhs, nullptr, st, row_start, row_end);
        else
            ell_launch<double, 0>(m, (const double *)x, (double *)y, nullptr, nullptr, st, row_start, row_end);
    }
    else
    {
        if (rhs)
            ell_launch<float, 1>(m, (const float *)x, (float *)y, (const float *)rhs, nullptr, st, row_start, row_end);
        else
            ell_launch<float, 0>(m, (const float *)x, (float *)y, nullptr, nullptr, st, row_start, row_end);
    }
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
            ell_launch<double, 1>(m, (const double *)x, (double *)y, (const double *)rhs, nullptr, st);
        else
            ell_launch<double, 0>(m, (const double *)x, (double *)y, nullptr, nullptr, st);
    }
    else
    {
        if (rhs)
            ell_launch<float, 1>(m, (const float *)x, (float *)y, (const float *)rhs, nullptr, st);
        else
            ell_launch<float, 0>(m, (const float *)x, (float *)y, nullptr, nullptr, st);
    }
    HSOLVE_API_END
}

int hsolve_ell_jacobi(hsolve_ell m, const void * rhs, const void * dinv, const void * x_old, void * x_new,
                      void * stream)
{
    HSOLVE_API_BEGIN
    require(m && rhs && dinv && x_old && x_new, HSOLVE_ERR_INVALID, "ell_jacobi: null pointer");
    require(x_old != x_new, HSOLVE_ERR_INVALID, "ell_jacobi: x_new must not alias x_old");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "ell_jacobi: matrix is not square");
    hipStream_t st = as_stream(stream);
    if (m->dtype)
        ell_launch<double, 2>(m, (const double *)x_old, (double *)x_new, (const double *)rhs, (const double *)dinv, st);
    else
        ell_launch<float, 2>(m, (const float *)x_old, (float *)x_new, (const float *)rhs, (const float *)dinv, st);
    HSOLVE_API_END
}

int hsolve_cg_konv_ell(hsolve_ell m, const void * rhs, void * x, double konv_rad, uint64_t max_iters,
                       void * iters_dev, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && rhs && x, HSOLVE_ERR_INVALID, "cg_konv_ell: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "cg_konv_ell: matrix is not square");
    require(m->rows <= (1u << 20), HSOLVE_ERR_SIZE, "cg_konv_ell: more than 2^20 rows for one workgroup");
    hipStream_t st = as_stream(stream);
    const uint64_t n = m->rows;
    const uint64_t cap = max_iters ? max_iters : 100 * n + 1000;
    const size_t es = m->dtype ? sizeof(double) : sizeof(float);
    void * work = stream_work(st, es * 4 * n + 256);
    with_ell_types(m, [&](auto Aj, auto Arl) {
        using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
        using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
        if (m->dtype)
        {
            EllRowOp<double, IT, LT> A{Aj, (const double *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            hipLaunchKernelGGL((cg_konv_ell_kernel<double, IT, LT>), dim3(1), dim3(kBlock), 0, st, A,
                               (const double *)rhs, (double *)x, (double *)work, (int64_t)n, konv_rad, cap,
                               (uint64_t *)iters_dev);
        }
        else
        {
            EllRowOp<float, IT, LT> A{Aj, (const float *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            hipLaunchKernelGGL((cg_konv_ell_kernel<float, IT, LT>), dim3(1), dim3(kBlock), 0, st, A,
                               (const float *)rhs, (float *)x, (float *)work, (int64_t)n, konv_rad, cap,
                               (uint64_t *)iters_dev);
        }
    });
    HIP_CHECK(hipGetLastError());
    HSOLVE_API_END
}

}  // extern "C"
