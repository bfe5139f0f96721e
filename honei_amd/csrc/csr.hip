// csr.hip -- SparseMatrixCSR product and defect on gfx950 (blocked CSR of
// honei/la/sparse_matrix_csr.hh: Ar row offsets in blocks, Aj block start
// columns, Ax `bs` values per block).
//
// Row sums must keep the reference's order (blocks ascending, positions
// within a block ascending, from +0) to be bit-identical to
// Product<tags::CPU> (honei/la/product.hh:208-245) and Defect<tags::CPU>
// (honei/math/defect.hh:386-418).  One lane therefore owns one row, but the
// matrix is not read row by row: a workgroup of up to 256 rows streams its
// contiguous element range [Ar[r0] bs, Ar[r1] bs) through LDS in chunks, every
// lane loading consecutive elements (coalesced Ax and Aj), forming the
// products Ax[e] * x[col(e)] there, and each lane then adds its own row's
// products from LDS in order.  The product is rounded before the add (no
// contraction), as on the CPU.  A handle packs rows into workgroups by
// element count (a table built once at upload), so short rows fill a chunk
// and a long row gets a workgroup alone.  Positions at or past the last column
// contribute +0 instead of being skipped: a row sum starts at +0 and can
// never become -0, so adding +0 leaves it unchanged.
#include "common.hh"
#include "hsolve.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace hsolve
{
    constexpr int kCsrRows = 256;                  // lanes per workgroup, at most one row each
    constexpr int kCsrJ = 8;                       // elements per lane per chunk
    constexpr int kCsrChunk = kCsrRows * kCsrJ;    // elements staged per pass (16 KiB FP64)

    // Products of the chunk [c0, c0 + cn) into LDS.  Loads use indices
    // clamped into the chunk, so every lane issues all kCsrJ loads at once
    // without branches; only the LDS stores are guarded.
    // REL: Aj holds int16 offsets from `base`, the workgroup's first row.
    template <typename DT, typename IT, bool BS1, bool REL>
    __device__ inline void csr_stage(DT * prod, const DT * __restrict__ x, const IT * __restrict__ Aj,
                                     const DT * __restrict__ Ax, uint64_t c0, uint32_t cn, uint64_t bs,
                                     uint64_t columns, uint64_t base)
    {
        const uint32_t last = cn - 1;
        // block index and position of element c0 (bs > 1 only)
        const uint64_t blk0 = BS1 ? 0 : c0 / bs;
        const uint32_t off0 = BS1 ? 0 : (uint32_t)(c0 - blk0 * bs);
        const uint32_t b32 = (uint32_t)bs;
        uint64_t col[kCsrJ];
#pragma unroll
        for (int j = 0; j < kCsrJ; ++j)
        {
            const uint32_t k = (uint32_t)j * kCsrRows + threadIdx.x;
            const uint32_t kk = k < last ? k : last;
            const uint32_t q = BS1 ? 0 : (off0 + kk) / b32;
            const IT a = BS1 ? Aj[c0 + kk] : Aj[blk0 + q];
            const uint64_t start = REL ? (uint64_t)((int64_t)base + (int64_t)a) : (uint64_t)a;
            col[j] = BS1 ? start : start + (off0 + kk - q * b32);
        }
        DT a[kCsrJ], xv[kCsrJ];
#pragma unroll
        for (int j = 0; j < kCsrJ; ++j)
        {
            const uint32_t k = (uint32_t)j * kCsrRows + threadIdx.x;
            const uint32_t kk = k < last ? k : last;
            a[j] = Ax[c0 + kk];
            xv[j] = x[col[j] < columns ? col[j] : 0];
        }
#pragma unroll
        for (int j = 0; j < kCsrJ; ++j)
        {
            const uint32_t k = (uint32_t)j * kCsrRows + threadIdx.x;
            if (k < cn)
                prod[k] = col[j] < columns ? a[j] * xv[j] : DT(0);
        }
    }

    // Adds this lane's products of the staged chunk [c0, c0 + cn) in order.
    template <typename DT>
    __device__ inline void csr_sum(DT & sum, const DT * prod, uint64_t my_b, uint64_t my_e, uint64_t c0, uint32_t cn)
    {
        // the row's part of the chunk, empty for rows that end before it or
        // start after it and for lanes without a row (my_b = my_e = 0)
        const uint64_t lo64 = my_b > c0 ? my_b : c0;
        const uint64_t hi64 = my_e < c0 + cn ? my_e : c0 + cn;
        if (hi64 <= lo64)
            return;
        const uint32_t hi = (uint32_t)(hi64 - c0);    // lo < hi <= cn
        uint32_t e = (uint32_t)(lo64 - c0);
        // four LDS reads in flight, added in order
        for (; e + 4 <= hi; e += 4)
        {
            const DT v0 = prod[e], v1 = prod[e + 1], v2 = prod[e + 2], v3 = prod[e + 3];
            sum += v0;
            sum += v1;
            sum += v2;
            sum += v3;
        }
        for (; e < hi; ++e)
            sum += prod[e];
    }

    // One workgroup sums the rows [r0, r1) whose elements are [e0, e1).
    // ADAPT: (r0, e0) and (r1, e1) come from the handle's row-block table
    // (consecutive rows packed up to kCsrChunk elements and kCsrRows rows,
    // a longer row alone); otherwise 256 consecutive rows from row_begin.
    // The first chunk is straight-line code, so the lane's row bounds load
    // alongside the chunk (a value loaded before a loop that uses it makes
    // the compiler drain every load at the loop entry).
    template <typename DT, typename IT, typename RT, bool BS1, bool DEFECT, bool ADAPT, bool REL>
    __global__ __launch_bounds__(kCsrRows) void csr_stream_kernel(const DT * __restrict__ x, DT * __restrict__ y,
            const IT * __restrict__ Aj, const DT * __restrict__ Ax, const RT * __restrict__ Ar,
            const RT * __restrict__ blocks, const DT * __restrict__ rhs, uint64_t row_begin, uint64_t row_end,
            uint64_t bs, uint64_t columns, uint64_t blk_first)
    {
        __shared__ DT prod[kCsrChunk];
        uint64_t r0, r1, e0, e1;
        if (ADAPT)
        {
            // table workgroup blk_first + blockIdx.x; its rows outside
            // [row_begin, row_end) are staged but not written
            const uint64_t b = blk_first + blockIdx.x;
            r0 = (uint64_t)blocks[2 * b];
            e0 = (uint64_t)blocks[2 * b + 1];
            r1 = (uint64_t)blocks[2 * b + 2];
            e1 = (uint64_t)blocks[2 * b + 3];
        }
        else
        {
            r0 = row_begin + (uint64_t)blockIdx.x * kCsrRows;
            r1 = r0 + kCsrRows < row_end ? r0 + kCsrRows : row_end;
            e0 = (uint64_t)Ar[r0] * bs;
            e1 = (uint64_t)Ar[r1] * bs;
        }
        const uint64_t row = r0 + threadIdx.x;
        const bool live = row < r1 && (! ADAPT || (row >= row_begin && row < row_end));
        const uint64_t rr = row < r1 ? row : r0;
        DT sum = DT(0);
        if (columns != 0 && e0 < e1)
        {
            uint32_t cn = (uint32_t)(e1 - e0 < (uint64_t)kCsrChunk ? e1 - e0 : (uint64_t)kCsrChunk);
            const uint64_t ab = (uint64_t)Ar[rr] * bs, ae = (uint64_t)Ar[rr + 1] * bs;
            csr_stage<DT, IT, BS1, REL>(prod, x, Aj, Ax, e0, cn, bs, columns, r0);
            const uint64_t my_b = live ? ab : 0;
            const uint64_t my_e = live ? ae : 0;
            lds_barrier();
            csr_sum(sum, prod, my_b, my_e, e0, cn);
            for (uint64_t c0 = e0 + kCsrChunk; c0 < e1; c0 += kCsrChunk)
            {
                cn = (uint32_t)(e1 - c0 < (uint64_t)kCsrChunk ? e1 - c0 : (uint64_t)kCsrChunk);
                lds_barrier();
                csr_stage<DT, IT, BS1, REL>(prod, x, Aj, Ax, c0, cn, bs, columns, r0);
                lds_barrier();
                csr_sum(sum, prod, my_b, my_e, c0, cn);
            }
        }
        if (live)
        {
            if (DEFECT)
                y[row] = rhs[row] - sum;
            else
                y[row] = sum;
        }
    }

    // ADAPT: nblk table workgroups from blk_first; else 256-row workgroups
    // over [row_begin, row_end).
    template <typename DT, typename IT, typename RT, bool ADAPT, bool REL = false>
    void csr_launch(const DT * x, DT * y, const IT * Aj, const DT * Ax, const RT * Ar, const RT * blocks,
                    uint64_t blk_first, uint64_t nblk, const DT * rhs, uint64_t row_begin, uint64_t row_end,
                    uint64_t bs, uint64_t columns, hipStream_t st)
    {
        const uint64_t g = ADAPT ? nblk : (row_end > row_begin ? (row_end - row_begin + kCsrRows - 1) / kCsrRows : 0);
        if (g == 0)
            return;
        require(g <= 0x7FFFFFFFull, HSOLVE_ERR_SIZE, "csr: too many row blocks for one launch");
        const unsigned grid = (unsigned)g;
#define HS_CSR_GO(B1, DF)                                                                                     \
        csr_stream_kernel<DT, IT, RT, B1, DF, ADAPT, REL><<<grid, kCsrRows, 0, st>>>(x, y, Aj, Ax, Ar, blocks, \
                rhs, row_begin, row_end, bs, columns, blk_first)
        if (bs == 1)
        {
            if (rhs)
                HS_CSR_GO(true, true);
            else
                HS_CSR_GO(true, false);
        }
        else
        {
            if (rhs)
                HS_CSR_GO(false, true);
            else
                HS_CSR_GO(false, false);
        }
#undef HS_CSR_GO
        HIP_CHECK(hipGetLastError());
    }

    template <typename DT>
    void csr_raw(const char * name, const void * x, void * y, const void * Aj, const void * Ax, const void * Ar,
                 const void * rhs, uint64_t row_begin, uint64_t row_end, uint64_t atomicsize, uint64_t columns,
                 void * stream)
    {
        require(x && y && Aj && Ax && Ar, HSOLVE_ERR_INVALID, std::string(name) + ": null pointer");
        require(x != y, HSOLVE_ERR_INVALID, std::string(name) + ": y must not alias x");
        require(atomicsize >= 1 && atomicsize <= 0xFFFFu, HSOLVE_ERR_INVALID,
                std::string(name) + ": atomicsize must be in [1, 65535]");
        csr_launch<DT, uint64_t, uint64_t, false>((const DT *)x, (DT *)y, (const uint64_t *)Aj, (const DT *)Ax,
                                                  (const uint64_t *)Ar, nullptr, 0, 0, (const DT *)rhs, row_begin,
                                                  row_end, atomicsize, columns, as_stream(stream));
    }
}

struct hsolve_csr_s
{
    int dtype;          // 0 float, 1 double
    int idx32;          // 1: uint32 Ar and Aj, 0: uint64
    uint64_t rows, cols, bs, nblocks, used;
    void * Ar;
    void * Aj;
    void * Ax;
    int rel;            // 1: Aj as int16 offsets from the first row of the element's workgroup
    void * blocks;      // row-block table: (first row, first element) pairs, nblk + 1 of them
    uint64_t nblk;
    std::vector<uint64_t> first_rows;   // host copy of the table's first rows (nblk + 1)
    // ELL form of the same matrix (blocksize 1, short rows of similar
    // length): a row's entries become its slots in the same order, so the
    // ELL kernels give the same sums; used for full-range products
    hsolve_ell ell = nullptr;
    int variant = 0;                    // 0 automatic, 1 CSR kernel only
    uint64_t device_bytes;
};

using namespace hsolve;

namespace
{
    template <typename DT, typename T>
    void csr_handle_launch_t(hsolve_csr m, const void * x, void * y, const void * rhs, uint64_t r0, uint64_t r1,
                             hipStream_t st)
    {
        if (r1 <= r0)
            return;
        // table workgroups that hold rows of [r0, r1)
        const auto & fr = m->first_rows;
        const uint64_t b0 = (uint64_t)(std::upper_bound(fr.begin(), fr.end(), r0) - fr.begin()) - 1;
        const uint64_t b1 = (uint64_t)(std::lower_bound(fr.begin(), fr.end(), r1) - fr.begin());
        if (m->rel)
            csr_launch<DT, int16_t, T, true, true>((const DT *)x, (DT *)y, (const int16_t *)m->Aj, (const DT *)m->Ax,
                                                   (const T *)m->Ar, (const T *)m->blocks, b0, b1 - b0,
                                                   (const DT *)rhs, r0, r1, m->bs, m->cols, st);
        else
            csr_launch<DT, T, T, true>((const DT *)x, (DT *)y, (const T *)m->Aj, (const DT *)m->Ax, (const T *)m->Ar,
                                       (const T *)m->blocks, b0, b1 - b0, (const DT *)rhs, r0, r1, m->bs, m->cols,
                                       st);
    }

    template <typename DT>
    void csr_handle_launch(hsolve_csr m, const void * x, void * y, const void * rhs, uint64_t r0, uint64_t r1,
                           hipStream_t st)
    {
        if (m->idx32)
            csr_handle_launch_t<DT, uint32_t>(m, x, y, rhs, r0, r1, st);
        else
            csr_handle_launch_t<DT, uint64_t>(m, x, y, rhs, r0, r1, st);
    }

    // Consecutive rows packed into workgroups of at most kCsrRows rows and
    // kCsrChunk elements; a row longer than that gets a workgroup alone.
    std::vector<uint64_t> csr_row_blocks(uint64_t rows, const uint64_t * Ar, uint64_t bs)
    {
        std::vector<uint64_t> t;
        uint64_t r = 0;
        while (r < rows)
        {
            const uint64_t e0 = Ar[r] * bs;
            t.push_back(r);
            t.push_back(e0);
            uint64_t r1 = r + 1;
            while (r1 < rows && r1 - r < (uint64_t)kCsrRows && Ar[r1 + 1] * bs - e0 <= (uint64_t)kCsrChunk)
                ++r1;
            r = r1;
        }
        t.push_back(rows);
        t.push_back(Ar[rows] * bs);
        return t;
    }
}

extern "C" {

int hsolve_product_csr_dv_float(const void * x, void * y, const void * Aj, const void * Ax, const void * Ar,
                                uint64_t row_start, uint64_t row_end, uint64_t atomicsize, uint64_t columns,
                                void * stream)
{
    HSOLVE_API_BEGIN
    csr_raw<float>("product_csr_dv_float", x, y, Aj, Ax, Ar, nullptr, row_start, row_end, atomicsize, columns,
                   stream);
    HSOLVE_API_END
}

int hsolve_product_csr_dv_double(const void * x, void * y, const void * Aj, const void * Ax, const void * Ar,
                                 uint64_t row_start, uint64_t row_end, uint64_t atomicsize, uint64_t columns,
                                 void * stream)
{
    HSOLVE_API_BEGIN
    csr_raw<double>("product_csr_dv_double", x, y, Aj, Ax, Ar, nullptr, row_start, row_end, atomicsize, columns,
                    stream);
    HSOLVE_API_END
}

int hsolve_defect_csr_dv_float(const void * rhs, void * result, const void * Aj, const void * Ax, const void * Ar,
                               const void * b, uint64_t rows, uint64_t atomicsize, uint64_t columns, void * stream)
{
    HSOLVE_API_BEGIN
    require(rhs != nullptr, HSOLVE_ERR_INVALID, "defect_csr_dv_float: null rhs");
    csr_raw<float>("defect_csr_dv_float", b, result, Aj, Ax, Ar, rhs, 0, rows, atomicsize, columns, stream);
    HSOLVE_API_END
}

int hsolve_defect_csr_dv_double(const void * rhs, void * result, const void * Aj, const void * Ax, const void * Ar,
                                const void * b, uint64_t rows, uint64_t atomicsize, uint64_t columns, void * stream)
{
    HSOLVE_API_BEGIN
    require(rhs != nullptr, HSOLVE_ERR_INVALID, "defect_csr_dv_double: null rhs");
    csr_raw<double>("defect_csr_dv_double", b, result, Aj, Ax, Ar, rhs, 0, rows, atomicsize, columns, stream);
    HSOLVE_API_END
}

int hsolve_csr_create(int dtype, uint64_t rows, uint64_t columns, uint64_t blocksize, const uint64_t * Ar_host,
                      const uint64_t * Aj_host, const void * Ax_host, uint64_t used_elements, hsolve_csr * out)
{
    HSOLVE_API_BEGIN
    require(out && Ar_host, HSOLVE_ERR_INVALID, "csr_create: null pointer");
    require(dtype == 0 || dtype == 1, HSOLVE_ERR_INVALID, "csr_create: dtype must be 0 (float) or 1 (double)");
    require(blocksize >= 1 && blocksize <= 0xFFFFu, HSOLVE_ERR_INVALID, "csr_create: blocksize must be in [1, 65535]");
    *out = nullptr;
    const uint64_t nb = Ar_host[rows];
    require(nb == 0 || (Aj_host && Ax_host), HSOLVE_ERR_INVALID, "csr_create: null Aj/Ax");
    for (uint64_t r = 0; r < rows; ++r)
        require(Ar_host[r] <= Ar_host[r + 1], HSOLVE_ERR_INVALID, "csr_create: Ar not ascending");
    require(nb <= UINT64_MAX / blocksize, HSOLVE_ERR_SIZE, "csr_create: element count overflows");
    const size_t vs = dtype ? sizeof(double) : sizeof(float);
    const bool idx32 = nb * blocksize <= 0xFFFFFFFFull && columns + blocksize <= 0xFFFFFFFFull;
    const size_t is = idx32 ? 4 : 8;
    std::vector<uint32_t> ar32, aj32;
    if (idx32)
    {
        ar32.resize(rows + 1);
        aj32.resize(nb);
        for (uint64_t r = 0; r <= rows; ++r)
            ar32[r] = (uint32_t)Ar_host[r];
        for (uint64_t i = 0; i < nb; ++i)
            aj32[i] = (uint32_t)(Aj_host[i] < columns ? Aj_host[i] : columns);
    }
    const std::vector<uint64_t> rb = csr_row_blocks(rows, Ar_host, blocksize);
    std::vector<uint32_t> rb32;
    if (idx32)
        rb32.assign(rb.begin(), rb.end());
    const uint64_t nblk = rb.size() / 2 - 1;
    // 16-bit block columns relative to the first row of their workgroup
    // when every one fits and a row's columns come in runs (at least half
    // of the steps between neighbouring entries of a row are <= 4: stencil
    // and FEM matrices, whose gathers coalesce and whose bytes dominate).
    // Measured on MI355X: Matrix A (9-point) 390 -> 359 us; a matrix with
    // scattered columns (Matrix B) went from 699 to 790 us with 16-bit
    // columns, so it keeps 32-bit ones.
    uint64_t steps = 0, short_steps = 0;
    for (uint64_t r = 0; idx32 && r < rows; ++r)
        for (uint64_t i = Ar_host[r] + 1; i < Ar_host[r + 1]; ++i)
        {
            ++steps;
            short_steps += (aj32[i] >= aj32[i - 1] && aj32[i] - aj32[i - 1] <= 4 * blocksize) ? 1 : 0;
        }
    std::vector<int16_t> aj16;
    bool rel = idx32 && nb > 0 && short_steps * 2 >= steps;
    for (uint64_t b = 0; rel && b < nblk; ++b)
    {
        const int64_t r0 = (int64_t)rb[2 * b];
        for (uint64_t i = Ar_host[rb[2 * b]]; rel && i < Ar_host[rb[2 * b + 2]]; ++i)
        {
            const int64_t d = (int64_t)aj32[i] - r0;
            rel = d >= -32768 && d <= 32767;
        }
    }
    if (rel)
    {
        aj16.resize(nb);
        for (uint64_t b = 0; b < nblk; ++b)
            for (uint64_t i = Ar_host[rb[2 * b]]; i < Ar_host[rb[2 * b + 2]]; ++i)
                aj16[i] = (int16_t)((int64_t)aj32[i] - (int64_t)rb[2 * b]);
    }
    const size_t js = rel ? sizeof(int16_t) : is;
    hsolve_csr_s * m = new hsolve_csr_s();
    m->dtype = dtype; m->idx32 = idx32 ? 1 : 0; m->rows = rows; m->cols = columns; m->bs = blocksize;
    m->nblocks = nb; m->used = used_elements; m->rel = rel ? 1 : 0;
    m->Ar = nullptr; m->Aj = nullptr; m->Ax = nullptr; m->blocks = nullptr;
    m->nblk = nblk;
    m->first_rows.resize(nblk + 1);
    for (uint64_t b = 0; b <= nblk; ++b)
        m->first_rows[b] = rb[2 * b];
    try
    {
        hipError_t e1 = hipMalloc(&m->Ar, is * (rows + 1));
        hipError_t e2 = hipMalloc(&m->Aj, js * (nb ? nb : 1));
        hipError_t e3 = hipMalloc(&m->Ax, vs * (nb ? nb * blocksize : 1));
        hipError_t e4 = hipMalloc(&m->blocks, is * rb.size());
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess)
            throw Error(HSOLVE_ERR_ALLOC, "csr_create: device allocation failed");
        if (idx32)
            HIP_CHECK(hipMemcpy(m->blocks, rb32.data(), is * rb.size(), hipMemcpyHostToDevice));
        else
            HIP_CHECK(hipMemcpy(m->blocks, rb.data(), is * rb.size(), hipMemcpyHostToDevice));
        if (idx32)
        {
            HIP_CHECK(hipMemcpy(m->Ar, ar32.data(), is * (rows + 1), hipMemcpyHostToDevice));
            if (rel)
                HIP_CHECK(hipMemcpy(m->Aj, aj16.data(), js * nb, hipMemcpyHostToDevice));
            else if (nb)
                HIP_CHECK(hipMemcpy(m->Aj, aj32.data(), is * nb, hipMemcpyHostToDevice));
        }
        else
        {
            HIP_CHECK(hipMemcpy(m->Ar, Ar_host, is * (rows + 1), hipMemcpyHostToDevice));
            if (nb)
                HIP_CHECK(hipMemcpy(m->Aj, Aj_host, is * nb, hipMemcpyHostToDevice));
        }
        if (nb)
            HIP_CHECK(hipMemcpy(m->Ax, Ax_host, vs * nb * blocksize, hipMemcpyHostToDevice));
    }
    catch (...)
    {
        if (m->Ar) (void)hipFree(m->Ar);
        if (m->Aj) (void)hipFree(m->Aj);
        if (m->Ax) (void)hipFree(m->Ax);
        if (m->blocks) (void)hipFree(m->blocks);
        delete m;
        throw;
    }
    m->device_bytes = is * (rows + 1) + js * nb + vs * nb * blocksize + is * rb.size();
    // the ELL form: blocksize 1, every column below `columns`, rows of at
    // most 16 entries and at most 50% padding
    if (blocksize == 1 && rows > 0 && nb > 0 && columns <= 0xFFFFFFFFull && rows <= 0xFFFFFFFFull)
    {
        uint64_t maxlen = 0;
        bool ok = true;
        for (uint64_t r = 0; r < rows; ++r)
            maxlen = std::max<uint64_t>(maxlen, Ar_host[r + 1] - Ar_host[r]);
        for (uint64_t i = 0; ok && i < nb; ++i)
            ok = Aj_host[i] < columns;
        const uint64_t stride = 32 * ((rows + 31) / 32);
        if (ok && maxlen >= 1 && maxlen <= 16 && rows * maxlen * 2 <= nb * 3)
        {
            std::vector<uint64_t> eaj(maxlen * stride, 0), arl(rows);
            std::vector<uint8_t> eax(vs * maxlen * stride, 0);
            const uint8_t * ax = (const uint8_t *)Ax_host;
            for (uint64_t r = 0; r < rows; ++r)
            {
                arl[r] = Ar_host[r + 1] - Ar_host[r];
                for (uint64_t k = 0; k < arl[r]; ++k)
                {
                    eaj[k * stride + r] = Aj_host[Ar_host[r] + k];
                    std::memcpy(&eax[(k * stride + r) * vs], &ax[(Ar_host[r] + k) * vs], vs);
                }
            }
            hsolve_ell e = nullptr;
            if (hsolve_ell_create(dtype, rows, columns, stride, maxlen, 1, eaj.data(), eax.data(), arl.data(), &e) ==
                HSOLVE_OK)
            {
                uint64_t eb = 0;
                (void)hsolve_ell_info(e, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &eb);
                m->ell = e;
                m->device_bytes += eb;
            }
        }
    }
    *out = m;
    HSOLVE_API_END
}

int hsolve_csr_destroy(hsolve_csr m)
{
    HSOLVE_API_BEGIN
    if (m)
    {
        HIP_CHECK(hipFree(m->Ar));
        HIP_CHECK(hipFree(m->Aj));
        HIP_CHECK(hipFree(m->Ax));
        HIP_CHECK(hipFree(m->blocks));
        if (m->ell)
            (void)hsolve_ell_destroy(m->ell);
        delete m;
    }
    HSOLVE_API_END
}

int hsolve_csr_info(hsolve_csr m, uint64_t * rows, uint64_t * columns, uint64_t * blocksize, uint64_t * nblocks,
                    uint64_t * used_elements, uint64_t * device_bytes)
{
    HSOLVE_API_BEGIN
    require(m != nullptr, HSOLVE_ERR_INVALID, "csr_info: null handle");
    if (rows) *rows = m->rows;
    if (columns) *columns = m->cols;
    if (blocksize) *blocksize = m->bs;
    if (nblocks) *nblocks = m->nblocks;
    if (used_elements) *used_elements = m->used;
    if (device_bytes) *device_bytes = m->device_bytes;
    HSOLVE_API_END
}

int hsolve_csr_set_variant(hsolve_csr m, int variant)
{
    HSOLVE_API_BEGIN
    require(m != nullptr, HSOLVE_ERR_INVALID, "csr_set_variant: null handle");
    require(variant == 0 || variant == 1, HSOLVE_ERR_INVALID, "csr_set_variant: variant must be 0 or 1");
    m->variant = variant;
    HSOLVE_API_END
}

int hsolve_csr_form(hsolve_csr m, int * form)
{
    HSOLVE_API_BEGIN
    require(m && form, HSOLVE_ERR_INVALID, "csr_form: null pointer");
    *form = m->ell ? 1 : 0;
    HSOLVE_API_END
}

int hsolve_csr_spmv(hsolve_csr m, const void * x, void * y, const void * rhs, uint64_t row_start, uint64_t row_end,
                    void * stream)
{
    HSOLVE_API_BEGIN
    require(m && x && y, HSOLVE_ERR_INVALID, "csr_spmv: null pointer");
    require(x != y, HSOLVE_ERR_INVALID, "csr_spmv: y must not alias x");
    if (row_end == 0)
        row_end = m->rows;
    require(row_start <= row_end && row_end <= m->rows, HSOLVE_ERR_INVALID, "csr_spmv: row range out of bounds");
    if (m->ell && m->variant == 0 && row_start == 0 && row_end == m->rows)
        return hsolve_ell_spmv(m->ell, x, y, rhs, stream);
    hipStream_t st = as_stream(stream);
    if (m->dtype)
        csr_handle_launch<double>(m, x, y, rhs, row_start, row_end, st);
    else
        csr_handle_launch<float>(m, x, y, rhs, row_start, row_end, st);
    HSOLVE_API_END
}

}  // extern "C"
