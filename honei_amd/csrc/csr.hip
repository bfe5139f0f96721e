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

#include <vector>

namespace hsolve
{
    constexpr int kCsrRows = 256;                  // lanes per workgroup, at most one row each
    constexpr int kCsrJ = 8;                       // elements per lane per chunk
    constexpr int kCsrChunk = kCsrRows * kCsrJ;    // elements staged per pass (16 KiB FP64)

    // Products of the chunk [c0, c0 + cn) into LDS.  Loads use indices
    // clamped into the chunk, so every lane issues all kCsrJ loads at once
    // without branches; only the LDS stores are guarded.
    template <typename DT, typename IT, bool BS1>
    __device__ inline void csr_stage(DT * prod, const DT * __restrict__ x, const IT * __restrict__ Aj,
                                     const DT * __restrict__ Ax, uint64_t c0, uint32_t cn, uint64_t bs,
                                     uint64_t columns)
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
            if (BS1)
                col[j] = (uint64_t)Aj[c0 + kk];
            else
            {
                const uint32_t q = (off0 + kk) / b32;
                col[j] = (uint64_t)Aj[blk0 + q] + (off0 + kk - q * b32);
            }
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
    template <typename DT, typename IT, typename RT, bool BS1, bool DEFECT, bool ADAPT>
    __global__ __launch_bounds__(kCsrRows) void csr_stream_kernel(const DT * __restrict__ x, DT * __restrict__ y,
            const IT * __restrict__ Aj, const DT * __restrict__ Ax, const RT * __restrict__ Ar,
            const RT * __restrict__ blocks, const DT * __restrict__ rhs, uint64_t row_begin, uint64_t row_end,
            uint64_t bs, uint64_t columns)
    {
        __shared__ DT prod[kCsrChunk];
        uint64_t r0, r1, e0, e1;
        if (ADAPT)
        {
            r0 = (uint64_t)blocks[2 * blockIdx.x];
            e0 = (uint64_t)blocks[2 * blockIdx.x + 1];
            r1 = (uint64_t)blocks[2 * blockIdx.x + 2];
            e1 = (uint64_t)blocks[2 * blockIdx.x + 3];
        }
        else
        {
            r0 = row_begin + (uint64_t)blockIdx.x * kCsrRows;
            r1 = r0 + kCsrRows < row_end ? r0 + kCsrRows : row_end;
            e0 = (uint64_t)Ar[r0] * bs;
            e1 = (uint64_t)Ar[r1] * bs;
        }
        const uint64_t row = r0 + threadIdx.x;
        const bool live = row < r1;
        const uint64_t rr = live ? row : r0;
        DT sum = DT(0);
        if (columns != 0 && e0 < e1)
        {
            uint32_t cn = (uint32_t)(e1 - e0 < (uint64_t)kCsrChunk ? e1 - e0 : (uint64_t)kCsrChunk);
            const uint64_t ab = (uint64_t)Ar[rr] * bs, ae = (uint64_t)Ar[rr + 1] * bs;
            csr_stage<DT, IT, BS1>(prod, x, Aj, Ax, e0, cn, bs, columns);
            const uint64_t my_b = live ? ab : 0;
            const uint64_t my_e = live ? ae : 0;
            lds_barrier();
            csr_sum(sum, prod, my_b, my_e, e0, cn);
            for (uint64_t c0 = e0 + kCsrChunk; c0 < e1; c0 += kCsrChunk)
            {
                cn = (uint32_t)(e1 - c0 < (uint64_t)kCsrChunk ? e1 - c0 : (uint64_t)kCsrChunk);
                lds_barrier();
                csr_stage<DT, IT, BS1>(prod, x, Aj, Ax, c0, cn, bs, columns);
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

    template <typename DT, typename IT, typename RT, bool ADAPT>
    void csr_launch(const DT * x, DT * y, const IT * Aj, const DT * Ax, const RT * Ar, const RT * blocks,
                    uint64_t nblk, const DT * rhs, uint64_t row_begin, uint64_t row_end, uint64_t bs,
                    uint64_t columns, hipStream_t st)
    {
        const uint64_t g = ADAPT ? nblk : (row_end > row_begin ? (row_end - row_begin + kCsrRows - 1) / kCsrRows : 0);
        if (g == 0)
            return;
        require(g <= 0x7FFFFFFFull, HSOLVE_ERR_SIZE, "csr: too many row blocks for one launch");
        const unsigned grid = (unsigned)g;
#define HS_CSR_GO(B1, DF)                                                                                     \
        csr_stream_kernel<DT, IT, RT, B1, DF, ADAPT><<<grid, kCsrRows, 0, st>>>(x, y, Aj, Ax, Ar, blocks, rhs, \
                row_begin, row_end, bs, columns)
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
                                                  (const uint64_t *)Ar, nullptr, 0, (const DT *)rhs, row_begin,
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
    void * blocks;      // row-block table: (first row, first element) pairs, nblk + 1 of them
    uint64_t nblk;
    uint64_t device_bytes;
};

using namespace hsolve;

namespace
{
    template <typename DT, typename T>
    void csr_handle_launch_t(hsolve_csr m, const void * x, void * y, const void * rhs, uint64_t r0, uint64_t r1,
                             hipStream_t st)
    {
        if (r0 == 0 && r1 == m->rows)
            csr_launch<DT, T, T, true>((const DT *)x, (DT *)y, (const T *)m->Aj, (const DT *)m->Ax, (const T *)m->Ar,
                                       (const T *)m->blocks, m->nblk, (const DT *)rhs, r0, r1, m->bs, m->cols, st);
        else
            csr_launch<DT, T, T, false>((const DT *)x, (DT *)y, (const T *)m->Aj, (const DT *)m->Ax, (const T *)m->Ar,
                                        nullptr, 0, (const DT *)rhs, r0, r1, m->bs, m->cols, st);
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
    hsolve_csr_s * m = new hsolve_csr_s();
    m->dtype = dtype; m->idx32 = idx32 ? 1 : 0; m->rows = rows; m->cols = columns; m->bs = blocksize;
    m->nblocks = nb; m->used = used_elements;
    m->Ar = nullptr; m->Aj = nullptr; m->Ax = nullptr; m->blocks = nullptr;
    m->nblk = rb.size() / 2 - 1;
    try
    {
        hipError_t e1 = hipMalloc(&m->Ar, is * (rows + 1));
        hipError_t e2 = hipMalloc(&m->Aj, is * (nb ? nb : 1));
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
            if (nb)
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
    m->device_bytes = is * (rows + 1) + is * nb + vs * nb * blocksize + is * rb.size();
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

int hsolve_csr_spmv(hsolve_csr m, const void * x, void * y, const void * rhs, uint64_t row_start, uint64_t row_end,
                    void * stream)
{
    HSOLVE_API_BEGIN
    require(m && x && y, HSOLVE_ERR_INVALID, "csr_spmv: null pointer");
    require(x != y, HSOLVE_ERR_INVALID, "csr_spmv: y must not alias x");
    if (row_end == 0)
        row_end = m->rows;
    require(row_start <= row_end && row_end <= m->rows, HSOLVE_ERR_INVALID, "csr_spmv: row range out of bounds");
    hipStream_t st = as_stream(stream);
    if (m->dtype)
        csr_handle_launch<double>(m, x, y, rhs, row_start, row_end, st);
    else
        csr_handle_launch<float>(m, x, y, rhs, row_start, row_end, st);
    HSOLVE_API_END
}

}  // extern "C"
