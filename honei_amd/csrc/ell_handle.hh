// ell_handle.hh -- the device-resident SparseMatrixELL handle and the
// dispatch over its storage variants.
#pragma once

#include "kernels_la.hh"

struct hsolve_ell_s
{
    int dtype;                 // 0 float, 1 double
    uint64_t rows, cols, stride, ncpr, threads, nnz;
    int idx_kind;              // 0: uint32 columns, 1: int16 offsets (column - row)
    int len_bytes;             // row-length type: 1, 2 or 4 bytes
    void * Aj;                 // ncpr * stride indices
    void * Ax;                 // ncpr * stride values
    void * Arl;                // rows row lengths
    uint64_t device_bytes;
    int variant;               // kernel variant bits (HSOLVE_ELL_*), 0 = automatic
    uint64_t max_offset;       // largest |column - row| over the stored entries
    uint64_t q1_side;          // M > 0 when rows = cols = M^2, threads = 1 and every stored
                               // column of row (y, x) is a neighbour (y + dy, x + dx),
                               // |dy|, |dx| <= 1, in ascending slot order; else 0
    // Copy for the LDS window kernel, built at upload when that kernel is
    // the automatic choice: Aj, Ax and Arl with each LDS block's rows in
    // descending order of length, and the block-local original row of every
    // storage position (uint16).  nullptr otherwise.
    void * sAj;
    void * sAx;
    void * sArl;
    void * sperm;
};

// kernel variant bits of hsolve_ell_set_variant
enum
{
    HSOLVE_ELL_ONE_ROW = 1,    // one row per lane (else two rows per lane)
    HSOLVE_ELL_FULL = 2,       // load all num_cols_per_row slots without waiting for row lengths
    HSOLVE_ELL_XCD = 4,        // XCD-contiguous block order
    HSOLVE_ELL_SET = 8,        // marks an explicit choice
    HSOLVE_ELL_LDS = 16,       // x window in LDS (banded matrices; else ignored)
    HSOLVE_ELL_SORT = 32,      // with LDS: rows sorted by length inside each block (when the copy exists)
    HSOLVE_ELL_PERSIST = 64    // with LDS: a fixed grid walks the row blocks, next lengths loaded ahead
};

namespace hsolve
{
    // A row of y = A x of an ELL handle for the generic kernels (the
    // reference's per-row order: slots k ascending, lanes t within a slot).
    template <typename DT, typename IT, typename LT> struct EllRowOp
    {
        const IT * Aj;
        const DT * Ax;
        const LT * Arl;
        uint64_t stride;
        uint32_t threads;
        // x: a pointer, or any value whose x[j] yields the j-th entry
        template <typename XV> __device__ __forceinline__ DT row(const XV & x, int64_t i) const
        {
            const uint32_t len = (uint32_t)Arl[i];
            if (threads == 1)
                return k::ell_row_sum<DT, IT, 0>(x, Aj + i, Ax + i, len, (uint64_t)i, stride);
            DT sum = DT(0);
            uint64_t j = (uint64_t)i * threads;
            for (uint32_t kk = 0; kk < len; ++kk, j += stride)
                for (uint32_t t = 0; t < threads; ++t)
                    sum += Ax[j + t] * x[k::col_of<IT>(Aj[j + t], (uint64_t)i)];
            return sum;
        }
    };

    // Calls f.template operator()<IT, LT>(Aj, Arl) with the handle's typed
    // index and row-length arrays.
    template <typename F>
    void with_ell_types(const hsolve_ell_s * m, F && f)
    {
        auto by_len = [&](auto it_tag) {
            using IT = decltype(it_tag);
            const IT * Aj = (const IT *)m->Aj;
            if (m->len_bytes == 1)
                f(Aj, (const uint8_t *)m->Arl);
            else if (m->len_bytes == 2)
                f(Aj, (const uint16_t *)m->Arl);
            else
                f(Aj, (const uint32_t *)m->Arl);
        };
        if (m->idx_kind == 1)
            by_len(int16_t(0));
        else
            by_len(uint32_t(0));
    }
}
