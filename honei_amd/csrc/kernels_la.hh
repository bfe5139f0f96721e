// kernels_la.hh -- gfx950 kernels of the linear-algebra layer: ELL and Q1
// products/defects, one-pass reductions and BLAS-1 elementwise ops.
//
// Numerics: the translation unit is compiled with -ffp-contract=off, so a
// product and the sum it feeds are rounded separately, exactly as in the
// reference's tags::CPU loops; per-row sums run in the same order as there
// (ascending ELL slot, Q1 order DD,LL,LD,LU,DL,DU,UL,UD,UU), which makes
// the products, defects and elementwise ops bit-identical to tags::CPU.
// Reductions use a fixed tree whose shape depends only on the length, so
// they are reproducible run to run (not equal to the sequential CPU sum).
#pragma once

#include "common.hh"

namespace hsolve
{
namespace k
{
    // ---------------------------------------------------------------- loads
    template <typename T> __device__ __forceinline__ T ld_stream(const T * p)
    {
        return __builtin_nontemporal_load(p);
    }

    // ------------------------------------------------------------------ ELL
    // Column index decoding: unsigned index types hold the column itself,
    // signed ones (int16_t) the offset of the column from the row, a
    // backend-internal compression chosen at upload when every stored
    // column lies within +-32767 of its row.
    template <typename IT> __device__ __forceinline__ uint64_t col_of(IT a, uint64_t row)
    {
        if constexpr (IT(-1) < IT(0))
            return (uint64_t)((int64_t)row + (int64_t)a);
        else
            return (uint64_t)a;
    }

    // Sum over the first `len` slots of one row (threads == 1 layout), in
    // ascending slot order as product.hh:192-202.  MAXK (compile-time bound
    // on num_cols_per_row) lets every slot's loads be issued before the
    // dependent sum, for memory-level parallelism.
    template <typename DT, typename IT, int MAXK, typename XV>
    __device__ __forceinline__ DT ell_row_sum(const XV & x, const IT * __restrict__ aj,
            const DT * __restrict__ ax, uint32_t len, uint64_t row, uint64_t stride)
    {
        DT sum = DT(0);
        if constexpr (MAXK > 0)
        {
            DT a[MAXK > 0 ? MAXK : 1];
            IT c[MAXK > 0 ? MAXK : 1];
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
            {
                if ((uint32_t)k < len)
                {
                    a[k] = ld_stream(ax + (uint64_t)k * stride);
                    c[k] = ld_stream(aj + (uint64_t)k * stride);
                }
            }
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
            {
                if ((uint32_t)k < len)
                    sum += a[k] * x[col_of<IT>(c[k], row)];
            }
        }
        else
        {
            for (uint32_t k = 0; k < len; ++k)
                sum += ld_stream(ax + (uint64_t)k * stride) * x[col_of<IT>(ld_stream(aj + (uint64_t)k * stride), row)];
        }
        return sum;
    }

    // One lane per row.  MODE 0: y = A x; 1: y = rhs - A x;
    // 2: y = x + (rhs - A x) * dinv (damped-Jacobi sweep: Defect, then
    //    ScaledSum(x, temp, dinv), jacobi.hh:268-279).
    template <typename DT, typename IT, typename LT, int MAXK, int MODE>
    __global__ __launch_bounds__(256) void ell_spmv_t1(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            const DT * __restrict__ dinv, uint64_t row_start, uint64_t row_end, uint64_t stride)
    {
        const uint64_t row = row_start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= row_end)
            return;
        const uint32_t len = (uint32_t)Arl[row];
        const DT sum = ell_row_sum<DT, IT, MAXK>(x, Aj + row, Ax + row, len, row, stride);
        if (MODE == 0)
            y[row] = sum;
        else if (MODE == 1)
            y[row] = rhs[row] - sum;
        else
        {
            const DT t = rhs[row] - sum;
            y[row] = x[row] + t * dinv[row];
        }
    }

    // XCD-aware block order: the hardware deals consecutive workgroups to the
    // 8 XCDs round-robin; this bijection hands every XCD one contiguous range
    // of logical blocks, so the x entries its blocks gather stay in its L2.
    __device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
    {
        const uint32_t q = nb >> 3, rem = nb & 7u;
        const uint32_t xcd = b & 7u, pos = b >> 3;
        return xcd * q + (xcd < rem ? xcd : rem) + pos;
    }

    template <typename T> struct Vec2;
    template <> struct Vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };
    template <> struct Vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
    template <> struct Vec2<int16_t> { typedef short type __attribute__((ext_vector_type(2))); };
    template <> struct Vec2<uint32_t> { typedef unsigned type __attribute__((ext_vector_type(2))); };

    // Two consecutive rows per lane (threads == 1 layout): the value and
    // index streams are read as 2-element vectors, halving the load
    // instructions per byte.  FULL loads every slot below num_cols_per_row
    // without waiting for the row lengths (for matrices whose rows are
    // nearly all full, e.g. stencils); the sums still stop at each row's
    // length, in ascending slot order.  Requires row_start even and the
    // slot arrays readable at row_end (stride > rows or rows even).
    // STENCIL (9-point structured matrices on a side-m grid, MAXK == 9): a
    // row with all 9 slots stored holds exactly the neighbours (y + dy,
    // x + dx) in ascending order, so its columns are row + offset and its
    // index slots are not read; only shorter (boundary) rows read them.
    template <typename DT, typename IT, typename LT, int MAXK, int MODE, bool FULL, bool XCD, bool STENCIL = false>
    __global__ __launch_bounds__(256) void ell_spmv_r2(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            const DT * __restrict__ dinv, uint64_t row_start, uint64_t row_end, uint64_t stride,
            uint32_t ncpr, int64_t m = 0)
    {
        typedef typename Vec2<DT>::type V;
        typedef typename Vec2<IT>::type VI;
        const uint32_t b = XCD ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
        const uint64_t r0 = row_start + ((uint64_t)b * blockDim.x + threadIdx.x) * 2;
        if (r0 >= row_end)
            return;
        const bool two = r0 + 1 < row_end;
        // the defect / sweep operands of both rows load up front (the second
        // row's index clamped to the first when there is none), not after
        // the sums
        const uint64_t r1 = two ? r0 + 1 : r0;
        DT b0 = DT(0), b1 = DT(0), x0 = DT(0), x1 = DT(0), d0 = DT(0), d1 = DT(0);
        if (MODE != 0)
        {
            b0 = rhs[r0];
            b1 = rhs[r1];
        }
        if (MODE == 2)
        {
            x0 = x[r0];
            x1 = x[r1];
            d0 = dinv[r0];
            d1 = dinv[r1];
        }
        const uint32_t len0 = (uint32_t)Arl[r0];
        const uint32_t len1 = two ? (uint32_t)Arl[r0 + 1] : 0u;
        const uint32_t lmax = len0 > len1 ? len0 : len1;
        // STENCIL: both rows full, columns from the offsets
        const bool implied = STENCIL && len0 == 9 && (len1 == 9 || ! two);
        DT a0[MAXK], a1[MAXK];
        IT c0[MAXK], c1[MAXK];
        // A slot past the rows' lengths (past num_cols_per_row with FULL)
        // re-reads slot 0, a cached line, instead of branching, so every
        // slot load issues at once (num_cols_per_row >= 1 at launch).
        const uint32_t lim = FULL ? ncpr : lmax;
#pragma unroll
        for (int k = 0; k < MAXK; ++k)
        {
            const uint64_t kk = (uint32_t)k < lim ? (uint64_t)k : 0;
            const V v = __builtin_nontemporal_load((const V *)(Ax + r0 + kk * stride));
            a0[k] = v.x;
            a1[k] = v.y;
            if (! implied)
            {
                const VI c = __builtin_nontemporal_load((const VI *)(Aj + r0 + kk * stride));
                c0[k] = (IT)c.x;
                c1[k] = (IT)c.y;
            }
        }
        DT s0 = DT(0), s1 = DT(0);
        if (implied)
        {
            // the two rows' 3 x 3 neighbourhoods share three windows of four
            // x values (12 loads instead of 18); slot k = 3 (dy + 1) + dx + 1
            DT w[3][4];
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                const int64_t c = (int64_t)r0 + (int64_t)(d - 1) * m - 1;
                w[d][0] = x[c];
                w[d][1] = x[c + 1];
                w[d][2] = x[c + 2];
                w[d][3] = x[two ? c + 3 : c + 2];
            }
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
            {
                s0 += a0[k] * w[k / 3][k % 3];
                if (two)
                    s1 += a1[k] * w[k / 3][k % 3 + 1];
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
            {
                if ((uint32_t)k < len0)
                    s0 += a0[k] * x[col_of<IT>(c0[k], r0)];
                if ((uint32_t)k < len1)
                    s1 += a1[k] * x[col_of<IT>(c1[k], r0 + 1)];
            }
        }
        if (MODE == 0)
        {
            y[r0] = s0;
            if (two)
                y[r0 + 1] = s1;
        }
        else if (MODE == 1)
        {
            y[r0] = b0 - s0;
            if (two)
                y[r0 + 1] = b1 - s1;
        }
        else
        {
            const DT t0 = b0 - s0;
            y[r0] = x0 + t0 * d0;
            if (two)
            {
                const DT t1 = b1 - s1;
                y[r0 + 1] = x1 + t1 * d1;
            }
        }
    }

    // Rows per block of ell_spmv_lds (two per lane): 1024 for FP64, 512
    // for FP32 (measured on MI355X, Matrix B: FP64 437 -> 417 us with 1024
    // rows, whose x window per row is half as large; FP32 245 -> 269 us).
    template <typename DT> constexpr int lds_rows() { return sizeof(DT) == 8 ? 1024 : 512; }

    // ell_spmv_r2 for matrices whose stored columns lie within +-maxoff of
    // their row (banded but unstructured, e.g. random columns in a window):
    // the block first copies x[w0, w1), the columns its rows can touch
    // (w0 rounded down to a 16-byte boundary), into LDS with direct
    // global-to-LDS loads, and the rows gather from there instead of sending
    // one scattered L2 request per nonzero.  Blocks in XCD-contiguous order,
    // so neighbouring windows overlap in one L2.  Dynamic LDS: (kLdsRows +
    // 2 maxoff) * sizeof(DT) + 16 bytes; x must be 16-byte aligned.  Same
    // sums, in the same order, as ell_spmv_r2.
    //
    // SORTED: the arrays hold each block's rows in descending order
    // of length (a copy built at upload) and perm[p] is the block-local
    // original row of storage position p.  Slot k of a wave's rows is then
    // used by a prefix of its lanes, so the cache lines of slots past a
    // row's length are not fetched; each row keeps its own slot order and
    // its result goes to its original index.
    template <typename DT, typename IT, typename LT, int MAXK, int MODE, bool SORTED = false>
    __global__ __launch_bounds__(lds_rows<DT>() / 2) void ell_spmv_lds(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            const DT * __restrict__ dinv, uint64_t rows, uint64_t cols, uint64_t stride, uint32_t ncpr,
            uint64_t maxoff, const uint16_t * __restrict__ perm, uint64_t blk0, uint64_t row_lo, uint64_t row_hi)
    {
        constexpr int kLdsRows = lds_rows<DT>();
        constexpr int kLdsThreads = kLdsRows / 2;
        extern __shared__ __align__(16) unsigned char lds_raw[];
        DT * xw = reinterpret_cast<DT *>(lds_raw);
        typedef typename Vec2<DT>::type V;
        typedef typename Vec2<IT>::type VI;
        // blocks blk0 .. blk0 + gridDim.x - 1; rows outside [row_lo, row_hi)
        // are computed with their block but not written
        const uint64_t rb0 = (blk0 + (uint64_t)xcd_block(blockIdx.x, gridDim.x)) * kLdsRows;
        constexpr uint32_t E = 16 / sizeof(DT);   // elements per 16-byte chunk
        const uint64_t w0 = (rb0 > maxoff ? rb0 - maxoff : 0) & ~(uint64_t)(E - 1);
        const uint64_t wend = rb0 + kLdsRows + maxoff;
        const uint64_t w1 = wend < cols ? wend : cols;
        // A block whose rows all lie past cols + maxoff (empty tail rows of a
        // matrix with rows > cols) has an empty window: w1 <= w0.
        const uint32_t wn = w1 > w0 ? (uint32_t)(w1 - w0) : 0u;
        const uint64_t r0 = rb0 + (uint64_t)threadIdx.x * 2;
        const bool live = r0 < rows;
        const bool two = r0 + 1 < rows;
        // Row lengths first, then every slot load of the lane's two rows at
        // once (a slot past both lengths re-reads slot 0, a cached line), then
        // the x window, which goes to LDS without passing through registers
        // (measured on Matrix B FP64: 90 VGPRs instead of 113, 409 us instead
        // of 422): all of it is in flight together.  No branch separates the
        // register loads (a lane past the last row reads the block's first
        // pair; Arl and perm hold one padding entry), so the compiler need not
        // drain them at a join.
        const uint64_t rr = live ? r0 : rb0;
        const uint32_t l0 = (uint32_t)Arl[rr], l1 = (uint32_t)Arl[rr + 1];
        uint64_t o0 = rr, o1 = rr + 1;
        if (SORTED)
        {
            // original rows of the two storage positions
            o0 = rb0 + perm[rr];
            o1 = rb0 + perm[rr + 1];
        }
        const uint32_t len0 = live ? l0 : 0u;
        const uint32_t len1 = two ? l1 : 0u;
        const uint32_t lmax = l0 > l1 ? l0 : l1;
        DT a0[MAXK], a1[MAXK];
        // the pair's two columns: one 32-bit word for 16-bit offsets (kept
        // packed until the gather), two words otherwise
        constexpr bool kPack = sizeof(IT) == 2;
        uint32_t cp[MAXK];
        IT c0[kPack ? 1 : MAXK], c1[kPack ? 1 : MAXK];
#pragma unroll
        for (int k = 0; k < MAXK; ++k)
        {
            const uint64_t kk = (uint32_t)k < lmax ? (uint64_t)k : 0;
            const V v = __builtin_nontemporal_load((const V *)(Ax + rr + kk * stride));
            a0[k] = v.x;
            a1[k] = v.y;
            if constexpr (kPack)
                cp[k] = __builtin_nontemporal_load((const uint32_t *)(Aj + rr + kk * stride));
            else
            {
                const VI c = __builtin_nontemporal_load((const VI *)(Aj + rr + kk * stride));
                c0[k] = (IT)c.x;
                c1[k] = (IT)c.y;
            }
        }
        // the x window straight into LDS: 16-byte chunks, one per lane and
        // load (global_load_lds_dwordx4 writes lane l's chunk at dst + 16 l);
        // the last wn % E elements by ordinary loads
        {
            const uint32_t nfull = wn / E;
            const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
            for (uint32_t g = wave; g * 64 < nfull; g += kLdsThreads / 64)
            {
                const uint32_t q = g * 64 + lane;
                if (q < nfull)
                    __builtin_amdgcn_global_load_lds(
                            (const void __attribute__((address_space(1))) *)(x + w0 + (uint64_t)q * E),
                            (void __attribute__((address_space(3))) *)(xw + (uint64_t)g * 64 * E), 16, 0, 0);
            }
            const uint32_t t = nfull * E + threadIdx.x;
            if (t < wn)
                xw[t] = x[w0 + t];
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        lds_barrier();
        if (! live)
            return;
        DT s0 = DT(0), s1 = DT(0);
#pragma unroll
        for (int k = 0; k < MAXK; ++k)
        {
            IT ca, cb;
            if constexpr (kPack)
            {
                ca = (IT)(uint16_t)(cp[k] & 0xFFFFu);
                cb = (IT)(uint16_t)(cp[k] >> 16);
            }
            else
            {
                ca = c0[k];
                cb = c1[k];
            }
            if ((uint32_t)k < len0)
                s0 += a0[k] * xw[col_of<IT>(ca, o0) - w0];
            if ((uint32_t)k < len1)
                s1 += a1[k] * xw[col_of<IT>(cb, o1) - w0];
        }
        const bool k0 = o0 >= row_lo && o0 < row_hi;
        const bool k1 = two && o1 >= row_lo && o1 < row_hi;
        if (MODE == 0)
        {
            if (k0)
                y[o0] = s0;
            if (k1)
                y[o1] = s1;
        }
        else if (MODE == 1)
        {
            if (k0)
                y[o0] = rhs[o0] - s0;
            if (k1)
                y[o1] = rhs[o1] - s1;
        }
        else
        {
            if (k0)
            {
                const DT t0 = rhs[o0] - s0;
                y[o0] = x[o0] + t0 * dinv[o0];
            }
            if (k1)
            {
                const DT t1 = rhs[o1] - s1;
                y[o1] = x[o1] + t1 * dinv[o1];
            }
        }
    }

    // Persistent form of ell_spmv_lds (same sums, same order, same results):
    // a fixed grid of workgroups walks the row blocks, each XCD a contiguous
    // range of them (consecutive blocks, whose windows overlap, run on one
    // L2 at about the same time).  While a workgroup works on one block, the
    // row lengths (and, SORTED, the original rows) of its next block are
    // already loading, so that block's slot and window loads issue at once
    // instead of after a round trip for the lengths.
    template <typename DT, typename IT, typename LT, int MAXK, int MODE, bool SORTED = false>
    __global__ __launch_bounds__(lds_rows<DT>() / 2) __attribute__((amdgpu_waves_per_eu(4, 8))) void ell_spmv_lds_p(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            const DT * __restrict__ dinv, uint64_t rows, uint64_t cols, uint64_t stride, uint32_t ncpr,
            uint64_t maxoff, const uint16_t * __restrict__ perm, uint32_t nblk)
    {
        constexpr int kLdsRows = lds_rows<DT>();
        constexpr int kLdsThreads = kLdsRows / 2;
        extern __shared__ __align__(16) unsigned char lds_raw[];
        DT * xw = reinterpret_cast<DT *>(lds_raw);
        typedef typename Vec2<DT>::type V;
        typedef typename Vec2<IT>::type VI;
        // this workgroup's XCD, its place among that XCD's workgroups, and
        // the XCD's range of blocks
        const uint32_t G = gridDim.x;
        const uint32_t xcd = blockIdx.x & 7u, pos = blockIdx.x >> 3;
        const uint32_t per_xcd = (G >> 3) + (xcd < (G & 7u) ? 1u : 0u);
        const uint32_t q = nblk >> 3, rem = nblk & 7u;
        const uint32_t bbeg = xcd * q + (xcd < rem ? xcd : rem);
        const uint32_t bcnt = q + (xcd < rem ? 1u : 0u);
        // lengths and block-local storage rows of the original rows of block
        // bbeg + i for this lane (lengths are at most 16: one byte each)
        auto meta = [&](uint32_t i, uint32_t & lens, uint32_t & loc) {
            const uint64_t rb0 = (uint64_t)(bbeg + i) * kLdsRows;
            const uint64_t r0 = rb0 + (uint64_t)threadIdx.x * 2;
            const uint64_t rr = r0 < rows ? r0 : rb0;
            lens = (uint32_t)Arl[rr] | ((uint32_t)Arl[rr + 1] << 8);
            loc = (uint32_t)(rr - rb0) | ((uint32_t)(rr + 1 - rb0) << 16);
            if (SORTED)
                loc = (uint32_t)perm[rr] | ((uint32_t)perm[rr + 1] << 16);
        };
        uint32_t nlens = 0, nloc = 0;
        if (pos < bcnt)
            meta(pos, nlens, nloc);
        for (uint32_t i = pos; i < bcnt; i += per_xcd)
        {
            const uint32_t lens = nlens, loc = nloc;
            if (i + per_xcd < bcnt)
                meta(i + per_xcd, nlens, nloc);
            const uint64_t rb0 = (uint64_t)(bbeg + i) * kLdsRows;
            const uint64_t w0 = rb0 > maxoff ? rb0 - maxoff : 0;
            const uint64_t wend = rb0 + kLdsRows + maxoff;
            const uint64_t w1 = wend < cols ? wend : cols;
            const uint32_t wn = w1 > w0 ? (uint32_t)(w1 - w0) : 0u;
            const uint64_t r0 = rb0 + (uint64_t)threadIdx.x * 2;
            const bool live = r0 < rows;
            const bool two = r0 + 1 < rows;
            const uint64_t rr = live ? r0 : rb0;
            const uint32_t l0 = lens & 0xffu, l1 = lens >> 8;
            const uint64_t o0 = rb0 + (loc & 0xffffu), o1 = rb0 + (loc >> 16);
            const uint32_t len0 = live ? l0 : 0u;
            const uint32_t len1 = two ? l1 : 0u;
            const uint32_t lmax = l0 > l1 ? l0 : l1;
            DT a0[MAXK], a1[MAXK];
            IT c0[MAXK], c1[MAXK];
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
            {
                const uint64_t kk = (uint32_t)k < lmax ? (uint64_t)k : 0;
                const V v = __builtin_nontemporal_load((const V *)(Ax + rr + kk * stride));
                const VI c = __builtin_nontemporal_load((const VI *)(Aj + rr + kk * stride));
                a0[k] = v.x;
                a1[k] = v.y;
                c0[k] = (IT)c.x;
                c1[k] = (IT)c.y;
            }
            // one pass of window loads covers the +-2048 window of a block
            // (FP64: 5120 values on 512 lanes; FP32: 4608 on 256); wider
            // windows take the loop below
            constexpr uint32_t kFill = sizeof(DT) == 8 ? 10 : 18;
            DT wv[kFill];
            if (wn != 0)
            {
#pragma unroll
                for (uint32_t f = 0; f < kFill; ++f)
                {
                    const uint32_t j = f * kLdsThreads + threadIdx.x;
                    wv[f] = x[w0 + (j < wn ? j : wn - 1)];
                }
            }
            // every lane is done with the previous block's window
            lds_barrier();
            if (wn != 0)
            {
#pragma unroll
                for (uint32_t f = 0; f < kFill; ++f)
                {
                    const uint32_t j = f * kLdsThreads + threadIdx.x;
                    if (j < wn)
                        xw[j] = wv[f];
                }
            }
            for (uint32_t base = kFill * kLdsThreads; base < wn; base += kFill * kLdsThreads)
            {
#pragma unroll
                for (uint32_t f = 0; f < kFill; ++f)
                {
                    const uint32_t j = base + f * kLdsThreads + threadIdx.x;
                    wv[f] = x[w0 + (j < wn ? j : wn - 1)];
                }
#pragma unroll
                for (uint32_t f = 0; f < kFill; ++f)
                {
                    const uint32_t j = base + f * kLdsThreads + threadIdx.x;
                    if (j < wn)
                        xw[j] = wv[f];
                }
            }
            lds_barrier();
            if (live)
            {
                DT s0 = DT(0), s1 = DT(0);
#pragma unroll
                for (int k = 0; k < MAXK; ++k)
                {
                    if ((uint32_t)k < len0)
                        s0 += a0[k] * xw[col_of<IT>(c0[k], o0) - w0];
                    if ((uint32_t)k < len1)
                        s1 += a1[k] * xw[col_of<IT>(c1[k], o1) - w0];
                }
                if (MODE == 0)
                {
                    y[o0] = s0;
                    if (two)
                        y[o1] = s1;
                }
                else if (MODE == 1)
                {
                    y[o0] = rhs[o0] - s0;
                    if (two)
                        y[o1] = rhs[o1] - s1;
                }
                else
                {
                    const DT t0 = rhs[o0] - s0;
                    y[o0] = x[o0] + t0 * dinv[o0];
                    if (two)
                    {
                        const DT t1 = rhs[o1] - s1;
                        y[o1] = x[o1] + t1 * dinv[o1];
                    }
                }
            }
        }
    }

    // ELL with threads > 1: still one lane per row, walking k then t so the
    // summation order is the reference's (product.hh:192-202).
    template <typename DT, typename IT, typename LT, int MODE>
    __global__ __launch_bounds__(256) void ell_spmv_tn(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            const DT * __restrict__ dinv, uint64_t row_start, uint64_t row_end, uint64_t stride,
            uint32_t threads)
    {
        const uint64_t row = row_start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= row_end)
            return;
        const uint32_t len = (uint32_t)Arl[row];
        uint64_t j = row * threads;
        DT sum = DT(0);
        for (uint32_t k = 0; k < len; ++k, j += stride)
            for (uint32_t t = 0; t < threads; ++t)
                sum += Ax[j + t] * x[col_of<IT>(Aj[j + t], row)];
        if (MODE == 0)
            y[row] = sum;
        else if (MODE == 1)
            y[row] = rhs[row] - sum;
        else
        {
            const DT t = rhs[row] - sum;
            y[row] = x[row] + t * dinv[row];
        }
    }

    // ------------------------------------------------------------------ Q1
    struct Bands
    {
        const void * b[9];   // LL LD LU DL DD DU UL UD UU
    };

    // Row i of a Q1 product.  ORD 0: Product<Q1>'s order (DD first, then
    // LL, LD, LU, DL, DU, UL, UD, UU; product.hh:370-403).  ORD 1: the
    // order of an ELL row with ascending columns (LL .. UU, summed from
    // zero), for bands that hold a 9-point structured ELL matrix with zeros
    // at its absent entries: a zero entry adds +-0 to a sum that is never
    // -0, so with finite x the result is the ELL product's, bit for bit.
    // x: a pointer, or any value whose x[j] yields the j-th entry
    template <typename DT, int ORD = 0, typename XV>
    __device__ __forceinline__ DT q1_row(const Bands & B, const XV & x, int64_t i, int64_t n, int64_t m)
    {
        const DT * const * b = reinterpret_cast<const DT * const *>(B.b);
        if (ORD == 1)
        {
            DT s = DT(0);
            if (i - m - 1 >= 0) s += ld_stream(b[0] + i) * x[i - m - 1];
            if (i - m >= 0)     s += ld_stream(b[1] + i) * x[i - m];
            if (i - m + 1 >= 0) s += ld_stream(b[2] + i) * x[i - m + 1];
            if (i - 1 >= 0)     s += ld_stream(b[3] + i) * x[i - 1];
            s += ld_stream(b[4] + i) * x[i];
            if (i + 1 < n)      s += ld_stream(b[5] + i) * x[i + 1];
            if (i + m - 1 < n)  s += ld_stream(b[6] + i) * x[i + m - 1];
            if (i + m < n)      s += ld_stream(b[7] + i) * x[i + m];
            if (i + m + 1 < n)  s += ld_stream(b[8] + i) * x[i + m + 1];
            return s;
        }
        DT s = ld_stream(b[4] + i) * x[i];
        if (i - m - 1 >= 0) s += ld_stream(b[0] + i) * x[i - m - 1];
        if (i - m >= 0)     s += ld_stream(b[1] + i) * x[i - m];
        if (i - m + 1 >= 0) s += ld_stream(b[2] + i) * x[i - m + 1];
        if (i - 1 >= 0)     s += ld_stream(b[3] + i) * x[i - 1];
        if (i + 1 < n)      s += ld_stream(b[5] + i) * x[i + 1];
        if (i + m - 1 < n)  s += ld_stream(b[6] + i) * x[i + m - 1];
        if (i + m < n)      s += ld_stream(b[7] + i) * x[i + m];
        if (i + m + 1 < n)  s += ld_stream(b[8] + i) * x[i + m + 1];
        return s;
    }

    // mode 0: y = A x; mode 1: y = rhs - A x;
    // mode 2: y = x + (rhs - A x) * dinv (one Jacobi sweep into y)
    template <typename DT, int MODE>
    __global__ __launch_bounds__(256) void q1_kernel(Bands B, const DT * __restrict__ x,
            DT * __restrict__ y, const DT * __restrict__ rhs, const DT * __restrict__ dinv,
            int64_t n, int64_t m, const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const DT s = q1_row<DT>(B, x, i, n, m);
        if (MODE == 0)
            y[i] = s;
        else if (MODE == 1)
            y[i] = rhs[i] - s;
        else
        {
            const DT t = rhs[i] - s;
            y[i] = x[i] + t * dinv[i];
        }
    }

    // 16-byte vectors: 4 floats or 2 doubles per lane
    template <typename DT> struct Vec16;
    template <> struct Vec16<float>
    {
        static constexpr int V = 4;
        typedef float type __attribute__((ext_vector_type(4)));
    };
    template <> struct Vec16<double>
    {
        static constexpr int V = 2;
        typedef double type __attribute__((ext_vector_type(2)));
    };

    // Q1 row sums of V consecutive rows i0..i0+V-1 that lie away from both
    // ends of the vector (every index guard of q1_row holds): bands come in
    // as one 16-byte load each, x from the three row windows.  Each row's
    // sum runs in q1_row's order.
    template <typename DT, int ORD = 0>
    __device__ __forceinline__ void q1_rows_v(const Bands & B, const DT * __restrict__ x, int64_t i0, int64_t m,
                                              DT (&out)[Vec16<DT>::V])
    {
        constexpr int V = Vec16<DT>::V;
        typedef typename Vec16<DT>::type VT;
        const DT * const * b = reinterpret_cast<const DT * const *>(B.b);
        VT bv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k)
            bv[k] = __builtin_nontemporal_load((const VT *)(b[k] + i0));
        DT xm[V + 2], xc[V + 2], xp[V + 2];
#pragma unroll
        for (int j = 0; j < V + 2; ++j)
        {
            xm[j] = x[i0 - m - 1 + j];
            xc[j] = x[i0 - 1 + j];
            xp[j] = x[i0 + m - 1 + j];
        }
        if (ORD == 1)
        {
#pragma unroll
            for (int v = 0; v < V; ++v)
            {
                DT s = DT(0);
                s += bv[0][v] * xm[v];
                s += bv[1][v] * xm[v + 1];
                s += bv[2][v] * xm[v + 2];
                s += bv[3][v] * xc[v];
                s += bv[4][v] * xc[v + 1];
                s += bv[5][v] * xc[v + 2];
                s += bv[6][v] * xp[v];
                s += bv[7][v] * xp[v + 1];
                s += bv[8][v] * xp[v + 2];
                out[v] = s;
            }
            return;
        }
#pragma unroll
        for (int v = 0; v < V; ++v)
        {
            DT s = bv[4][v] * xc[v + 1];
            s += bv[0][v] * xm[v];
            s += bv[1][v] * xm[v + 1];
            s += bv[2][v] * xm[v + 2];
            s += bv[3][v] * xc[v];
            s += bv[5][v] * xc[v + 2];
            s += bv[6][v] * xp[v];
            s += bv[7][v] * xp[v + 1];
            s += bv[8][v] * xp[v + 2];
            out[v] = s;
        }
    }

    // Vectorised form of q1_kernel: V = 16 / sizeof(DT) rows per lane; lane
    // `lane` covers rows lane*V .. lane*V+V-1.  Needs 16-byte aligned bands,
    // y, rhs and dinv (checked by the launchers).
    template <typename DT, int MODE, int ORD = 0>
    __device__ __forceinline__ void q1_body_v(const Bands & B, const DT * __restrict__ x,
            DT * __restrict__ y, const DT * __restrict__ rhs, const DT * __restrict__ dinv,
            int64_t n, int64_t m, int64_t lane)
    {
        constexpr int V = Vec16<DT>::V;
        typedef typename Vec16<DT>::type VT;
        const int64_t i0 = lane * V;
        if (i0 >= n)
            return;
        if (i0 - m - 1 < 0 || i0 + V + m >= n)
        {
            for (int v = 0; v < V && i0 + v < n; ++v)
            {
                const int64_t i = i0 + v;
                const DT s = q1_row<DT, ORD>(B, x, i, n, m);
                if (MODE == 0)
                    y[i] = s;
                else if (MODE == 1)
                    y[i] = rhs[i] - s;
                else
                {
                    const DT t = rhs[i] - s;
                    y[i] = x[i] + t * dinv[i];
                }
            }
            return;
        }
        DT s[V];
        q1_rows_v<DT, ORD>(B, x, i0, m, s);
        VT o;
        if (MODE == 0)
        {
#pragma unroll
            for (int v = 0; v < V; ++v)
                o[v] = s[v];
        }
        else
        {
            const VT r = __builtin_nontemporal_load((const VT *)(rhs + i0));
            if (MODE == 1)
            {
#pragma unroll
                for (int v = 0; v < V; ++v)
                    o[v] = r[v] - s[v];
            }
            else
            {
                const VT di = __builtin_nontemporal_load((const VT *)(dinv + i0));
#pragma unroll
                for (int v = 0; v < V; ++v)
                {
                    const DT t = r[v] - s[v];
                    o[v] = x[i0 + v] + t * di[v];
                }
            }
        }
        *(VT *)(y + i0) = o;
    }

    template <typename DT, int MODE>
    __global__ __launch_bounds__(256) void q1_kernel_v(Bands B, const DT * __restrict__ x,
            DT * __restrict__ y, const DT * __restrict__ rhs, const DT * __restrict__ dinv,
            int64_t n, int64_t m, const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        q1_body_v<DT, MODE>(B, x, y, rhs, dinv, n, m, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    }

    inline bool aligned16(const void * p) { return ((uintptr_t)p & 15u) == 0; }

    // Launches the vectorised Q1 kernel when every streamed array is 16-byte
    // aligned, else the one-row-per-lane form.  MODE as q1_kernel.
    template <typename DT, int MODE>
    inline void q1_launch(const Bands & B, const DT * x, DT * y, const DT * rhs, const DT * dinv, int64_t n,
                          int64_t m, const int * active, hipStream_t st)
    {
        bool vec = aligned16(y) && (MODE == 0 || aligned16(rhs)) && (MODE != 2 || aligned16(dinv));
        for (int k = 0; k < 9; ++k)
            vec = vec && aligned16(B.b[k]);
        if (vec)
        {
            constexpr int V = Vec16<DT>::V;
            const uint64_t lanes = ((uint64_t)n + V - 1) / V;
            hipLaunchKernelGGL((q1_kernel_v<DT, MODE>), dim3(grid_for(lanes, 256)), dim3(256), 0, st, B, x, y, rhs,
                               dinv, n, m, active);
        }
        else
            hipLaunchKernelGGL((q1_kernel<DT, MODE>), dim3(grid_for((uint64_t)n, 256)), dim3(256), 0, st, B, x, y, rhs,
                               dinv, n, m, active);
    }

    // ------------------------------------------------------------ reductions
    __device__ __forceinline__ double wave_sum(double v)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off, 64);
        return v;
    }

    __device__ __forceinline__ double block_sum(double v, double * lds)
    {
        v = wave_sum(v);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0)
            lds[wid] = v;
        __syncthreads();
        double r = 0.0;
        if (threadIdx.x == 0)
        {
            const int nw = blockDim.x >> 6;
            for (int w = 0; w < nw; ++w)
                r += lds[w];
        }
        __syncthreads();
        return r;   // valid in thread 0
    }

    // Two sums at once: wave shuffles, then the waves' totals in order.
    // Valid in thread 0 on return.
    __device__ __forceinline__ void block_sum2(double & a, double & b, double * lds)
    {
        a = wave_sum(a);
        b = wave_sum(b);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0)
        {
            lds[2 * wid] = a;
            lds[2 * wid + 1] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double ra = 0.0, rb = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            {
                ra += lds[2 * w];
                rb += lds[2 * w + 1];
            }
            a = ra;
            b = rb;
        }
        __syncthreads();
    }

    // Two-value reduction over the blocks of a launch, in two launches: each
    // block stores its pair of block sums (put_partials2: wave shuffles, then
    // the waves in order; no fence and no atomic), and a following one-block
    // launch of 256 threads forms the totals (sum_partials2: thread-strided
    // sums of the partials in block order, then the block tree), so the
    // result depends only on the grid.  The kernel boundary orders the
    // partial stores before the final sum.  (A last-block finish instead
    // needs a device-scope fence per block, which costs ~50 ns of serialised
    // time on gfx950 and dominates a reduction over thousands of blocks.)
    __device__ __forceinline__ void put_partials2(double a, double b, double * partials, unsigned bx)
    {
        __shared__ double lds[32];
        block_sum2(a, b, lds);
        if (threadIdx.x == 0)
        {
            partials[2 * bx] = a;
            partials[2 * bx + 1] = b;
        }
    }
    // totals over nb partial pairs; valid on thread 0
    __device__ __forceinline__ void sum_partials2(const double * partials, unsigned nb, double & a, double & b)
    {
        __shared__ double lds[32];
        double va = 0.0, vb = 0.0;
        for (unsigned i = threadIdx.x; i < nb; i += blockDim.x)
        {
            va += partials[2 * i];
            vb += partials[2 * i + 1];
        }
        block_sum2(va, vb, lds);
        a = va;
        b = vb;
    }

    // sum_i a_i * b_i (b == nullptr: sum a_i^2), accumulated in double:
    // block sums here, the total in dot_finish_kernel (one block: thread-
    // strided sums of the partials in index order, then the block tree).
    template <typename DT>
    __global__ __launch_bounds__(256) void dot_kernel(const DT * __restrict__ a,
            const DT * __restrict__ b, uint64_t n, double * partials)
    {
        __shared__ double lds[16];
        double v = 0.0;
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const double ai = (double)a[i];
            v += ai * (b ? (double)b[i] : ai);
        }
        const double bt = block_sum(v, lds);
        if (threadIdx.x == 0)
            partials[blockIdx.x] = bt;
    }

    template <typename DT>
    __global__ __launch_bounds__(256) void dot_finish_kernel(const double * partials, unsigned nb, DT * out, int root)
    {
        __shared__ double lds[16];
        double v = 0.0;
        for (unsigned b = threadIdx.x; b < nb; b += blockDim.x)
            v += partials[b];
        const double total = block_sum(v, lds);
        if (threadIdx.x == 0)
            *out = root ? (DT)sqrt(total) : (DT)total;
    }

    // ------------------------------------------------------------ BLAS-1
    // op codes of the elementwise kernel
    enum Elem
    {
        E_SCALED_SUM_S = 0,   // r = x + b*y
        E_SCALED_SUM_V,       // x = x + y*z
        E_SUM,                // x = x + y
        E_SCALE,              // x = x*b
        E_DIFF2,              // x = x - y
        E_DIFF3,              // r = x - y
        E_EPROD3,             // r = x*y
        E_EINV,               // x = 1/x (x != 0)
        E_FILL                // r = b
    };

    template <typename DT, int OP>
    __global__ __launch_bounds__(256) void elem_kernel(DT * r, const DT * x,
            const DT * y, const DT * z, DT b, uint64_t n)
    {
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            if (OP == E_SCALED_SUM_S)
                r[i] = b * y[i] + x[i];
            else if (OP == E_SCALED_SUM_V)
                r[i] = r[i] + y[i] * z[i];
            else if (OP == E_SUM)
                r[i] = r[i] + y[i];
            else if (OP == E_SCALE)
                r[i] = r[i] * b;
            else if (OP == E_DIFF2)
                r[i] = r[i] - y[i];
            else if (OP == E_DIFF3)
                r[i] = x[i] - y[i];
            else if (OP == E_EPROD3)
                r[i] = x[i] * y[i];
            else if (OP == E_EINV)
            {
                const DT v = r[i];
                if (v != DT(0))
                    r[i] = DT(1) / v;
            }
            else if (OP == E_FILL)
                r[i] = b;
        }
    }

    template <typename SRC, typename DST>
    __global__ __launch_bounds__(256) void convert_kernel(const SRC * __restrict__ src,
            DST * __restrict__ dst, uint64_t n)
    {
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            dst[i] = (DST)src[i];
    }

    static __global__ __launch_bounds__(256) void copy16_kernel(const uint4 * __restrict__ src,
            uint4 * __restrict__ dst, uint64_t n)
    {
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            dst[i] = src[i];
    }

    static __global__ __launch_bounds__(256) void zero_bytes_kernel(uint8_t * __restrict__ p, uint64_t bytes)
    {
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += stride)
            p[i] = 0;
    }
}
}
