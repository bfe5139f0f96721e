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

    // ------------------------------------------------------------ ELL (T=1)
    // One lane per row.  Slots are walked in ascending order up to the row's
    // length; MAXK (compile-time bound on num_cols_per_row) lets the loads of
    // all slots be issued before the dependent sum, for memory-level
    // parallelism, while the sum itself keeps the reference order.
    template <typename DT, typename IT, typename LT, int MAXK, bool DEFECT>
    __global__ __launch_bounds__(256) void ell_spmv_t1(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            uint64_t row_start, uint64_t row_end, uint64_t stride)
    {
        const uint64_t row = row_start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= row_end)
            return;
        const uint32_t len = (uint32_t)Arl[row];
        const IT * aj = Aj + row;
        const DT * ax = Ax + row;
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
                    sum += a[k] * x[c[k]];
            }
        }
        else
        {
            for (uint32_t k = 0; k < len; ++k)
                sum += ld_stream(ax + (uint64_t)k * stride) * x[ld_stream(aj + (uint64_t)k * stride)];
        }
        y[row] = DEFECT ? rhs[row] - sum : sum;
    }

    // ELL with threads > 1: still one lane per row, walking k then t so the
    // summation order is the reference's (product.hh:192-202).
    template <typename DT, typename IT, typename LT, bool DEFECT>
    __global__ __launch_bounds__(256) void ell_spmv_tn(
            const DT * __restrict__ x, DT * __restrict__ y, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            uint64_t row_start, uint64_t row_end, uint64_t stride, uint32_t threads)
    {
        const uint64_t row = row_start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= row_end)
            return;
        const uint32_t len = (uint32_t)Arl[row];
        uint64_t j = row * threads;
        DT sum = DT(0);
        for (uint32_t k = 0; k < len; ++k, j += stride)
            for (uint32_t t = 0; t < threads; ++t)
                sum += Ax[j + t] * x[Aj[j + t]];
        y[row] = DEFECT ? rhs[row] - sum : sum;
    }

    // ELL damped-Jacobi sweep: x_new = x_old + (rhs - A x_old) * dinv
    // (Defect, then ScaledSum(x, temp, dinv), jacobi.hh:268-279).
    template <typename DT, typename IT, typename LT, int MAXK>
    __global__ __launch_bounds__(256) void ell_jacobi_t1(
            const DT * __restrict__ x_old, DT * __restrict__ x_new, const IT * __restrict__ Aj,
            const DT * __restrict__ Ax, const LT * __restrict__ Arl, const DT * __restrict__ rhs,
            const DT * __restrict__ dinv, uint64_t rows, uint64_t stride)
    {
        const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= rows)
            return;
        const uint32_t len = (uint32_t)Arl[row];
        DT sum = DT(0);
        if constexpr (MAXK > 0)
        {
            DT a[MAXK > 0 ? MAXK : 1];
            IT c[MAXK > 0 ? MAXK : 1];
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
                if ((uint32_t)k < len)
                {
                    a[k] = ld_stream(Ax + row + (uint64_t)k * stride);
                    c[k] = ld_stream(Aj + row + (uint64_t)k * stride);
                }
#pragma unroll
            for (int k = 0; k < MAXK; ++k)
                if ((uint32_t)k < len)
                    sum += a[k] * x_old[c[k]];
        }
        else
        {
            for (uint32_t k = 0; k < len; ++k)
                sum += Ax[row + (uint64_t)k * stride] * x_old[Aj[row + (uint64_t)k * stride]];
        }
        const DT t = rhs[row] - sum;
        x_new[row] = x_old[row] + t * dinv[row];
    }

    // ------------------------------------------------------------------ Q1
    struct Bands
    {
        const void * b[9];   // LL LD LU DL DD DU UL UD UU
    };

    template <typename DT>
    __device__ __forceinline__ DT q1_row(const Bands & B, const DT * __restrict__ x, int64_t i,
                                         int64_t n, int64_t m)
    {
        const DT * const * b = reinterpret_cast<const DT * const *>(B.b);
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

    // Finishes a one-pass reduction: the block that takes the last ticket
    // sums all block partials in index order and writes the result.
    template <typename RT, typename F>
    __device__ __forceinline__ void finish_reduce(double block_total, double * partials,
            unsigned * ticket, F && write_result)
    {
        __shared__ double lds[16];
        __shared__ bool last;
        if (threadIdx.x == 0)
        {
            partials[blockIdx.x] = block_total;
            __threadfence();
            const unsigned t = atomicAdd(ticket, 1u);
            last = (t == gridDim.x - 1);
        }
        __syncthreads();
        if (! last)
            return;
        __threadfence();
        double v = 0.0;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
            v += __builtin_nontemporal_load(partials + b);
        const double total = block_sum(v, lds);
        if (threadIdx.x == 0)
        {
            write_result(total);
            *ticket = 0u;
        }
    }

    // sum_i a_i * b_i (b == nullptr: sum a_i^2), accumulated in double.
    template <typename DT>
    __global__ __launch_bounds__(256) void dot_kernel(const DT * __restrict__ a,
            const DT * __restrict__ b, uint64_t n, double * partials, unsigned * ticket,
            DT * out, int root)
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
        finish_reduce<DT>(bt, partials, ticket, [&](double total) {
            *out = root ? (DT)sqrt(total) : (DT)total;
        });
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
