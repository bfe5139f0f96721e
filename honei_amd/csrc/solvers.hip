// solvers.hip -- device-resident solvers of the MI355X backend:
//   * CGSolver<Tag, methods::NONE> / <Tag, methods::VAR> (honei/math/cg.hh)
//     on Q1 bands or an ELL handle, FP32 or FP64;
//   * Multigrid<Tag, OuterTag, NONE, JAC, CYCLE::V, MIXED> on Q1 bands
//     (honei/math/multigrid.hh:1616-2393): FP32 V-cycles inside an FP64
//     iterative-refinement loop.
//
// Design: every scalar of the iteration (alpha, lambda, norms, iteration
// counts, convergence flags) lives on the device.  Each reduction is fused
// into the kernel that produces its operand and finished by the last block
// to arrive, which also performs the scalar update and the stopping test.
// Every kernel of a step first reads a `run` flag and returns if the solve
// has already stopped, so the host enqueues whole chunks of iterations and
// synchronises once per chunk, while the iteration count stays exactly the
// reference's.
#include "ell_handle.hh"
#include "kernels_mg.hh"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace hsolve
{
namespace
{
    constexpr int kSolverBlocks = 2048;   // fixed grid of the fused reduction kernels

    // --------------------------------------------------------------- operators
    // A row of y = A x for the fused kernels.
    template <typename DT> struct Q1Op
    {
        k::Bands B;
        int64_t n, m;
        __device__ __forceinline__ DT row(const DT * __restrict__ x, int64_t i) const
        {
            return k::q1_row<DT>(B, x, i, n, m);
        }
    };

    template <typename DT, typename IT, typename LT> struct EllOp
    {
        const IT * Aj;
        const DT * Ax;
        const LT * Arl;
        uint64_t stride;
        uint32_t threads;
        __device__ __forceinline__ DT row(const DT * __restrict__ x, int64_t i) const
        {
            const uint32_t len = (uint32_t)Arl[i];
            if (threads == 1)
                return k::ell_row_sum<DT, IT, 0>(x, Aj + i, Ax + i, len, (uint64_t)i, stride);
            DT sum = DT(0);
            uint64_t j = (uint64_t)i * threads;
            for (uint32_t k = 0; k < len; ++k, j += stride)
                for (uint32_t t = 0; t < threads; ++t)
                    sum += Ax[j + t] * x[k::col_of<IT>(Aj[j + t], (uint64_t)i)];
            return sum;
        }
    };

    // --------------------------------------------------------------- reductions
    __device__ __forceinline__ void block_sum2(double & a, double & b, double * lds)
    {
        a = k::wave_sum(a);
        b = k::wave_sum(b);
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

    // Last-block finish of a two-value reduction; `fin(a, b)` runs on one
    // thread of the last block with the totals.
    template <typename F>
    __device__ __forceinline__ void finish2(double a, double b, double * partials, unsigned * ticket, F && fin)
    {
        __shared__ double lds[32];
        __shared__ bool last;
        block_sum2(a, b, lds);
        if (threadIdx.x == 0)
        {
            partials[2 * blockIdx.x] = a;
            partials[2 * blockIdx.x + 1] = b;
            __threadfence();
            last = (atomicAdd(ticket, 1u) == gridDim.x - 1);
        }
        __syncthreads();
        if (! last)
            return;
        __threadfence();
        double va = 0.0, vb = 0.0;
        for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x)
        {
            va += __builtin_nontemporal_load(partials + 2 * i);
            vb += __builtin_nontemporal_load(partials + 2 * i + 1);
        }
        block_sum2(va, vb, lds);
        if (threadIdx.x == 0)
        {
            fin(va, vb);
            *ticket = 0u;
        }
    }

    // --------------------------------------------------------------- CG state
    template <typename DT> struct CgScalars
    {
        DT alpha, alpha_old, alpha_new, temp, lambda, talpha;
        DT initial_defect, current_defect;
        DT eps_relative;
        uint64_t iters, max_iters;
        int run;          // 1 while iterating
        int converged;    // 1 when a stopping test fired
    };

    template <typename DT> struct Eps;
    template <> struct Eps<float> { static constexpr float v = 1.19209290e-07F; };
    template <> struct Eps<double> { static constexpr double v = 2.2204460492503131e-16; };

    template <typename DT> __device__ __forceinline__ DT guard(DT v)
    {
        return (fabs(v) > Eps<DT>::v) ? v : Eps<DT>::v;
    }
    template <typename DT> __device__ __forceinline__ DT guard_signed(DT v)
    {
        return (fabs(v) > Eps<DT>::v) ? v : Eps<DT>::v * (v / (DT)fabs(v));
    }

    // r = b - A x; p = r (NONE) or p = P*r (VAR); reductions r.r and r.p
    template <typename DT, typename Op, bool VAR>
    __global__ __launch_bounds__(256) void cg_init_kernel(Op A, const DT * __restrict__ b, const DT * __restrict__ x,
            DT * __restrict__ r, DT * __restrict__ p, const DT * __restrict__ P, int64_t n, double * partials,
            unsigned * ticket, CgScalars<DT> * s)
    {
        double acc_rr = 0.0, acc_rp = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const DT sum = A.row(x, i);
            const DT ri = b[i] - sum;
            r[i] = ri;
            const DT pi = VAR ? P[i] * ri : ri;
            p[i] = pi;
            acc_rr += (double)ri * (double)ri;
            acc_rp += (double)ri * (double)pi;
        }
        finish2(acc_rr, acc_rp, partials, ticket, [&](double rr, double rp) {
            s->iters = 0;
            s->converged = 0;
            s->run = (s->max_iters > 0) ? 1 : 0;
            if (VAR)
            {
                s->initial_defect = (DT)sqrt((DT)rr);
                s->alpha_new = (DT)rp;
            }
            else
            {
                s->alpha = (DT)rr;
                s->initial_defect = (DT)sqrt((DT)rr);
            }
            s->current_defect = s->initial_defect;
        });
    }

    // v = A p; temp = v.p; lambda
    template <typename DT, typename Op, bool VAR>
    __global__ __launch_bounds__(256) void cg_spmv_kernel(Op A, const DT * __restrict__ p, DT * __restrict__ v,
            int64_t n, double * partials, unsigned * ticket, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double acc = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const DT vi = A.row(p, i);
            v[i] = vi;
            acc += (double)vi * (double)p[i];
        }
        finish2(acc, 0.0, partials, ticket, [&](double vp, double) {
            const DT temp = (DT)vp;
            s->temp = temp;
            if (VAR)
                s->lambda = s->alpha_new / guard_signed<DT>(temp);
            else
                s->lambda = s->alpha / guard<DT>(temp);
        });
    }

    // x += lambda p; r += (-lambda) v; r.r (and r.(P r) for VAR); stopping test
    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void cg_update_kernel(DT * __restrict__ x, DT * __restrict__ r,
            const DT * __restrict__ p, const DT * __restrict__ v, const DT * __restrict__ P, int64_t n,
            double * partials, unsigned * ticket, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        const DT lambda = s->lambda;
        const DT mlambda = -lambda;
        double acc_rr = 0.0, acc_rz = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            x[i] = x[i] + lambda * p[i];
            const DT ri = r[i] + mlambda * v[i];
            r[i] = ri;
            acc_rr += (double)ri * (double)ri;
            if (VAR)
                acc_rz += (double)ri * (double)(P[i] * ri);
        }
        finish2(acc_rr, acc_rz, partials, ticket, [&](double rr, double rz) {
            const DT eps = s->eps_relative;
            if (VAR)
            {
                ++s->iters;
                const DT cur = (DT)sqrt((DT)rr);
                s->current_defect = cur;
                if (cur < eps * s->initial_defect || cur < eps)
                {
                    s->converged = 1;
                    s->run = 0;
                    return;
                }
                if (s->iters == s->max_iters)
                {
                    s->run = 0;
                    return;
                }
                s->alpha = s->alpha_new;
                s->alpha_new = (DT)rz;
                s->talpha = s->alpha_new / guard_signed<DT>(s->alpha);
            }
            else
            {
                s->alpha_old = s->alpha;
                s->alpha = (DT)rr;
                s->talpha = s->alpha / guard<DT>(s->alpha_old);
                ++s->iters;
                const DT cur = (DT)sqrt(s->alpha);
                s->current_defect = cur;
                if (cur < eps * s->initial_defect || cur < eps)
                {
                    s->converged = 1;
                    s->run = 0;
                }
                else if (s->iters == s->max_iters)
                    s->run = 0;
            }
        });
    }

    // p = p * talpha + r (NONE) or p = p * talpha + P*r (VAR)
    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void cg_direction_kernel(DT * __restrict__ p, const DT * __restrict__ r,
            const DT * __restrict__ P, int64_t n, const CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        const DT talpha = s->talpha;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const DT t = p[i] * talpha;
            p[i] = t + (VAR ? P[i] * r[i] : r[i]);
        }
    }

    // Device workspace of one CG solve.
    template <typename DT> struct CgWork
    {
        DT *r = nullptr, *p = nullptr, *v = nullptr;
        double * partials = nullptr;
        unsigned * ticket = nullptr;
        CgScalars<DT> * s = nullptr;
        void * block = nullptr;

        CgWork(uint64_t n, hipStream_t st)
        {
            const size_t vb = ((sizeof(DT) * n + 255) / 256) * 256;
            const size_t bytes = 3 * vb + sizeof(double) * 2 * kSolverBlocks + 256 + sizeof(CgScalars<DT>) + 256;
            hipError_t e = hipMalloc(&block, bytes);
            if (e != hipSuccess)
                throw Error(HSOLVE_ERR_ALLOC, "CG workspace allocation failed");
            char * c = (char *)block;
            r = (DT *)c; c += vb;
            p = (DT *)c; c += vb;
            v = (DT *)c; c += vb;
            partials = (double *)c; c += sizeof(double) * 2 * kSolverBlocks;
            ticket = (unsigned *)c; c += 256;
            s = (CgScalars<DT> *)c;
            HIP_CHECK(hipMemsetAsync(ticket, 0, 256, st));
        }
        ~CgWork()
        {
            if (block)
                (void)hipFree(block);
        }
    };

    template <typename DT, typename Op, bool VAR>
    void run_cg(const Op & A, uint64_t n, const DT * b, DT * x, const DT * P, uint64_t max_iters, double eps,
                uint64_t * used_iters, double * defect_out, hipStream_t st, int * converged_out)
    {
        CgWork<DT> w(n, st);
        CgScalars<DT> h;
        std::memset(&h, 0, sizeof(h));
        h.eps_relative = (DT)eps;
        h.max_iters = max_iters;
        HIP_CHECK(hipMemcpyAsync(w.s, &h, sizeof(h), hipMemcpyHostToDevice, st));
        const unsigned grid = std::min<unsigned>(grid_for(n, kBlock), kSolverBlocks);
        hipLaunchKernelGGL((cg_init_kernel<DT, Op, VAR>), dim3(grid), dim3(kBlock), 0, st, A, b, x, w.r, w.p, P,
                           (int64_t)n, w.partials, w.ticket, w.s);
        HIP_CHECK(hipGetLastError());
        // chunks of iterations between host checks; every kernel of an
        // iteration that starts after the stop is a no-op
        uint64_t enqueued = 0;
        uint64_t chunk = 8;
        while (true)
        {
            for (uint64_t c = 0; c < chunk && enqueued < max_iters; ++c, ++enqueued)
            {
                hipLaunchKernelGGL((cg_spmv_kernel<DT, Op, VAR>), dim3(grid), dim3(kBlock), 0, st, A, w.p, w.v,
                                   (int64_t)n, w.partials, w.ticket, w.s);
                hipLaunchKernelGGL((cg_update_kernel<DT, VAR>), dim3(grid), dim3(kBlock), 0, st, x, w.r, w.p, w.v, P,
                                   (int64_t)n, w.partials, w.ticket, w.s);
                hipLaunchKernelGGL((cg_direction_kernel<DT, VAR>), dim3(grid), dim3(kBlock), 0, st, w.p, w.r, P,
                                   (int64_t)n, (const CgScalars<DT> *)w.s);
            }
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(&h, w.s, sizeof(h), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            if (h.run == 0 || enqueued >= max_iters)
                break;
            if (chunk < 64)
                chunk *= 2;
        }
        if (used_iters)
            *used_iters = h.iters;
        if (defect_out)
            *defect_out = (h.initial_defect > 0) ? (double)h.current_defect / (double)h.initial_defect
                                                 : (double)h.current_defect;
        if (converged_out)
            *converged_out = h.converged;
    }

    k::Bands bands9(const void * const * b)
    {
        require(b != nullptr, HSOLVE_ERR_INVALID, "null band array");
        k::Bands B;
        for (int i = 0; i < 9; ++i)
        {
            require(b[i] != nullptr, HSOLVE_ERR_INVALID, "null band pointer");
            B.b[i] = b[i];
        }
        return B;
    }

    template <typename DT, bool VAR>
    void run_cg_ell(const hsolve_ell_s * m, const DT * b, DT * x, const DT * P, uint64_t max_iters, double eps,
                    uint64_t * used, double * defect, hipStream_t st, int * conv)
    {
        const uint64_t n = m->rows;
        with_ell_types(m, [&](auto Aj, auto Arl) {
            using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
            using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
            EllOp<DT, IT, LT> A{Aj, (const DT *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            run_cg<DT, EllOp<DT, IT, LT>, VAR>(A, n, b, x, P, max_iters, eps, used, defect, st, conv);
        });
    }
}  // namespace
}  // namespace hsolve

// =========================================================================
// Mixed-precision MG-IR
// =========================================================================
namespace hsolve
{
namespace
{
    struct MgScalars
    {
        int outer_active;       // 1 while the outer loop runs
        int run;                // 1 while the current inner solve runs
        int pad0, pad1;
        uint64_t outer_iters, inner_total, inner_last, cg_iters;
        double initial_defect, def_norm, scale, inv;
        float inner_initial, inner_defect;
    };

    // boundary injection of b into the initial guess (multigrid.hh:2282-2297)
    __global__ void inject_boundary_kernel(const double * __restrict__ b, double * __restrict__ ig, int64_t n,
                                           int64_t sq)
    {
        const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        // first row, last row, first column
        if (t < sq)
        {
            ig[t] = b[t];
            ig[n - sq + t] = b[n - sq + t];
            const int64_t k = t * sq;
            if (k < n)
                ig[k] = b[k];
        }
    }

    // od = b - A64 x (FP64 Q1 defect) and ||od||; FINISH 0: initial, 1: outer step
    template <int FINISH>
    __global__ __launch_bounds__(256) void outer_defect_kernel(k::Bands B, const double * __restrict__ b,
            const double * __restrict__ x, double * __restrict__ od, int64_t n, int64_t m, double * partials,
            unsigned * ticket, MgScalars * s, double outer_eps)
    {
        if (s->outer_active == 0)
            return;
        double acc = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const double prod = k::q1_row<double>(B, x, i, n, m);
            const double di = b[i] - prod;
            od[i] = di;
            acc += di * di;
        }
        finish2(acc, 0.0, partials, ticket, [&](double dd, double) {
            const double nrm = sqrt(dd);
            if (FINISH == 0)
            {
                s->initial_defect = nrm;
                s->def_norm = nrm;
                s->scale = 1.0;
                s->inv = 1.0;
                s->outer_iters = 0;
                s->inner_total = 0;
            }
            else
            {
                s->def_norm = nrm;
                s->outer_iters += 1;
                s->inner_total += s->inner_last;
                if (nrm < outer_eps * s->initial_defect)
                    s->outer_active = 0;
                else
                {
                    s->scale = nrm;
                    s->inv = 1.0 / nrm;
                }
            }
        });
    }

    // rhs32 = (float)(od * inv); opens the inner solve (run = 1, count = 0)
    __global__ __launch_bounds__(256) void outer_to_inner_kernel(const double * __restrict__ od,
            float * __restrict__ rhs32, int64_t n, MgScalars * s, int first)
    {
        if (s->outer_active == 0)
            return;
        const double inv = s->inv;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            rhs32[i] = first ? (float)od[i] : (float)(od[i] * inv);
    }

    __global__ void open_inner_kernel(MgScalars * s)
    {
        if (threadIdx.x == 0 && blockIdx.x == 0)
            s->run = s->outer_active;
    }

    // result += scale * (double)x32
    __global__ __launch_bounds__(256) void outer_accumulate_kernel(double * __restrict__ result,
            const float * __restrict__ x32, int64_t n, const MgScalars * s)
    {
        if (s->outer_active == 0)
            return;
        const double scale = s->scale;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            result[i] = result[i] + scale * (double)x32[i];
    }

    // ||d||_2 of the finest FP32 defect; MODE 0: initial norm of the inner
    // solve, MODE 1: end of a V-cycle (count it, stop on convergence)
    template <int MODE>
    __global__ __launch_bounds__(256) void inner_norm_kernel(const float * __restrict__ d, int64_t n,
            double * partials, unsigned * ticket, MgScalars * s, float tol, int convergence_check)
    {
        if (s->run == 0)
            return;
        double acc = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const double v = (double)d[i];
            acc += v * v;
        }
        finish2(acc, 0.0, partials, ticket, [&](double ss, double) {
            const float nrm = sqrtf((float)ss);
            if (MODE == 0)
            {
                s->inner_initial = nrm;
                s->inner_defect = nrm;
                if (convergence_check && nrm <= tol)
                    s->run = 0;   // nothing to do; the count keeps the previous solve's value
                else
                    s->inner_last = 0;
            }
            else
            {
                s->inner_defect = nrm;
                s->inner_last += 1;
                if (nrm <= s->inner_initial * tol)
                    s->run = 0;
            }
        });
    }

    // device copy guarded by a run flag
    template <typename DT>
    __global__ __launch_bounds__(256) void copy_if_kernel(const DT * __restrict__ src, DT * __restrict__ dst,
                                                          int64_t n, const int * active)
    {
        if (*active == 0)
            return;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            dst[i] = src[i];
    }

    // dst += src, guarded by a run flag (Sum<Tag>(x, c))
    template <typename DT>
    __global__ __launch_bounds__(256) void add_if_kernel(const DT * __restrict__ src, DT * __restrict__ dst,
                                                         int64_t n, const int * active)
    {
        if (*active == 0)
            return;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            dst[i] = dst[i] + src[i];
    }

    // counts a V-cycle that ran (one-level hierarchies skip the norm)
    __global__ void count_cycle_kernel(MgScalars * s)
    {
        if (threadIdx.x == 0 && blockIdx.x == 0 && s->run)
            s->inner_last += 1;
    }

    // adapt_correction_factor == 0: alpha = (d.c) / ((A c).c)
    __global__ __launch_bounds__(256) void adapt_alpha_kernel(k::Bands B, const float * __restrict__ d,
            const float * __restrict__ c, int64_t n, int64_t m, double * partials, unsigned * ticket,
            float * alpha_out, const int * active)
    {
        if (*active == 0)
            return;
        double a1 = 0.0, a2 = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const float t = k::q1_row<float>(B, c, i, n, m);
            a1 += (double)d[i] * (double)c[i];
            a2 += (double)t * (double)c[i];
        }
        finish2(a1, a2, partials, ticket, [&](double d1, double d2) { *alpha_out = (float)d1 / (float)d2; });
    }

    template <typename DT>
    __global__ __launch_bounds__(256) void axpy_dev_alpha_kernel(DT * __restrict__ x, const DT * __restrict__ c,
                                                                 const DT * alpha, int64_t n, const int * active)
    {
        if (*active == 0)
            return;
        const DT a = *alpha;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            x[i] = x[i] + a * c[i];
    }

    // FillMatrix<POISSON, DIRICHLET_0> on the device (fill_matrix.hh:37-76)
    template <typename DT>
    __global__ void fill_poisson_kernel(DT * ll, DT * ld, DT * lu, DT * dl, DT * dd, DT * du, DT * ul, DT * ud,
                                        DT * uu, int64_t N, int64_t M)
    {
        const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= N)
            return;
        const int64_t i = k + 1;
        const bool bnd = (i <= M || i > N - M || (i % M) == 1 || (i % M) == 0);
        const DT off = bnd ? DT(0.0) : DT(-1.0 / 3.0);
        ll[k] = off; ld[k] = off; lu[k] = off; dl[k] = off; du[k] = off; ul[k] = off; ud[k] = off; uu[k] = off;
        dd[k] = bnd ? DT(1.0) : DT(8.0 / 3.0);
    }

    // dinv = (float)((double)(1/dd) * 0.7): ElementInverse then Scale by a
    // double factor, as the reference tests build MGInfo::diags_inverted
    __global__ void dinv_kernel(const float * __restrict__ dd, float * __restrict__ dinv, int64_t n)
    {
        const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n)
            return;
        float v = dd[k];
        if (v != 0.0f)
            v = 1.0f / v;
        dinv[k] = (float)((double)v * 0.7);
    }

    struct Level
    {
        uint64_t n = 0;
        int64_t side = 0;
        float * bands = nullptr;    // 9 * n
        float * dinv = nullptr;
        float * x = nullptr;
        float * tmp = nullptr;
        float * c = nullptr;
        float * d = nullptr;
        float * rhs = nullptr;
        k::Bands B() const
        {
            k::Bands r;
            for (int i = 0; i < 9; ++i)
                r.b[i] = bands + (uint64_t)i * n;
            return r;
        }
    };
}  // namespace
}  // namespace hsolve

struct hsolve_mg_s
{
    hsolve_mg_params p;
    std::vector<hsolve::Level> lv;
    uint64_t n = 0;
    int64_t side = 0;
    double * bands64 = nullptr;   // 9 * n
    double * od = nullptr;        // outer defect
    double * result = nullptr;
    double * ig = nullptr;
    double * partials = nullptr;
    unsigned * ticket = nullptr;
    hsolve::MgScalars * s = nullptr;
    float * alpha_dev = nullptr;
    float * cg_work = nullptr;
    std::vector<void *> allocations;
    uint64_t bytes = 0;

    void * alloc(size_t b)
    {
        void * ptr = nullptr;
        hipError_t e = hipMalloc(&ptr, b ? b : 1);
        if (e != hipSuccess)
            throw hsolve::Error(HSOLVE_ERR_ALLOC, "MG hierarchy allocation of " + std::to_string(b) + " bytes failed");
        allocations.push_back(ptr);
        bytes += b;
        return ptr;
    }

    ~hsolve_mg_s()
    {
        for (void * a : allocations)
            (void)hipFree(a);
    }
};

namespace hsolve
{
namespace
{
    unsigned grid_cap(uint64_t n, unsigned cap) { return std::min<unsigned>(grid_for(n, kBlock), cap); }

    void validate_params(const hsolve_mg_params * p)
    {
        require(p != nullptr, HSOLVE_ERR_INVALID, "mg: null params");
        require(p->min_level >= 1 && p->min_level <= p->max_level && p->max_level <= 14, HSOLVE_ERR_INVALID,
                "mg: need 1 <= min_level <= max_level <= 14");
        require(p->n_pre_smooth >= 1, HSOLVE_ERR_INVALID, "mg: n_pre_smooth must be >= 1");
        require(p->n_max_iter >= 1, HSOLVE_ERR_INVALID, "mg: n_max_iter must be >= 1");
        const uint64_t nmin = ((1ull << p->min_level) + 1) * ((1ull << p->min_level) + 1);
        require(nmin <= 16641, HSOLVE_ERR_INVALID, "mg: coarsest level above 129^2");
    }

    hsolve_mg_s * mg_alloc(const hsolve_mg_params * p)
    {
        validate_params(p);
        std::unique_ptr<hsolve_mg_s> mg(new hsolve_mg_s());
        mg->p = *p;
        mg->side = (int64_t)(1ull << p->max_level) + 1;
        mg->n = (uint64_t)(mg->side * mg->side);
        mg->lv.resize(p->max_level + 1);
        for (uint64_t L = p->min_level; L <= p->max_level; ++L)
        {
            Level & l = mg->lv[L];
            l.side = (int64_t)(1ull << L) + 1;
            l.n = (uint64_t)(l.side * l.side);
            l.bands = (float *)mg->alloc(sizeof(float) * 9 * l.n);
            l.dinv = (float *)mg->alloc(sizeof(float) * l.n);
            l.x = (float *)mg->alloc(sizeof(float) * l.n);
            l.tmp = (float *)mg->alloc(sizeof(float) * l.n);
            l.c = (float *)mg->alloc(sizeof(float) * l.n);
            l.d = (float *)mg->alloc(sizeof(float) * l.n);
            l.rhs = (float *)mg->alloc(sizeof(float) * l.n);
        }
        const uint64_t n = mg->n;
        mg->bands64 = (double *)mg->alloc(sizeof(double) * 9 * n);
        mg->od = (double *)mg->alloc(sizeof(double) * n);
        mg->result = (double *)mg->alloc(sizeof(double) * n);
        mg->ig = (double *)mg->alloc(sizeof(double) * n);
        mg->partials = (double *)mg->alloc(sizeof(double) * 2 * kSolverBlocks);
        mg->ticket = (unsigned *)mg->alloc(256);
        mg->s = (MgScalars *)mg->alloc(sizeof(MgScalars));
        mg->alpha_dev = (float *)mg->alloc(256);
        const uint64_t nmin = mg->lv[p->min_level].n;
        mg->cg_work = (float *)mg->alloc(sizeof(float) * 4 * nmin + 256);
        HIP_CHECK(hipMemset(mg->ticket, 0, 256));
        for (uint64_t L = p->min_level; L <= p->max_level; ++L)
            HIP_CHECK(hipMemset(mg->lv[L].x, 0, sizeof(float) * mg->lv[L].n));
        return mg.release();
    }

    // One V-cycle (multigrid.hh:1660-1905), all kernels gated on s->run.
    void enqueue_vcycle(hsolve_mg_s * mg, bool first_cycle, hipStream_t st)
    {
        const hsolve_mg_params & p = mg->p;
        const uint64_t minl = p.min_level, maxl = p.max_level;
        const int * run = &mg->s->run;
        k::Mask mk;
        for (int i = 0; i < 8; ++i)
            mk.dir[i] = (p.macro_border_mask[i] == 2) ? 1 : 0;
        const uint64_t npre = p.n_pre_smooth;

        auto sweeps = [&](Level & l, float *& xcur, float *& xalt, const float * rhs, uint64_t count) {
            for (uint64_t sIt = 0; sIt < count; ++sIt)
            {
                hipLaunchKernelGGL((k::q1_kernel<float, 2>), dim3(grid_for(l.n, kBlock)), dim3(kBlock), 0, st, l.B(),
                                   (const float *)xcur, xalt, rhs, (const float *)l.dinv, (int64_t)l.n, (int64_t)l.side,
                                   run);
                std::swap(xcur, xalt);
            }
        };

        if (minl != maxl)
        {
            // ---- restriction loop
            for (uint64_t L = maxl; L > minl; --L)
            {
                Level & l = mg->lv[L];
                const unsigned g = grid_for(l.n, kBlock);
                // defect of this level before smoothing: finest level keeps its
                // defect in d (rhs in the first cycle with initial_zero);
                // coarser levels use the restricted defect stored in rhs
                const float * dsrc = (L == maxl) ? ((first_cycle && p.initial_zero) ? l.rhs : l.d) : l.rhs;
                if (L == maxl)
                {
                    // c = dinv*d, (npre-1) sweeps on c with rhs d, x += c
                    if (npre == 1)
                        hipLaunchKernelGGL((k::jacobi_zero1<float, 1>), dim3(g), dim3(kBlock), 0, st, dsrc,
                                           (const float *)l.dinv, l.x, (int64_t)l.n, run);
                    else if (npre == 2)
                        hipLaunchKernelGGL((k::jacobi_zero2_q1<float, 1>), dim3(g), dim3(kBlock), 0, st, l.B(), dsrc,
                                           (const float *)l.dinv, l.x, (int64_t)l.n, (int64_t)l.side, run);
                    else
                    {
                        hipLaunchKernelGGL((k::jacobi_zero2_q1<float, 0>), dim3(g), dim3(kBlock), 0, st, l.B(), dsrc,
                                           (const float *)l.dinv, l.c, (int64_t)l.n, (int64_t)l.side, run);
                        float * cc = l.c;
                        float * ca = l.tmp;
                        sweeps(l, cc, ca, dsrc, npre - 2);
                        hipLaunchKernelGGL((add_if_kernel<float>), dim3(grid_cap(l.n, 16384)), dim3(kBlock), 0, st,
                                           (const float *)cc, l.x, (int64_t)l.n, run);
                    }
                }
                else
                {
                    // x = dinv*d, (npre-1) sweeps on x with rhs d
                    if (npre == 1)
                        hipLaunchKernelGGL((k::jacobi_zero1<float, 0>), dim3(g), dim3(kBlock), 0, st, dsrc,
                                           (const float *)l.dinv, l.x, (int64_t)l.n, run);
                    else
                    {
                        hipLaunchKernelGGL((k::jacobi_zero2_q1<float, 0>), dim3(g), dim3(kBlock), 0, st, l.B(), dsrc,
                                           (const float *)l.dinv, l.x, (int64_t)l.n, (int64_t)l.side, run);
                        float * xc = l.x;
                        float * xa = l.tmp;
                        sweeps(l, xc, xa, dsrc, npre - 2);
                        if (xc != l.x)
                            hipLaunchKernelGGL((copy_if_kernel<float>), dim3(grid_cap(l.n, 16384)), dim3(kBlock), 0, st,
                                               (const float *)xc, l.x, (int64_t)l.n, run);
                    }
                }
                // d = rhs - A x
                hipLaunchKernelGGL((k::q1_kernel<float, 1>), dim3(g), dim3(kBlock), 0, st, l.B(), (const float *)l.x,
                                   l.d, (const float *)l.rhs, (const float *)nullptr, (int64_t)l.n, (int64_t)l.side, run);
                // restrict into the coarser level's rhs (= its defect)
                Level & lc = mg->lv[L - 1];
                hipLaunchKernelGGL((k::restriction_kernel<float>), dim3(grid_for(lc.n, kBlock)), dim3(kBlock), 0, st,
                                   lc.rhs, (const float *)l.d, (int64_t)lc.side, mk, run);
            }
        }
        // ---- coarse grid solve: ConjugateGradients<NONE>(A, d, x, FLT_EPSILON)
        {
            Level & l = mg->lv[minl];
            // with one level the coarse solve sees the finest defect d
            const float * crhs = (minl == maxl) ? ((first_cycle && p.initial_zero) ? l.rhs : l.d) : l.rhs;
            hipLaunchKernelGGL((k::cg_konv_q1_kernel<float>), dim3(1), dim3(kBlock), 0, st, l.B(),
                               crhs, l.x, mg->cg_work, (int64_t)l.n, (int64_t)l.side,
                               (double)1.19209290e-07F, (uint64_t)(100 * l.n + 1000), &mg->s->cg_iters, run);
            if (minl == maxl)
            {
                hipLaunchKernelGGL((k::q1_kernel<float, 1>), dim3(grid_for(l.n, kBlock)), dim3(kBlock), 0, st, l.B(),
                                   (const float *)l.x, l.d, (const float *)l.rhs, (const float *)nullptr,
                                   (int64_t)l.n, (int64_t)l.side, run);
                return;
            }
        }
        // ---- prolongation loop
        const float adapt = (float)p.adapt_correction_factor;
        for (uint64_t L = minl + 1; L <= maxl; ++L)
        {
            Level & l = mg->lv[L];
            Level & lc = mg->lv[L - 1];
            const unsigned g = grid_for(l.n, kBlock);
            if (adapt != 0.0f)
            {
                hipLaunchKernelGGL((k::prolongation_kernel<float, true>), dim3(g), dim3(kBlock), 0, st, l.x,
                                   (const float *)lc.x, (int64_t)l.side, adapt, mk, run);
            }
            else
            {
                hipLaunchKernelGGL((k::prolongation_kernel<float, false>), dim3(g), dim3(kBlock), 0, st, l.c,
                                   (const float *)lc.x, (int64_t)l.side, 0.0f, mk, run);
                hipLaunchKernelGGL(adapt_alpha_kernel, dim3(grid_cap(l.n, kSolverBlocks)), dim3(kBlock), 0, st, l.B(),
                                   (const float *)l.d, (const float *)l.c, (int64_t)l.n, (int64_t)l.side, mg->partials,
                                   mg->ticket, mg->alpha_dev, run);
                hipLaunchKernelGGL((axpy_dev_alpha_kernel<float>), dim3(grid_cap(l.n, 16384)), dim3(kBlock), 0, st, l.x,
                                   (const float *)l.c, (const float *)mg->alpha_dev, (int64_t)l.n, run);
            }
            // post-smoothing: n_pre_smooth sweeps (multigrid.hh:1851)
            float * xc = l.x;
            float * xa = l.tmp;
            sweeps(l, xc, xa, l.rhs, npre);
            if (xc != l.x)
                hipLaunchKernelGGL((copy_if_kernel<float>), dim3(grid_cap(l.n, 16384)), dim3(kBlock), 0, st,
                                   (const float *)xc, l.x, (int64_t)l.n, run);
            if (L == maxl)
                hipLaunchKernelGGL((k::q1_kernel<float, 1>), dim3(g), dim3(kBlock), 0, st, l.B(), (const float *)l.x,
                                   l.d, (const float *)l.rhs, (const float *)nullptr, (int64_t)l.n, (int64_t)l.side,
                                   run);
        }
    }

    void enqueue_inner(hsolve_mg_s * mg, hipStream_t st)
    {
        const hsolve_mg_params & p = mg->p;
        Level & top = mg->lv[p.max_level];
        const float tol = (float)p.tolerance;
        hipLaunchKernelGGL(open_inner_kernel, dim3(1), dim3(64), 0, st, mg->s);
        // initial defect of the inner solve
        const float * d0 = top.rhs;
        if (! p.initial_zero)
        {
            hipLaunchKernelGGL((k::q1_kernel<float, 1>), dim3(grid_for(top.n, kBlock)), dim3(kBlock), 0, st, top.B(),
                               (const float *)top.x, top.d, (const float *)top.rhs, (const float *)nullptr,
                               (int64_t)top.n, (int64_t)top.side, (const int *)&mg->s->run);
            d0 = top.d;
        }
        hipLaunchKernelGGL((inner_norm_kernel<0>), dim3(grid_cap(top.n, kSolverBlocks)), dim3(kBlock), 0, st, d0,
                           (int64_t)top.n, mg->partials, mg->ticket, mg->s, tol, (int)p.convergence_check);
        for (uint64_t it = 0; it < p.n_max_iter; ++it)
        {
            enqueue_vcycle(mg, it == 0, st);
            if (p.min_level == p.max_level)
            {
                // one-level case: the cycle loop is left right after the coarse solve
                hipLaunchKernelGGL(count_cycle_kernel, dim3(1), dim3(64), 0, st, mg->s);
                break;
            }
            hipLaunchKernelGGL((inner_norm_kernel<1>), dim3(grid_cap(top.n, kSolverBlocks)), dim3(kBlock), 0, st,
                               (const float *)top.d, (int64_t)top.n, mg->partials, mg->ticket, mg->s, tol, 0);
        }
    }
}  // namespace
}  // namespace hsolve

using namespace hsolve;

extern "C" {

int hsolve_cg_q1(int dtype, const void * const * bands, uint64_t size, const void * b, void * x, uint64_t max_iters,
                 double eps, uint64_t * used_iters, double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(b && x, HSOLVE_ERR_INVALID, "cg_q1: null vector");
    const int64_t m = (int64_t)std::sqrt((double)size);
    require((uint64_t)(m * m) == size && size > 0, HSOLVE_ERR_SIZE, "cg_q1: size is not a square");
    hipStream_t st = as_stream(stream);
    int conv = 0;
    if (dtype)
    {
        Q1Op<double> A{bands9(bands), (int64_t)size, m};
        run_cg<double, Q1Op<double>, false>(A, size, (const double *)b, (double *)x, nullptr, max_iters, eps,
                                            used_iters, defect_out, st, &conv);
    }
    else
    {
        Q1Op<float> A{bands9(bands), (int64_t)size, m};
        run_cg<float, Q1Op<float>, false>(A, size, (const float *)b, (float *)x, nullptr, max_iters, eps, used_iters,
                                          defect_out, st, &conv);
    }
    if (! conv && max_iters > 0)
    {
        set_last_error("cg_q1: max_iters reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

int hsolve_cg_ell(hsolve_ell m, const void * b, void * x, uint64_t max_iters, double eps, uint64_t * used_iters,
                  double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && b && x, HSOLVE_ERR_INVALID, "cg_ell: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "cg_ell: matrix is not square");
    int conv = 0;
    if (m->dtype)
        run_cg_ell<double, false>(m, (const double *)b, (double *)x, nullptr, max_iters, eps, used_iters, defect_out,
                                  as_stream(stream), &conv);
    else
        run_cg_ell<float, false>(m, (const float *)b, (float *)x, nullptr, max_iters, eps, used_iters, defect_out,
                                 as_stream(stream), &conv);
    if (! conv && max_iters > 0)
    {
        set_last_error("cg_ell: max_iters reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

int hsolve_pcg_ell(hsolve_ell m, const void * P, const void * b, void * x, uint64_t max_iters, double eps,
                   uint64_t * used_iters, double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && P && b && x, HSOLVE_ERR_INVALID, "pcg_ell: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "pcg_ell: matrix is not square");
    int conv = 0;
    if (m->dtype)
        run_cg_ell<double, true>(m, (const double *)b, (double *)x, (const double *)P, max_iters, eps, used_iters,
                                 defect_out, as_stream(stream), &conv);
    else
        run_cg_ell<float, true>(m, (const float *)b, (float *)x, (const float *)P, max_iters, eps, used_iters,
                                defect_out, as_stream(stream), &conv);
    if (! conv && max_iters > 0)
    {
        set_last_error("pcg_ell: max_iters reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

int hsolve_mg_create(const hsolve_mg_params * params, const double * const * bands64_host,
                     const float * const * bands32_host, const float * const * dinv32_host, hsolve_mg * out)
{
    HSOLVE_API_BEGIN
    require(out && bands64_host && bands32_host, HSOLVE_ERR_INVALID, "mg_create: null pointer");
    *out = nullptr;
    std::unique_ptr<hsolve_mg_s> mg(mg_alloc(params));
    for (int b = 0; b < 9; ++b)
    {
        require(bands64_host[b] != nullptr, HSOLVE_ERR_INVALID, "mg_create: null FP64 band");
        HIP_CHECK(hipMemcpy(mg->bands64 + (uint64_t)b * mg->n, bands64_host[b], sizeof(double) * mg->n,
                            hipMemcpyHostToDevice));
    }
    for (uint64_t L = params->min_level; L <= params->max_level; ++L)
    {
        Level & l = mg->lv[L];
        for (int b = 0; b < 9; ++b)
        {
            const float * src = bands32_host[L * 9 + b];
            require(src != nullptr, HSOLVE_ERR_INVALID, "mg_create: null FP32 band");
            HIP_CHECK(hipMemcpy(l.bands + (uint64_t)b * l.n, src, sizeof(float) * l.n, hipMemcpyHostToDevice));
        }
        if (dinv32_host && dinv32_host[L])
            HIP_CHECK(hipMemcpy(l.dinv, dinv32_host[L], sizeof(float) * l.n, hipMemcpyHostToDevice));
        else
            hipLaunchKernelGGL(dinv_kernel, dim3(grid_for(l.n, kBlock)), dim3(kBlock), 0, 0,
                               (const float *)(l.bands + 4 * l.n), l.dinv, (int64_t)l.n);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    *out = mg.release();
    HSOLVE_API_END
}

int hsolve_mg_create_poisson(const hsolve_mg_params * params, hsolve_mg * out)
{
    HSOLVE_API_BEGIN
    require(out != nullptr, HSOLVE_ERR_INVALID, "mg_create_poisson: null out");
    *out = nullptr;
    std::unique_ptr<hsolve_mg_s> mg(mg_alloc(params));
    {
        double * b = mg->bands64;
        const uint64_t n = mg->n;
        hipLaunchKernelGGL((fill_poisson_kernel<double>), dim3(grid_for(n, kBlock)), dim3(kBlock), 0, 0, b, b + n,
                           b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n, b + 7 * n, b + 8 * n, (int64_t)n,
                           mg->side);
    }
    for (uint64_t L = params->min_level; L <= params->max_level; ++L)
    {
        Level & l = mg->lv[L];
        float * b = l.bands;
        const uint64_t n = l.n;
        hipLaunchKernelGGL((fill_poisson_kernel<float>), dim3(grid_for(n, kBlock)), dim3(kBlock), 0, 0, b, b + n,
                           b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n, b + 7 * n, b + 8 * n, (int64_t)n,
                           l.side);
        hipLaunchKernelGGL(dinv_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, 0, (const float *)(b + 4 * n),
                           l.dinv, (int64_t)n);
    }
    HIP_CHECK(hipGetLastError());
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
    *bytes = mg->bytes;
    HSOLVE_API_END
}

int hsolve_mg_solve(hsolve_mg mg, const void * b_dev, void * x_dev, uint64_t * outer_iters, uint64_t * inner_iters,
                    double * final_defect, double * initial_defect, void * stream)
{
    HSOLVE_API_BEGIN
    require(mg && b_dev && x_dev, HSOLVE_ERR_INVALID, "mg_solve: null pointer");
    hipStream_t st = as_stream(stream);
    const hsolve_mg_params & p = mg->p;
    const uint64_t n = mg->n;
    const double * b = (const double *)b_dev;
    double * x = (double *)x_dev;
    Level & top = mg->lv[p.max_level];
    k::Bands B64;
    for (int i = 0; i < 9; ++i)
        B64.b[i] = mg->bands64 + (uint64_t)i * n;
    const double outer_eps = (double)(float)p.tolerance;

    // initial state: outer loop active, iterates of coarser levels zero
    MgScalars h;
    std::memset(&h, 0, sizeof(h));
    h.outer_active = 1;
    HIP_CHECK(hipMemcpyAsync(mg->s, &h, sizeof(h), hipMemcpyHostToDevice, st));
    for (uint64_t L = p.min_level; L < p.max_level; ++L)
        hipLaunchKernelGGL((k::elem_kernel<float, k::E_FILL>), dim3(grid_cap(mg->lv[L].n, 16384)), dim3(kBlock), 0, st,
                           mg->lv[L].x, (const float *)nullptr, (const float *)nullptr, (const float *)nullptr, 0.0f,
                           mg->lv[L].n);
    // initial guess with the Dirichlet values of b injected (multigrid.hh:2277-2298)
    HIP_CHECK(hipMemcpyAsync(mg->ig, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(inject_boundary_kernel, dim3(grid_for(mg->side, kBlock)), dim3(kBlock), 0, st, b, mg->ig,
                       (int64_t)n, mg->side);
    hipLaunchKernelGGL((k::convert_kernel<double, float>), dim3(grid_cap(n, 16384)), dim3(kBlock), 0, st,
                       (const double *)mg->ig, top.x, n);
    hipLaunchKernelGGL((k::elem_kernel<double, k::E_FILL>), dim3(grid_cap(n, 16384)), dim3(kBlock), 0, st, mg->result,
                       (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, 0.0, n);
    hipLaunchKernelGGL((outer_defect_kernel<0>), dim3(grid_cap(n, kSolverBlocks)), dim3(kBlock), 0, st, B64, b,
                       (const double *)mg->ig, mg->od, (int64_t)n, mg->side, mg->partials, mg->ticket, mg->s,
                       outer_eps);
    HIP_CHECK(hipGetLastError());

    const uint64_t cap = p.max_outer ? p.max_outer : (uint64_t)-1;
    uint64_t enqueued = 0;
    uint64_t chunk = 1;
    while (true)
    {
        for (uint64_t c = 0; c < chunk && enqueued < cap; ++c, ++enqueued)
        {
            hipLaunchKernelGGL(outer_to_inner_kernel, dim3(grid_cap(n, 16384)), dim3(kBlock), 0, st,
                               (const double *)mg->od, top.rhs, (int64_t)n, mg->s, enqueued == 0 ? 1 : 0);
            enqueue_inner(mg, st);
            hipLaunchKernelGGL(outer_accumulate_kernel, dim3(grid_cap(n, 16384)), dim3(kBlock), 0, st, mg->result,
                               (const float *)top.x, (int64_t)n, (const MgScalars *)mg->s);
            hipLaunchKernelGGL((outer_defect_kernel<1>), dim3(grid_cap(n, kSolverBlocks)), dim3(kBlock), 0, st, B64, b,
                               (const double *)mg->result, mg->od, (int64_t)n, mg->side, mg->partials, mg->ticket,
                               mg->s, outer_eps);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(&h, mg->s, sizeof(h), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (h.outer_active == 0 || enqueued >= cap)
            break;
        if (chunk < 4)
            chunk *= 2;
    }
    HIP_CHECK(hipMemcpyAsync(x, mg->result, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (outer_iters) *outer_iters = h.outer_iters;
    if (inner_iters) *inner_iters = h.inner_total;
    if (final_defect) *final_defect = h.def_norm;
    if (initial_defect) *initial_defect = h.initial_defect;
    if (h.outer_active)
    {
        set_last_error("mg_solve: max_outer reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

}  // extern "C"
