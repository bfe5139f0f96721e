// solvers.hip -- device-resident Krylov solvers of the MI355X backend:
// CGSolver<Tag, methods::NONE> / <Tag, methods::VAR> (honei/math/cg.hh) on
// Q1 bands or an ELL handle, FP32 or FP64.  (The multigrid solver lives in
// mg_engine.hip / mg_solve.hip.)
//
// Design: every scalar of the iteration (alpha, lambda, norms, iteration
// counts, convergence flags) lives on the device.  Each reduction is fused
// into the kernel that produces its operand (block partials) and finished
// by a one-block kernel, which also performs the scalar update and the
// stopping test.
// Every kernel of a step first reads a `run` flag and returns if the solve
// has already stopped, so the host enqueues whole chunks of iterations and
// synchronises once per chunk, while the iteration count stays exactly the
// reference's.
#include "ell_handle.hh"
#include "kernels_la.hh"

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
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
        template <typename XV> __device__ __forceinline__ DT row(const XV & x, int64_t i) const
        {
            return k::q1_row<DT>(B, x, i, n, m);
        }
    };

    template <typename DT, typename IT, typename LT> using EllOp = EllRowOp<DT, IT, LT>;


    // --------------------------------------------------------------- CG state
    template <typename DT> struct CgScalars
    {
        DT alpha, alpha_old, alpha_new, temp, lambda, talpha;
        DT initial_defect, current_defect;
        DT eps_relative;
        uint64_t iters, max_iters;
        int run;          // 1 while iterating
        int converged;    // 1 when a stopping test fired
        // 1: ConjugateGradients<NONE>'s fixed-iteration ELL form
        // (conjugate_gradients.hh:203-248): no epsilon guards on v.p and
        // alpha_old, thresholds 1e-8 * ||r0|| and 1e-8 formed in double
        int plain;
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
    // (block partials; cg_init_finish forms the totals)
    template <typename DT, typename Op, bool VAR>
    __global__ __launch_bounds__(256) void cg_init_kernel(Op A, const DT * __restrict__ b, const DT * __restrict__ x,
            DT * __restrict__ r, DT * __restrict__ p, const DT * __restrict__ P, int64_t n, double * partials)
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
        k::put_partials2(acc_rr, acc_rp, partials, blockIdx.x);
    }

    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void cg_init_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        double rr, rp;
        k::sum_partials2(partials, nb, rr, rp);
        if (threadIdx.x == 0)
        {
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
        }
    }

    // v = A p; partials of v.p
    template <typename DT, typename Op>
    __global__ __launch_bounds__(256) void cg_spmv_kernel(Op A, const DT * __restrict__ p, DT * __restrict__ v,
            int64_t n, double * partials, const CgScalars<DT> * s)
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
        k::put_partials2(acc, 0.0, partials, blockIdx.x);
    }

    // temp = v.p; lambda
    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void cg_spmv_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double vp, unused;
        k::sum_partials2(partials, nb, vp, unused);
        if (threadIdx.x == 0)
        {
            const DT temp = (DT)vp;
            s->temp = temp;
            if (VAR)
                s->lambda = s->alpha_new / guard_signed<DT>(temp);
            else
                s->lambda = s->alpha / (s->plain ? temp : guard<DT>(temp));
        }
    }

    // x += lambda p; r += (-lambda) v; partials of r.r (and r.(P r) for VAR)
    // FUSED (small grids): every block first forms temp = v.p from the SpMV's
    // partials itself (the same sum in the same order in every block) and
    // lambda from it, so cg_spmv_finish is not launched; block 0 stores
    // temp and lambda.  The SpMV's partials are read from sp_partials, this
    // kernel's own go to `partials`, a different buffer.
    template <typename DT, bool VAR, bool FUSED = false>
    __global__ __launch_bounds__(256) void cg_update_kernel(DT * __restrict__ x, DT * __restrict__ r,
            const DT * __restrict__ p, const DT * __restrict__ v, const DT * __restrict__ P, int64_t n,
            double * partials, CgScalars<DT> * s, const double * sp_partials = nullptr, unsigned sp_nb = 0)
    {
        if (s->run == 0)
            return;
        DT lambda;
        if (FUSED)
        {
            __shared__ DT sh_lambda;
            double vp, unused;
            k::sum_partials2(sp_partials, sp_nb, vp, unused);
            if (threadIdx.x == 0)
            {
                const DT temp = (DT)vp;
                const DT lam = VAR ? s->alpha_new / guard_signed<DT>(temp)
                                   : s->alpha / (s->plain ? temp : guard<DT>(temp));
                sh_lambda = lam;
                if (blockIdx.x == 0)
                {
                    s->temp = temp;
                    s->lambda = lam;
                }
            }
            __syncthreads();
            lambda = sh_lambda;
        }
        else
            lambda = s->lambda;
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
        k::put_partials2(acc_rr, acc_rz, partials, blockIdx.x);
    }

    // the stopping test after an update and the next direction's factor,
    // on a copy of the scalars (rr = r.r, rz = r.(P r) for VAR)
    template <typename DT, bool VAR>
    __device__ __forceinline__ void cg_test(CgScalars<DT> & s, double rr, double rz)
    {
        const DT eps = s.eps_relative;
        if (VAR)
        {
            ++s.iters;
            const DT cur = (DT)sqrt((DT)rr);
            s.current_defect = cur;
            if (cur < eps * s.initial_defect || cur < eps)
            {
                s.converged = 1;
                s.run = 0;
                return;
            }
            if (s.iters == s.max_iters)
            {
                s.run = 0;
                return;
            }
            s.alpha = s.alpha_new;
            s.alpha_new = (DT)rz;
            s.talpha = s.alpha_new / guard_signed<DT>(s.alpha);
        }
        else
        {
            s.alpha_old = s.alpha;
            s.alpha = (DT)rr;
            s.talpha = s.alpha / (s.plain ? s.alpha_old : guard<DT>(s.alpha_old));
            ++s.iters;
            const DT cur = (DT)sqrt(s.alpha);
            s.current_defect = cur;
            const bool stop = s.plain ? ((double)cur < 1e-08 * (double)s.initial_defect || (double)cur < 1e-08)
                                      : (cur < eps * s.initial_defect || cur < eps);
            if (stop)
            {
                s.converged = 1;
                s.run = 0;
            }
            else if (s.iters == s.max_iters)
                s.run = 0;
        }
    }

    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void cg_update_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double rr, rz;
        k::sum_partials2(partials, nb, rr, rz);
        if (threadIdx.x != 0)
            return;
        CgScalars<DT> t = *s;
        cg_test<DT, VAR>(t, rr, rz);
        *s = t;
    }

    // The next direction p * talpha + r (NONE) or p * talpha + P*r (VAR),
    // formed where it is read: x[j] of an operator row.
    template <typename DT, bool VAR> struct NextDir
    {
        const DT * p;
        const DT * r;
        const DT * P;
        DT talpha;
        __device__ __forceinline__ DT operator[](int64_t j) const
        {
            const DT t = p[j] * talpha;
            return t + (VAR ? P[j] * r[j] : r[j]);
        }
    };

    // Small grids (launch-bound solves): the direction update folded into
    // the next SpMV.  p_new = p_old * talpha + (P)r is formed at every read,
    // with the operations of cg_direction_kernel, so every block sees the
    // same values; each row stores its own p_new into the other direction
    // buffer (p_old stays intact while other blocks still read it).  talpha
    // and the stop flag come from cg_update_finish.
    template <typename DT, typename Op, bool VAR>
    __global__ __launch_bounds__(256) void cg_spmv_dir_kernel(Op A, const DT * __restrict__ p_old,
            const DT * __restrict__ r, const DT * __restrict__ P, DT * __restrict__ p_new, DT * __restrict__ v,
            int64_t n, double * partials, const CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        const NextDir<DT, VAR> d{p_old, r, P, s->talpha};
        double acc = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const DT vi = A.row(d, i);
            const DT pi = d[i];
            p_new[i] = pi;
            v[i] = vi;
            acc += (double)vi * (double)pi;
        }
        k::put_partials2(acc, 0.0, partials, blockIdx.x);
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

    // ----------------------------------------- ConjugateGradients<Tag, JAC>
    // conjugate_gradients.hh:1110-1155 with the ELL cg_kernel :885-913.
    // Scalars: temp = beta (r.z), alpha = step, talpha = the direction factor.
    template <typename DT> __device__ __forceinline__ DT div_guard(DT num, DT den)
    {
        return fabs(den) >= Eps<DT>::v ? num / den : num / Eps<DT>::v;
    }

    // r = b - A x; z = dd * r; d = z; partials of r.r and r.z
    template <typename DT, typename Op>
    __global__ __launch_bounds__(256) void jac_init_kernel(Op A, const DT * __restrict__ b,
            const DT * __restrict__ x, const DT * __restrict__ dd, DT * __restrict__ r, DT * __restrict__ z,
            DT * __restrict__ d, int64_t n, double * partials)
    {
        double acc_rr = 0.0, acc_rz = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const DT ri = b[i] - A.row(x, i);
            r[i] = ri;
            const DT zi = dd[i] * ri;
            z[i] = zi;
            d[i] = zi;
            acc_rr += (double)ri * (double)ri;
            acc_rz += (double)ri * (double)zi;
        }
        k::put_partials2(acc_rr, acc_rz, partials, blockIdx.x);
    }

    template <typename DT>
    __global__ __launch_bounds__(256) void jac_init_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        double rr, rz;
        k::sum_partials2(partials, nb, rr, rz);
        if (threadIdx.x == 0)
        {
            s->iters = 0;
            s->converged = 0;
            s->run = (s->max_iters > 0) ? 1 : 0;
            s->initial_defect = (DT)sqrt((DT)rr);
            s->current_defect = s->initial_defect;
            s->temp = (DT)rz;
        }
    }

    // alpha = (r.z) / (d.t), t = A d computed by cg_spmv_kernel
    template <typename DT>
    __global__ __launch_bounds__(256) void jac_spmv_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double dt, unused;
        k::sum_partials2(partials, nb, dt, unused);
        if (threadIdx.x == 0)
            s->alpha = div_guard<DT>(s->temp, (DT)dt);
    }

    // x += alpha d; r -= alpha t; z = dd * r; partials of z.r and r.r
    // FUSED (small grids): every block forms d.t from the SpMV's partials
    // (sp_partials, the same sum in the same order in every block) and alpha
    // from it, so jac_spmv_finish is not launched; block 0 stores alpha.
    // Other blocks read only s->temp and s->run here.
    template <typename DT, bool FUSED = false>
    __global__ __launch_bounds__(256) void jac_update_kernel(DT * __restrict__ x, DT * __restrict__ r,
            DT * __restrict__ z, const DT * __restrict__ d, const DT * __restrict__ t, const DT * __restrict__ dd,
            int64_t n, double * partials, CgScalars<DT> * s, const double * sp_partials = nullptr,
            unsigned sp_nb = 0)
    {
        if (s->run == 0)
            return;
        DT alpha;
        if (FUSED)
        {
            __shared__ DT sh_alpha;
            double dt, unused;
            k::sum_partials2(sp_partials, sp_nb, dt, unused);
            if (threadIdx.x == 0)
            {
                const DT a = div_guard<DT>(s->temp, (DT)dt);
                sh_alpha = a;
                if (blockIdx.x == 0)
                    s->alpha = a;
            }
            __syncthreads();
            alpha = sh_alpha;
        }
        else
            alpha = s->alpha;
        const DT malpha = -alpha;
        double acc_zr = 0.0, acc_rr = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            x[i] = x[i] + alpha * d[i];
            const DT ri = r[i] + malpha * t[i];
            r[i] = ri;
            const DT zi = dd[i] * ri;
            z[i] = zi;
            acc_zr += (double)zi * (double)ri;
            acc_rr += (double)ri * (double)ri;
        }
        k::put_partials2(acc_zr, acc_rr, partials, blockIdx.x);
    }

    // beta = (z.r) / (r.z)_old; the stopping test ||r|| < eps ||r0|| or the cap
    template <typename DT>
    __global__ __launch_bounds__(256) void jac_update_finish(const double * partials, unsigned nb,
                                                             CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double zr, rr;
        k::sum_partials2(partials, nb, zr, rr);
        if (threadIdx.x != 0)
            return;
        s->talpha = div_guard<DT>((DT)zr, s->temp);
        s->temp = (DT)zr;
        ++s->iters;
        const DT cur = (DT)sqrt((DT)rr);
        s->current_defect = cur;
        if (cur < s->initial_defect * s->eps_relative)
        {
            s->converged = 1;
            s->run = 0;
        }
        else if (s->iters == s->max_iters)
            s->run = 0;
    }

    // d = z + beta d (ScaledSum(d, z, temp_0, beta) with temp_0 a copy of d)
    template <typename DT>
    __global__ __launch_bounds__(256) void jac_direction_kernel(DT * __restrict__ d, const DT * __restrict__ z,
            int64_t n, const CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        const DT beta = s->talpha;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
            d[i] = beta * d[i] + z[i];
    }

    // The next direction beta * d + z of jac_direction_kernel, formed where
    // it is read.
    template <typename DT> struct JacNextDir
    {
        const DT * d;
        const DT * z;
        DT beta;
        __device__ __forceinline__ DT operator[](int64_t j) const
        {
            const DT t = beta * d[j];
            return t + z[j];
        }
    };

    // Small grids: the direction update folded into the next SpMV, as
    // cg_spmv_dir_kernel (t = A d_new, partials of t.d_new; each row stores
    // its d_new into the other direction buffer).
    template <typename DT, typename Op>
    __global__ __launch_bounds__(256) void jac_spmv_dir_kernel(Op A, const DT * __restrict__ d_old,
            const DT * __restrict__ z, DT * __restrict__ d_new, DT * __restrict__ t, int64_t n, double * partials,
            const CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        const JacNextDir<DT> nd{d_old, z, s->talpha};
        double acc = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const DT ti = A.row(nd, i);
            const DT di = nd[i];
            d_new[i] = di;
            t[i] = ti;
            acc += (double)ti * (double)di;
        }
        k::put_partials2(acc, 0.0, partials, blockIdx.x);
    }

    // One device array, freed with the object.
    template <typename DT> struct DeviceArray
    {
        DT * p = nullptr;
        explicit DeviceArray(uint64_t n)
        {
            if (hipMalloc((void **)&p, sizeof(DT) * (n ? n : 1)) != hipSuccess)
                throw Error(HSOLVE_ERR_ALLOC, "solver workspace allocation failed");
        }
        ~DeviceArray()
        {
            if (p)
                (void)hipFree(p);
        }
    };

    // Device workspace of one CG solve: allocated here, or carved from a
    // caller's block of at least bytes_for(n) bytes (256-byte aligned).
    template <typename DT> struct CgWork
    {
        DT *r = nullptr, *p = nullptr, *v = nullptr;
        DT * p2 = nullptr;                // the other direction buffer of cg_spmv_dir_kernel
        double * partials = nullptr;
        double * partials2 = nullptr;     // a second set, for kernels that read one set and write the other
        CgScalars<DT> * s = nullptr;
        void * block = nullptr;
        bool owned = false;

        static size_t vector_bytes(uint64_t n) { return ((sizeof(DT) * n + 255) / 256) * 256; }
        static size_t bytes_for(uint64_t n)
        {
            return 4 * vector_bytes(n) + sizeof(double) * 4 * kSolverBlocks + sizeof(CgScalars<DT>) + 256;
        }

        explicit CgWork(uint64_t n, void * external = nullptr)
        {
            const size_t vb = vector_bytes(n);
            if (external)
                block = external;
            else
            {
                hipError_t e = hipMalloc(&block, bytes_for(n));
                if (e != hipSuccess)
                    throw Error(HSOLVE_ERR_ALLOC, "CG workspace allocation failed");
                owned = true;
            }
            char * c = (char *)block;
            r = (DT *)c; c += vb;
            p = (DT *)c; c += vb;
            v = (DT *)c; c += vb;
            p2 = (DT *)c; c += vb;
            partials = (double *)c; c += sizeof(double) * 2 * kSolverBlocks;
            partials2 = (double *)c; c += sizeof(double) * 2 * kSolverBlocks;
            s = (CgScalars<DT> *)c;
        }
        ~CgWork()
        {
            if (owned && block)
                (void)hipFree(block);
        }
    };

    // work: NULL (the workspace is allocated for this solve) or a device
    // block of at least CgWork<DT>::bytes_for(n) bytes
    template <typename DT, typename Op, bool VAR>
    void run_cg(const Op & A, uint64_t n, const DT * b, DT * x, const DT * P, uint64_t max_iters, double eps,
                uint64_t * used_iters, double * defect_out, hipStream_t st, int * converged_out, int plain = 0,
                void * work = nullptr)
    {
        CgWork<DT> w(n, work);
        CgScalars<DT> h;
        std::memset(&h, 0, sizeof(h));
        h.eps_relative = (DT)eps;
        h.max_iters = max_iters;
        h.plain = plain;
        HIP_CHECK(hipMemcpyAsync(w.s, &h, sizeof(h), hipMemcpyHostToDevice, st));
        const unsigned grid = std::min<unsigned>(grid_for(n, kBlock), kSolverBlocks);
        const bool fused = grid <= 512;
        hipLaunchKernelGGL((cg_init_kernel<DT, Op, VAR>), dim3(grid), dim3(kBlock), 0, st, A, b, x, w.r, w.p, P,
                           (int64_t)n, w.partials);
        hipLaunchKernelGGL((cg_init_finish<DT, VAR>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials, grid,
                           w.s);
        HIP_CHECK(hipGetLastError());
        // chunks of iterations between host checks; every kernel of an
        // iteration that starts after the stop is a no-op
        uint64_t enqueued = 0;
        uint64_t chunk = 8;
        if (fused)
        {
            // small grids (launch-bound solves), three launches per
            // iteration: the SpMV forms the direction itself, the update forms
            // lambda from the SpMV's partials.  pv[cp] holds the current
            // direction.
            DT * pv[2] = {w.p, w.p2};
            int cp = 0;
            bool first = true;      // pv[0] still holds the direction cg_init_kernel formed
            while (true)
            {
                for (uint64_t c = 0; c < chunk && enqueued < max_iters; ++c, ++enqueued)
                {
                    if (first)
                    {
                        hipLaunchKernelGGL((cg_spmv_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A, pv[cp],
                                           w.v, (int64_t)n, w.partials, (const CgScalars<DT> *)w.s);
                        first = false;
                    }
                    else
                    {
                        hipLaunchKernelGGL((cg_spmv_dir_kernel<DT, Op, VAR>), dim3(grid), dim3(kBlock), 0, st, A,
                                           (const DT *)pv[cp], (const DT *)w.r, P, pv[cp ^ 1], w.v, (int64_t)n,
                                           w.partials, (const CgScalars<DT> *)w.s);
                        cp ^= 1;
                    }
                    hipLaunchKernelGGL((cg_update_kernel<DT, VAR, true>), dim3(grid), dim3(kBlock), 0, st, x, w.r,
                                       pv[cp], w.v, P, (int64_t)n, w.partials2, w.s, (const double *)w.partials,
                                       grid);
                    hipLaunchKernelGGL((cg_update_finish<DT, VAR>), dim3(1), dim3(kBlock), 0, st,
                                       (const double *)w.partials2, grid, w.s);
                }
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipMemcpyAsync(&h, w.s, sizeof(h), hipMemcpyDeviceToHost, st));
                HIP_CHECK(hipStreamSynchronize(st));
                if (h.run == 0 || enqueued >= max_iters)
                    break;
                if (chunk < 64)
                    chunk *= 2;
            }
        }
        else
        {
            while (true)
            {
                for (uint64_t c = 0; c < chunk && enqueued < max_iters; ++c, ++enqueued)
                {
                    hipLaunchKernelGGL((cg_spmv_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A, w.p, w.v,
                                       (int64_t)n, w.partials, (const CgScalars<DT> *)w.s);
                    hipLaunchKernelGGL((cg_spmv_finish<DT, VAR>), dim3(1), dim3(kBlock), 0, st,
                                       (const double *)w.partials, grid, w.s);
                    hipLaunchKernelGGL((cg_update_kernel<DT, VAR>), dim3(grid), dim3(kBlock), 0, st, x, w.r, w.p,
                                       w.v, P, (int64_t)n, w.partials, w.s, (const double *)nullptr, 0u);
                    hipLaunchKernelGGL((cg_update_finish<DT, VAR>), dim3(1), dim3(kBlock), 0, st,
                                       (const double *)w.partials, grid, w.s);
                    hipLaunchKernelGGL((cg_direction_kernel<DT, VAR>), dim3(grid), dim3(kBlock), 0, st, w.p, w.r,
                                       P, (int64_t)n, (const CgScalars<DT> *)w.s);
                }
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipMemcpyAsync(&h, w.s, sizeof(h), hipMemcpyDeviceToHost, st));
                HIP_CHECK(hipStreamSynchronize(st));
                if (h.run == 0 || enqueued >= max_iters)
                    break;
                if (chunk < 64)
                    chunk *= 2;
            }
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

    // ConjugateGradients<JAC> on an ELL handle: CgWork's r, p, v hold r, d, t;
    // z has its own vector.
    template <typename DT, typename Op>
    void run_pcg_jac(const Op & A, uint64_t n, const DT * b, DT * x, const DT * dd, uint64_t max_iters, double eps,
                     uint64_t * used_iters, double * defect_out, hipStream_t st, int * converged_out)
    {
        CgWork<DT> w(n);
        DeviceArray<DT> zbuf(n);
        DT * z = zbuf.p;
        CgScalars<DT> h;
        std::memset(&h, 0, sizeof(h));
        h.eps_relative = (DT)eps;
        h.max_iters = max_iters;
        HIP_CHECK(hipMemcpyAsync(w.s, &h, sizeof(h), hipMemcpyHostToDevice, st));
        const unsigned grid = std::min<unsigned>(grid_for(n, kBlock), kSolverBlocks);
        hipLaunchKernelGGL((jac_init_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A, b, (const DT *)x, dd, w.r,
                           z, w.p, (int64_t)n, w.partials);
        hipLaunchKernelGGL((jac_init_finish<DT>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials, grid,
                           w.s);
        HIP_CHECK(hipGetLastError());
        uint64_t enqueued = 0, chunk = 8;
        // small grids (launch-bound solves): three launches per iteration, as
        // run_cg; pv[cp] holds the current direction
        DT * pv[2] = {w.p, w.p2};
        int cp = 0;
        bool first = true;
        while (grid <= 512)
        {
            for (uint64_t c = 0; c < chunk && enqueued < max_iters; ++c, ++enqueued)
            {
                if (first)
                {
                    hipLaunchKernelGGL((cg_spmv_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A,
                                       (const DT *)pv[cp], w.v, (int64_t)n, w.partials, (const CgScalars<DT> *)w.s);
                    first = false;
                }
                else
                {
                    hipLaunchKernelGGL((jac_spmv_dir_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A,
                                       (const DT *)pv[cp], (const DT *)z, pv[cp ^ 1], w.v, (int64_t)n, w.partials,
                                       (const CgScalars<DT> *)w.s);
                    cp ^= 1;
                }
                hipLaunchKernelGGL((jac_update_kernel<DT, true>), dim3(grid), dim3(kBlock), 0, st, x, w.r, z,
                                   (const DT *)pv[cp], (const DT *)w.v, dd, (int64_t)n, w.partials2, w.s,
                                   (const double *)w.partials, grid);
                hipLaunchKernelGGL((jac_update_finish<DT>), dim3(1), dim3(kBlock), 0, st,
                                   (const double *)w.partials2, grid, w.s);
            }
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(&h, w.s, sizeof(h), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            if (h.run == 0 || enqueued >= max_iters)
                break;
            if (chunk < 64)
                chunk *= 2;
        }
        while (grid > 512)
        {
            for (uint64_t c = 0; c < chunk && enqueued < max_iters; ++c, ++enqueued)
            {
                hipLaunchKernelGGL((cg_spmv_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A, (const DT *)w.p,
                                   w.v, (int64_t)n, w.partials, (const CgScalars<DT> *)w.s);
                hipLaunchKernelGGL((jac_spmv_finish<DT>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials,
                                   grid, w.s);
                hipLaunchKernelGGL((jac_update_kernel<DT>), dim3(grid), dim3(kBlock), 0, st, x, w.r, z,
                                   (const DT *)w.p, (const DT *)w.v, dd, (int64_t)n, w.partials, w.s,
                                   (const double *)nullptr, 0u);
                hipLaunchKernelGGL((jac_update_finish<DT>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials,
                                   grid, w.s);
                hipLaunchKernelGGL((jac_direction_kernel<DT>), dim3(grid), dim3(kBlock), 0, st, w.p, (const DT *)z,
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

    template <typename DT, bool VAR>
    void run_cg_ell(const hsolve_ell_s * m, const DT * b, DT * x, const DT * P, uint64_t max_iters, double eps,
                    uint64_t * used, double * defect, hipStream_t st, int * conv, int plain = 0,
                    void * work = nullptr)
    {
        const uint64_t n = m->rows;
        with_ell_types(m, [&](auto Aj, auto Arl) {
            using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
            using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
            EllOp<DT, IT, LT> A{Aj, (const DT *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            run_cg<DT, EllOp<DT, IT, LT>, VAR>(A, n, b, x, P, max_iters, eps, used, defect, st, conv, plain, work);
        });
    }

    // ------------------------- CGSolver<VAR> with a SparseMatrixELL P
    // (cg.hh:141-219): z = Product(P, r) by the handle's SpMV, then the
    // partials of r.r and r.z in a separate pass.

    // partials of (r.r, r.z)
    template <typename DT>
    __global__ __launch_bounds__(256) void pm_dot_kernel(const DT * __restrict__ r, const DT * __restrict__ z,
            int64_t n, double * partials, const CgScalars<DT> * s)
    {
        if (s && s->run == 0)
            return;
        double acc_rr = 0.0, acc_rz = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            const double ri = (double)r[i];
            acc_rr += ri * ri;
            acc_rz += ri * (double)z[i];
        }
        k::put_partials2(acc_rr, acc_rz, partials, blockIdx.x);
    }

    // after x += lambda p, r -= lambda v: ||r|| and the stopping tests
    template <typename DT>
    __global__ __launch_bounds__(256) void pm_test_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double rr, unused;
        k::sum_partials2(partials, nb, rr, unused);
        if (threadIdx.x != 0)
            return;
        ++s->iters;
        const DT cur = (DT)sqrt((DT)rr);
        s->current_defect = cur;
        const DT eps = s->eps_relative;
        if (cur < eps * s->initial_defect || cur < eps)
        {
            s->converged = 1;
            s->run = 0;
        }
        else if (s->iters == s->max_iters)
            s->run = 0;
    }

    // alpha = alpha_new; alpha_new = r.z; talpha = alpha_new / alpha (guarded)
    template <typename DT>
    __global__ __launch_bounds__(256) void pm_dir_finish(const double * partials, unsigned nb, CgScalars<DT> * s)
    {
        if (s->run == 0)
            return;
        double unused, rz;
        k::sum_partials2(partials, nb, unused, rz);
        if (threadIdx.x != 0)
            return;
        s->alpha = s->alpha_new;
        s->alpha_new = (DT)rz;
        s->talpha = s->alpha_new / guard_signed<DT>(s->alpha);
    }

    template <typename DT, typename Op>
    void run_pcg_pmat(const Op & A, hsolve_ell Pm, uint64_t n, const DT * b, DT * x, uint64_t max_iters, double eps,
                      uint64_t * used_iters, double * defect_out, hipStream_t st, int * converged_out)
    {
        CgWork<DT> w(n);
        DeviceArray<DT> zbuf(n);
        DT * z = zbuf.p;
        CgScalars<DT> h;
        std::memset(&h, 0, sizeof(h));
        h.eps_relative = (DT)eps;
        h.max_iters = max_iters;
        HIP_CHECK(hipMemcpyAsync(w.s, &h, sizeof(h), hipMemcpyHostToDevice, st));
        const unsigned grid = std::min<unsigned>(grid_for(n, kBlock), kSolverBlocks);
        auto precon = [&](const DT * in, DT * out) {
            const int rc = hsolve_ell_spmv(Pm, in, out, nullptr, (void *)st);
            if (rc != HSOLVE_OK)
                throw Error(rc, std::string("pcg_ell_pmat: preconditioner product failed: ") + hsolve_last_error());
        };
        // r = b - A x; p = P r; initial ||r|| and alpha_new = r.p
        hipLaunchKernelGGL((cg_init_kernel<DT, Op, false>), dim3(grid), dim3(kBlock), 0, st, A, b, (const DT *)x,
                           w.r, w.p, (const DT *)nullptr, (int64_t)n, w.partials);
        precon(w.r, w.p);
        hipLaunchKernelGGL((pm_dot_kernel<DT>), dim3(grid), dim3(kBlock), 0, st, (const DT *)w.r, (const DT *)w.p,
                           (int64_t)n, w.partials, (const CgScalars<DT> *)nullptr);
        hipLaunchKernelGGL((cg_init_finish<DT, true>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials, grid,
                           w.s);
        HIP_CHECK(hipGetLastError());
        uint64_t enqueued = 0, chunk = 8;
        while (true)
        {
            for (uint64_t c = 0; c < chunk && enqueued < max_iters; ++c, ++enqueued)
            {
                hipLaunchKernelGGL((cg_spmv_kernel<DT, Op>), dim3(grid), dim3(kBlock), 0, st, A, (const DT *)w.p,
                                   w.v, (int64_t)n, w.partials, (const CgScalars<DT> *)w.s);
                hipLaunchKernelGGL((cg_spmv_finish<DT, true>), dim3(1), dim3(kBlock), 0, st,
                                   (const double *)w.partials, grid, w.s);
                hipLaunchKernelGGL((cg_update_kernel<DT, false>), dim3(grid), dim3(kBlock), 0, st, x, w.r,
                                   (const DT *)w.p, (const DT *)w.v, (const DT *)nullptr, (int64_t)n, w.partials, w.s,
                                   (const double *)nullptr, 0u);
                hipLaunchKernelGGL((pm_test_finish<DT>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials,
                                   grid, w.s);
                // z = P r (after a stop this writes z only, which nothing reads)
                precon(w.r, z);
                hipLaunchKernelGGL((pm_dot_kernel<DT>), dim3(grid), dim3(kBlock), 0, st, (const DT *)w.r,
                                   (const DT *)z, (int64_t)n, w.partials, (const CgScalars<DT> *)w.s);
                hipLaunchKernelGGL((pm_dir_finish<DT>), dim3(1), dim3(kBlock), 0, st, (const double *)w.partials,
                                   grid, w.s);
                // p = p * talpha + z
                hipLaunchKernelGGL((cg_direction_kernel<DT, false>), dim3(grid), dim3(kBlock), 0, st, w.p,
                                   (const DT *)z, (const DT *)nullptr, (int64_t)n, (const CgScalars<DT> *)w.s);
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

    // ------------------------------------------ allocation-free CG workspace
    // hsolve_cg_ell_noalloc carves its CG workspace (four vectors, two sets
    // of reduction partials, the scalars) from this fixed array instead of
    // allocating it.  Every device the library runs on holds its own copy
    // until the process ends; every call on a device shares that copy, so
    // calls that overlap on one device (other threads, other streams) race.
    constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
    static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

    constexpr int kMaxDevices = 64;
    std::mutex g_scratch_mutex;
    void * g_scratch_ptr[kMaxDevices] = {};

    // The device address of the current device's scratch, looked up once per
    // device with that device current.  (The array's name in host code is a
    // host-side stand-in, never a device address.)
    void * noalloc_scratch()
    {
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        require(dev >= 0 && dev < kMaxDevices, HSOLVE_ERR_INVALID, "cg_ell_noalloc: device index out of range");
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (g_scratch_ptr[dev] == nullptr)
        {
            void * ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || ptr == nullptr)
                throw Error(HSOLVE_ERR_HIP, std::string("cg_ell_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                                                hipGetErrorString(e));
            g_scratch_ptr[dev] = ptr;
        }
        return g_scratch_ptr[dev];
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

int hsolve_pcg_jac_ell(hsolve_ell m, const void * dd_inverted, const void * b, void * x, uint64_t max_iters,
                       double eps, uint64_t * used_iters, double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && dd_inverted && b && x, HSOLVE_ERR_INVALID, "pcg_jac_ell: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "pcg_jac_ell: matrix is not square");
    int conv = 0;
    hipStream_t st = as_stream(stream);
    with_ell_types(m, [&](auto Aj, auto Arl) {
        using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
        using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
        if (m->dtype)
        {
            EllOp<double, IT, LT> A{Aj, (const double *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            run_pcg_jac<double>(A, m->rows, (const double *)b, (double *)x, (const double *)dd_inverted, max_iters,
                                eps, used_iters, defect_out, st, &conv);
        }
        else
        {
            EllOp<float, IT, LT> A{Aj, (const float *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            run_pcg_jac<float>(A, m->rows, (const float *)b, (float *)x, (const float *)dd_inverted, max_iters,
                               eps, used_iters, defect_out, st, &conv);
        }
    });
    if (! conv && max_iters > 0)
    {
        set_last_error("pcg_jac_ell: max_iters reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

int hsolve_pcg_ell_pmat(hsolve_ell m, hsolve_ell P, const void * b, void * x, uint64_t max_iters, double eps,
                        uint64_t * used_iters, double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && P && b && x, HSOLVE_ERR_INVALID, "pcg_ell_pmat: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "pcg_ell_pmat: matrix is not square");
    require(P->rows == m->rows && P->cols == m->rows, HSOLVE_ERR_SIZE,
            "pcg_ell_pmat: preconditioner size differs from the matrix");
    require(P->dtype == m->dtype, HSOLVE_ERR_INVALID, "pcg_ell_pmat: preconditioner value type differs");
    int conv = 0;
    hipStream_t st = as_stream(stream);
    with_ell_types(m, [&](auto Aj, auto Arl) {
        using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
        using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
        if (m->dtype)
        {
            EllOp<double, IT, LT> A{Aj, (const double *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            run_pcg_pmat<double>(A, P, m->rows, (const double *)b, (double *)x, max_iters, eps, used_iters,
                                 defect_out, st, &conv);
        }
        else
        {
            EllOp<float, IT, LT> A{Aj, (const float *)m->Ax, Arl, m->stride, (uint32_t)m->threads};
            run_pcg_pmat<float>(A, P, m->rows, (const float *)b, (float *)x, max_iters, eps, used_iters, defect_out,
                                st, &conv);
        }
    });
    if (! conv && max_iters > 0)
    {
        set_last_error("pcg_ell_pmat: max_iters reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

int hsolve_cg_iters_ell(hsolve_ell m, const void * rhs, void * x, uint64_t max_iters, uint64_t * used_iters,
                        double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && rhs && x, HSOLVE_ERR_INVALID, "cg_iters_ell: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "cg_iters_ell: matrix is not square");
    int conv = 0;
    if (m->dtype)
        run_cg_ell<double, false>(m, (const double *)rhs, (double *)x, nullptr, max_iters, 1e-8, used_iters,
                                  defect_out, as_stream(stream), &conv, 1);
    else
        run_cg_ell<float, false>(m, (const float *)rhs, (float *)x, nullptr, max_iters, 1e-8, used_iters,
                                 defect_out, as_stream(stream), &conv, 1);
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

int hsolve_cg_ell_noalloc(hsolve_ell m, const void * b, void * x, uint64_t max_iters, double eps,
                          uint64_t * used_iters, double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && b && x, HSOLVE_ERR_INVALID, "cg_ell_noalloc: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "cg_ell_noalloc: matrix is not square");
    const uint64_t n = m->rows;
    const size_t need = m->dtype ? CgWork<double>::bytes_for(n) : CgWork<float>::bytes_for(n);
    if (need > kNoAllocScratchBytes)
        throw Error(HSOLVE_ERR_SIZE, "cg_ell_noalloc: " + std::to_string(n) + " rows need " + std::to_string(need) +
                                         " bytes of workspace, above the fixed scratch of " +
                                         std::to_string(kNoAllocScratchBytes) + " bytes; use hsolve_cg_ell");
    void * work = noalloc_scratch();
    int conv = 0;
    if (m->dtype)
        run_cg_ell<double, false>(m, (const double *)b, (double *)x, nullptr, max_iters, eps, used_iters, defect_out,
                                  as_stream(stream), &conv, 0, work);
    else
        run_cg_ell<float, false>(m, (const float *)b, (float *)x, nullptr, max_iters, eps, used_iters, defect_out,
                                 as_stream(stream), &conv, 0, work);
    if (! conv && max_iters > 0)
    {
        set_last_error("cg_ell_noalloc: max_iters reached");
        return HSOLVE_ERR_NOT_CONVERGED;
    }
    HSOLVE_API_END
}

}  // extern "C"
