// kernels_mg.hh -- gfx950 kernels of the multigrid layer: Q1 transfers
// (restriction, prolongation, prolongation + correction), the fused first
// Jacobi sweeps and the one-workgroup coarse-grid CG.
//
// The transfers are written as gathers (one lane per output node) but
// replay, per node, the exact sequence of additions that the reference's
// scattering loops perform (honei/math/restriction.hh:43-172,
// prolongation.hh:43-139), in double as the reference does (its constants
// 0.5/0.25 are double), so results are bit-identical to tags::CPU.
#pragma once

#include "kernels_la.hh"

namespace hsolve
{
namespace k
{
    struct Mask
    {
        uint8_t dir[8];   // 1 where macro_border_mask[i] == 2 (Dirichlet)
    };

    // ApplyDirichletBoundaries (apply_dirichlet_boundaries.hh:31-80) as a
    // per-node predicate on a side-M grid.
    __device__ __forceinline__ bool dirichlet_node(int64_t k, int64_t M, const Mask & mk)
    {
        const int64_t N = M * M;
        if (mk.dir[0] && k == 0) return true;
        if (mk.dir[1] && k == M - 1) return true;
        if (mk.dir[2] && k == N - 1) return true;
        if (mk.dir[3] && k == N - M) return true;
        if (mk.dir[4] && k >= 1 && k < M - 1) return true;
        if (mk.dir[5] && k >= 2 * M - 1 && k < N - M && (k % M) == M - 1) return true;
        if (mk.dir[6] && k >= N - M + 1 && k < N - 1) return true;
        if (mk.dir[7] && k >= M && k < N - M && (k % M) == 0) return true;
        return false;
    }

    template <typename DT>
    __device__ __forceinline__ DT add_d(DT acc, double w, DT v)
    {
        return (DT)((double)acc + w * (double)v);
    }

    // Restriction<Tag, NONE>: one lane per coarse node (r, s).  The
    // reference visits the node in its outer loop iterations i = r, r+1, s,
    // s+1 (those in [1, Nc-1]) and, inside one iteration, in the statement
    // order (a)..(h) below.
    template <typename DT>
    __global__ __launch_bounds__(256) void restriction_kernel(DT * __restrict__ coarse,
            const DT * __restrict__ fine, int64_t Nc, Mask mk, const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (idx >= Nc * Nc)
            return;
        const int64_t r = idx / Nc, s = idx - r * Nc, Nf = 2 * Nc - 1;
        const DT * f = fine;
        DT v = f[(2 * r) * Nf + 2 * s];
        int64_t c[4] = {r, r + 1, s, s + 1};
        // sort 4 values (r < r+1 and s < s+1 already)
        if (c[2] < c[0]) { int64_t t = c[0]; c[0] = c[2]; c[2] = t; }
        if (c[3] < c[1]) { int64_t t = c[1]; c[1] = c[3]; c[3] = t; }
        if (c[1] < c[0]) { int64_t t = c[0]; c[0] = c[1]; c[1] = t; }
        if (c[3] < c[2]) { int64_t t = c[2]; c[2] = c[3]; c[3] = t; }
        if (c[2] < c[1]) { int64_t t = c[1]; c[1] = c[2]; c[2] = t; }
        int64_t prev = -1;
        for (int q = 0; q < 4; ++q)
        {
            const int64_t i = c[q];
            if (i == prev)
                continue;
            prev = i;
            if (i < 1 || i > Nc - 1)
                continue;
            if (r == i)     v = add_d(v, 0.5, f[(2 * r - 1) * Nf + 2 * s]);         // (a)
            if (r == i - 1) v = add_d(v, 0.5, f[(2 * r + 1) * Nf + 2 * s]);         // (b)
            if (s == i - 1) v = add_d(v, 0.5, f[(2 * r) * Nf + 2 * s + 1]);         // (c)
            if (s == i)     v = add_d(v, 0.5, f[(2 * r) * Nf + 2 * s - 1]);         // (d)
            if (r == i - 1 && s <= Nc - 2) v = add_d(v, 0.25, f[(2 * r + 1) * Nf + 2 * s + 1]); // (e)
            if (r == i - 1 && s >= 1)      v = add_d(v, 0.25, f[(2 * r + 1) * Nf + 2 * s - 1]); // (f)
            if (r == i && s <= Nc - 2)     v = add_d(v, 0.25, f[(2 * r - 1) * Nf + 2 * s + 1]); // (g)
            if (r == i && s >= 1)          v = add_d(v, 0.25, f[(2 * r - 1) * Nf + 2 * s - 1]); // (h)
        }
        if (dirichlet_node(idx, Nc, mk))
            v = DT(0);
        coarse[idx] = v;
    }

    // Value of the bilinear prolongation at fine node (R, S) before the
    // Dirichlet mask, in the reference's addition order.
    template <typename DT>
    __device__ __forceinline__ DT prolongate_node(const DT * __restrict__ c, int64_t R, int64_t S,
                                                  int64_t Nc)
    {
        const bool re = (R & 1) == 0, se = (S & 1) == 0;
        if (re && se)
            return c[(R >> 1) * Nc + (S >> 1)];
        DT v = DT(0);
        if (re)
        {   // fine row 2j, column 2i-1
            const int64_t j = R >> 1, i = (S + 1) >> 1;
            v = add_d(v, 0.5, c[j * Nc + i - 1]);
            v = add_d(v, 0.5, c[j * Nc + i]);
        }
        else if (se)
        {   // fine row 2i-1, column 2j
            const int64_t i = (R + 1) >> 1, j = S >> 1;
            v = add_d(v, 0.5, c[(i - 1) * Nc + j]);
            v = add_d(v, 0.5, c[i * Nc + j]);
        }
        else
        {   // fine row 2i-1, column 2j+1
            const int64_t i = (R + 1) >> 1, j = (S - 1) >> 1;
            v = add_d(v, 0.25, c[(i - 1) * Nc + j]);
            v = add_d(v, 0.25, c[(i - 1) * Nc + j + 1]);
            v = add_d(v, 0.25, c[i * Nc + j]);
            v = add_d(v, 0.25, c[i * Nc + j + 1]);
        }
        return v;
    }

    // AXPY == false: fine = P coarse (Prolongation<Tag, NONE>).
    // AXPY == true:  fine += alpha * (P coarse) (Prolongation then
    //                ScaledSum(x, c, alpha): x = x + alpha*c).
    template <typename DT, bool AXPY>
    __global__ __launch_bounds__(256) void prolongation_kernel(DT * __restrict__ fine,
            const DT * __restrict__ coarse, int64_t Nf, DT alpha, Mask mk,
            const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (idx >= Nf * Nf)
            return;
        const int64_t R = idx / Nf, S = idx - R * Nf, Nc = (Nf + 1) / 2;
        DT v = prolongate_node<DT>(coarse, R, S, Nc);
        if (dirichlet_node(idx, Nf, mk))
            v = DT(0);
        if (AXPY)
            fine[idx] = fine[idx] + alpha * v;
        else
            fine[idx] = v;
    }

    template <typename DT>
    __global__ __launch_bounds__(256) void dirichlet_kernel(DT * __restrict__ grid, int64_t M, Mask mk)
    {
        const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (idx >= M * M)
            return;
        if (dirichlet_node(idx, M, mk))
            grid[idx] = DT(0);
    }

    // First two damped-Jacobi sweeps from a zero start in one pass:
    //   c1 = dinv * d                         (Jacobi MG overload (1), ElementProduct)
    //   c2 = c1 + (d - A c1) * dinv           (one sweep with rhs d)
    // with c1 recomputed at the neighbours instead of stored.  ACC selects
    // the output: 0 -> out = c2, 1 -> out = out + c2 (Sum<Tag>(x, c)).
    template <typename DT, int ACC>
    __global__ __launch_bounds__(256) void jacobi_zero2_q1(Bands B, const DT * __restrict__ d,
            const DT * __restrict__ dinv, DT * __restrict__ out, int64_t n, int64_t m,
            const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const DT * const * b = reinterpret_cast<const DT * const *>(B.b);
#define HS_C1(j) (dinv[(j)] * d[(j)])
        DT s = ld_stream(b[4] + i) * HS_C1(i);
        if (i - m - 1 >= 0) s += ld_stream(b[0] + i) * HS_C1(i - m - 1);
        if (i - m >= 0)     s += ld_stream(b[1] + i) * HS_C1(i - m);
        if (i - m + 1 >= 0) s += ld_stream(b[2] + i) * HS_C1(i - m + 1);
        if (i - 1 >= 0)     s += ld_stream(b[3] + i) * HS_C1(i - 1);
        if (i + 1 < n)      s += ld_stream(b[5] + i) * HS_C1(i + 1);
        if (i + m - 1 < n)  s += ld_stream(b[6] + i) * HS_C1(i + m - 1);
        if (i + m < n)      s += ld_stream(b[7] + i) * HS_C1(i + m);
        if (i + m + 1 < n)  s += ld_stream(b[8] + i) * HS_C1(i + m + 1);
#undef HS_C1
        const DT di = dinv[i];
        const DT c1 = di * d[i];
        const DT t = d[i] - s;
        const DT c2 = c1 + t * di;
        if (ACC)
            out[i] = out[i] + c2;
        else
            out[i] = c2;
    }

    // out = dinv * d (ACC 0) or out += dinv * d (ACC 1): the single first
    // sweep when only one smoothing step is configured.
    template <typename DT, int ACC>
    __global__ __launch_bounds__(256) void jacobi_zero1(const DT * __restrict__ d,
            const DT * __restrict__ dinv, DT * __restrict__ out, int64_t n,
            const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const DT c1 = dinv[i] * d[i];
        if (ACC)
            out[i] = out[i] + c1;
        else
            out[i] = c1;
    }

    // Sequential sum of a[i]*b[i] by one thread, as DotProduct<tags::CPU>.
    template <typename DT>
    __device__ __forceinline__ void seq_dot(const DT * a, const DT * b, int64_t n, DT * dst)
    {
        if (threadIdx.x == 0)
        {
            DT r = DT(0);
            for (int64_t i = 0; i < n; ++i)
                r += a[i] * b[i];
            *dst = r;
        }
        __syncthreads();
    }

    template <typename DT> struct Eps;
    template <> struct Eps<float> { static constexpr float v = 1.19209290e-07F; };
    template <> struct Eps<double> { static constexpr double v = 2.2204460492503131e-16; };

    // ConjugateGradients<Tag, NONE>::value(Q1, rhs, x, konv_rad) in one
    // workgroup (conjugate_gradients.hh:490-526, kernel :108-151).  Vectors
    // g, u, energy, new_gradient live in `work` (4n); reductions are the
    // sequential CPU sums, so the iteration is bit-identical to tags::CPU.
    // `active` (may be null): when it points to 0 the kernel does nothing.
    template <typename DT>
    __global__ __launch_bounds__(256) void cg_konv_q1_kernel(Bands B, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, int64_t m, double konv_rad,
            uint64_t max_iters, uint64_t * iters_out, const int * active)
    {
        if (active && *active == 0)
            return;
        __shared__ DT sh[4];
        DT * g = work;
        DT * u = work + n;
        DT * en = work + 2 * n;
        DT * ng = work + 3 * n;
        const DT eps = Eps<DT>::v;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
            x[i] = DT(0);
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        {
            const DT p = q1_row<DT>(B, x, i, n, m);
            g[i] = p - rhs[i];
        }
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
            u[i] = g[i] * DT(-1.);
        seq_dot<DT>(x, x, n, &sh[0]);
        DT norm_x = sh[0];
        DT norm_x_last = norm_x - DT(1);
        uint64_t it = 0;
        while ((double)(norm_x - norm_x_last) > konv_rad && it < max_iters)
        {
            seq_dot<DT>(g, g, n, &sh[1]);
            DT upper = sh[1];
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
                en[i] = q1_row<DT>(B, u, i, n, m);
            __syncthreads();
            seq_dot<DT>(en, u, n, &sh[2]);
            DT lower = sh[2];
            const DT f1 = (fabs(lower) >= eps) ? DT(upper / lower) : DT(upper / eps);
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
            {
                const DT t = u[i] * f1;
                en[i] = t + x[i];
            }
            __syncthreads();
            lower = upper;
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
            {
                const DT p = q1_row<DT>(B, en, i, n, m);
                ng[i] = p - rhs[i];
            }
            __syncthreads();
            seq_dot<DT>(ng, ng, n, &sh[3]);
            upper = sh[3];
            const DT f2 = (fabs(lower) >= eps) ? DT(upper / lower) : DT(upper / eps);
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
            {
                const DT t = u[i] * f2;
                u[i] = t - ng[i];
                g[i] = ng[i];
                x[i] = en[i];
            }
            __syncthreads();
            ++it;
            norm_x_last = norm_x;
            seq_dot<DT>(x, x, n, &sh[0]);
            norm_x = sh[0];
            __syncthreads();
        }
        if (iters_out && threadIdx.x == 0)
            *iters_out = it;
    }
}
}
