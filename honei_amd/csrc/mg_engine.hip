// mg_engine.hip -- kernels and tick sequence of the FP32-MG / FP64-IR engine
// (declarations and the slot/tick model: mg_engine.hh).
//
// Reference: Multigrid<Tag, OuterTag, NONE, JAC, CYCLE::V, MIXED>::value
// (honei/math/multigrid.hh:2245-2393) around the FP32 V-cycle
// (multigrid.hh:1616-1905):
//   outer: x0 with b's Dirichlet values injected (:2277-2298); result = 0;
//          d = b - A x0 (FP64); loop { rhs32 = (float)(d / |d|) (unscaled
//          the first time); inner FP32 MG solve; result += |d| * x32;
//          d = b - A result; stop when |d| < eps |d0| }
//   inner: |d| of the initial FP32 defect, V-cycles until
//          |d_k| <= tol |d_0| or n_max_iter cycles
//   cycle: per level down: Jacobi sweeps from zero (x = dinv*d, then
//          n_pre_smooth-1 sweeps), defect, restriction; coarse
//          ConjugateGradients<NONE> (konv_rad form); per level up:
//          prolongation + correction, n_pre_smooth post-sweeps, defect on
//          the finest level.
#include "mg_engine.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace hsolve
{
namespace mg
{
namespace
{
    __device__ __forceinline__ bool live(const Slot & sl) { return sl.s != nullptr; }

    // Virtual block coordinates of a slot launch (grid x [, y] per slot,
    // blockIdx.z = slot).  The hardware hands out workgroups in linear order
    // (x fastest), dealing consecutive ones to the 8 XCDs in turn.  Logical
    // blocks are ordered row block major, slot fastest, and every XCD gets
    // one contiguous range of them (k::xcd_block): all slots of a row block
    // run on one XCD, one after another, so the level matrix that they
    // share is fetched into that XCD's L2 once per row block, and the rows
    // +-side away that a stencil row gathers stay in the same L2.
    struct VB
    {
        unsigned x, y, z;
    };
    __device__ __forceinline__ VB vblock()
    {
        if (gridDim.z == 1)     // one slot: the hardware order
            return VB{blockIdx.x, blockIdx.y, 0u};
        const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const unsigned L = k::xcd_block(lin, gridDim.x * gridDim.y * gridDim.z);
        const unsigned z = L % gridDim.z, r = L / gridDim.z;
        return VB{r % gridDim.x, r / gridDim.x, z};
    }

    // Row i of y = A x' for an engine ELL matrix, x'_c = gx(c), in the
    // reference's order (slots ascending, summed from zero:
    // honei/la/product.hh:148-206).  With 9 stored slots per row (the Q1
    // stencils) every slot load and gather is issued before the first add;
    // padding slots hold valid columns (checked at upload) and are loaded but
    // not summed.
    template <typename DT, typename GX>
    __device__ __forceinline__ DT ell_row_g(const EllMat<DT> & E, int64_t i, GX && gx)
    {
        const uint32_t l = E.len[i];
        DT s = DT(0);
        if (E.ncpr == 9)
        {
            uint32_t c[9];
            DT v[9], xv[9];
#pragma unroll
            for (int k = 0; k < 9; ++k)
            {
                const uint64_t j = (uint64_t)i + (uint64_t)k * E.stride;
                c[k] = E.col[j];
                v[k] = E.val[j];
            }
#pragma unroll
            for (int k = 0; k < 9; ++k)
                xv[k] = gx(c[k]);
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if ((uint32_t)k < l)
                    s += v[k] * xv[k];
            return s;
        }
        uint64_t j = (uint64_t)i;
        for (uint32_t k = 0; k < l; ++k, j += E.stride)
            s += E.val[j] * gx(E.col[j]);
        return s;
    }

    template <typename DT>
    __device__ __forceinline__ DT ell_row(const EllMat<DT> & E, const DT * __restrict__ x, int64_t i)
    {
        return ell_row_g<DT>(E, i, [&](uint32_t c) { return x[c]; });
    }

    struct LevelSizes
    {
        int64_t n[kMaxLevels];
    };

    // One level's matrix and smoother data for the kernels: Q1 bands (B, m =
    // grid side) or an ELL matrix (E), chosen by the kernels' MK argument.
    struct LevelArgs
    {
        k::Bands B;          // B.b[k] = B.b[0] + k * bs
        EllMat32 E;
        const float * dinv;
        int64_t n, m;
        int64_t bs;          // band stride in elements
        const uint64_t * codes = nullptr;   // value codes of bands and dinv, or null
        const float * ctab = nullptr;       // their tables (16 per array)
    };

    // Matrix kinds of the level kernels: MK 0 Q1 bands (Product<Q1>'s sum
    // order), MK 1 ELL matrices, MK 2 bands holding 9-point structured ELL
    // matrices (ascending-column order, the same bits as MK 1).
    template <int MK> constexpr bool kBands = MK != 1;
    template <int MK> constexpr int kOrd = MK == 2 ? 1 : 0;

    template <int MK>
    __device__ __forceinline__ float level_row(const LevelArgs & a, const float * __restrict__ x, int64_t i)
    {
        if (kBands<MK>)
            return k::q1_row<float, kOrd<MK>>(a.B, x, i, a.n, a.m);
        return ell_row<float>(a.E, x, i);
    }

    // the level operator for the one-workgroup coarse CG
    template <int MK> struct LevelRowOp
    {
        LevelArgs a;
        __device__ __forceinline__ float row(const float * __restrict__ x, int64_t i) const
        {
            return level_row<MK>(a, x, i);
        }
    };

    // ------------------------------------------------------------ tick gates
    __global__ void plan_kernel(const Slot * T, unsigned nslots)
    {
        const unsigned z = blockIdx.x * blockDim.x + threadIdx.x;
        if (z >= nslots)
            return;
        Scalars * s = T[z].s;
        if (s == nullptr)
            return;
        const int ph = s->phase;
        s->do_start = (ph == PH_START) ? 1 : 0;
        s->do_outer = (ph == PH_SOLVE && s->need_outer) ? 1 : 0;
        s->open = 0;
        if (ph == PH_START || ph == PH_SOLVE)
            s->ticks += 1;
    }

    // ------------------------------------------------------ start of a solve
    // result = initial guess (0 or the caller's) with b injected on the first
    // row, the last row and the first column (multigrid.hh:2282-2297);
    // x32 of the finest level = (float)result
    __global__ __launch_bounds__(256) void start_init_kernel(const Slot * T, int L, int64_t n, int64_t sq, int fx)
    {
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl) || sl.s->do_start == 0)
            return;
        for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        {
            double g = sl.s->zero_guess ? 0.0 : sl.result[i];
            if (i < sq || i >= n - sq || (i % sq) == 0)
                g = sl.b[i];
            sl.result[i] = g;
            sl.v[L][fx][i] = (float)g;
        }
    }

    // iterates of the coarser levels start at zero
    __global__ __launch_bounds__(256) void start_zero_kernel(const Slot * T, int minl, int L, LevelSizes sz)
    {
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl) || sl.s->do_start == 0)
            return;
        for (int l = minl; l < L; ++l)
            for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < sz.n[l]; i += (int64_t)gridDim.x * blockDim.x)
                sl.v[l][F_X][i] = 0.0f;
    }

    // ------------------------------------------------------- FP64 outer step
    // result += scale * (double)x32
    __global__ __launch_bounds__(256) void outer_accumulate_kernel(const Slot * T, int L, int64_t n, int fx)
    {
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl) || sl.s->do_outer == 0)
            return;
        const double scale = sl.s->scale;
        for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            sl.result[i] = sl.result[i] + scale * (double)sl.v[L][fx][i];
    }

    // od = b - A64 result and its norm; opens the next inner solve or ends
    // the outer loop.  On a start: the initial defect.  MK 0: Q1 bands B
    // (two rows per lane, 16-byte loads inside); MK 1: the ELL matrix E.
    //
    // CB: the interior rows read the bands as value codes (one 64-bit word
    // per row, band k in bits 4k .. 4k + 3) and look them up in ctab
    // (16 per band, copied to LDS); the same values, the same sums.
    template <int MK, bool CB = false>
    __global__ __launch_bounds__(256) void outer_defect_kernel(k::Bands B, EllMat64 E, const Slot * T, int64_t n,
            int64_t m, const uint64_t * __restrict__ codes, const double * __restrict__ ctab)
    {
        __shared__ double tab[CB ? 144 : 1];
        if constexpr (CB)
        {
            if (threadIdx.x < 144)
                tab[threadIdx.x] = ctab[threadIdx.x];
            __syncthreads();
        }
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl))
            return;
        Scalars * s = sl.s;
        const int start = s->do_start;
        if (start == 0 && s->do_outer == 0)
            return;
        const double * __restrict__ b = sl.b;
        const double * __restrict__ x = sl.result;
        double * __restrict__ od = sl.od;
        double acc = 0.0;
        if (MK == 1)
        {
            // one row per lane and pass, lanes on consecutive rows
            const int64_t stride1 = (int64_t)gridDim.x * blockDim.x;
            for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < n; i += stride1)
            {
                const double di = b[i] - ell_row<double>(E, x, i);
                od[i] = di;
                acc += di * di;
            }
        }
        const int64_t stride = (int64_t)gridDim.x * blockDim.x * 2;
        for (int64_t i0 = ((int64_t)vb.x * blockDim.x + threadIdx.x) * 2; kBands<MK> && i0 < n; i0 += stride)
        {
            if (i0 - m - 1 >= 0 && i0 + 2 + m < n)
            {
                // b and x are the caller's arrays (8-byte aligned), od is the slot's
                double prod[2];
                if constexpr (CB)
                {
                    // q1_rows_v's sums (Product<Q1>'s order: dd first) with
                    // the band values from the codes of rows i0, i0 + 1
                    typedef uint64_t U2 __attribute__((ext_vector_type(2)));
                    const U2 cw = __builtin_nontemporal_load((const U2 *)(codes + i0));
                    double xm[4], xc[4], xp[4];
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj)
                    {
                        xm[jj] = x[i0 - m - 1 + jj];
                        xc[jj] = x[i0 - 1 + jj];
                        xp[jj] = x[i0 + m - 1 + jj];
                    }
#pragma unroll
                    for (int v = 0; v < 2; ++v)
                    {
                        const uint64_t w = cw[v];
                        auto bk = [&](int k) { return tab[16 * k + (int)((w >> (4 * k)) & 15u)]; };
                        double sum = bk(4) * xc[v + 1];
                        sum += bk(0) * xm[v];
                        sum += bk(1) * xm[v + 1];
                        sum += bk(2) * xm[v + 2];
                        sum += bk(3) * xc[v];
                        sum += bk(5) * xc[v + 2];
                        sum += bk(6) * xp[v];
                        sum += bk(7) * xp[v + 1];
                        sum += bk(8) * xp[v + 2];
                        prod[v] = sum;
                    }
                }
                else
                    k::q1_rows_v<double, kOrd<MK>>(B, x, i0, m, prod);
                typedef k::Vec16<double>::type V2;
                V2 o;
                o[0] = __builtin_nontemporal_load(b + i0) - prod[0];
                o[1] = __builtin_nontemporal_load(b + i0 + 1) - prod[1];
                *(V2 *)(od + i0) = o;
                acc += o[0] * o[0];
                acc += o[1] * o[1];
            }
            else
            {
                for (int64_t i = i0; i < i0 + 2 && i < n; ++i)
                {
                    const double prod = k::q1_row<double, kOrd<MK>>(B, x, i, n, m);
                    const double di = b[i] - prod;
                    od[i] = di;
                    acc += di * di;
                }
            }
        }
        k::put_partials2(acc, 0.0, sl.partials, vb.x);
    }

    // |od| from the partials of outer_defect_kernel (nb blocks per slot), then
    // the outer step's decisions; one block per slot (blockIdx.z)
    __global__ __launch_bounds__(256) void outer_defect_finish_kernel(const Slot * T, unsigned nb, double outer_eps,
                                                                      uint64_t max_outer)
    {
        const Slot & sl = T[blockIdx.z];
        if (! live(sl))
            return;
        Scalars * s = sl.s;
        const int start = s->do_start;
        if (start == 0 && s->do_outer == 0)
            return;
        double dd, unused;
        k::sum_partials2(sl.partials, nb, dd, unused);
        if (threadIdx.x != 0)
            return;
        {
            const double nrm = sqrt(dd);
            if (start)
            {
                s->initial_defect = nrm;
                s->def_norm = nrm;
                s->scale = 1.0;
                s->inv = 1.0;
                s->outer_iters = 0;
                s->inner_total = 0;
                s->phase = PH_SOLVE;
                s->open = 1;
            }
            else
            {
                s->def_norm = nrm;
                s->outer_iters += 1;
                s->inner_total += s->inner_last;
                s->need_outer = 0;
                if (nrm < outer_eps * s->initial_defect)
                {
                    s->outer_active = 0;
                    s->phase = PH_DONE;
                }
                else
                {
                    s->scale = nrm;
                    s->inv = 1.0 / nrm;
                    if (max_outer != 0 && s->outer_iters >= max_outer)
                    {
                        s->outer_active = 0;
                        s->status = 1;
                        s->phase = PH_DONE;
                    }
                    else
                        s->open = 1;
                }
            }
        }
    }

    // rhs32 = (float)(od * inv) (inv == 1 on the first step, so the cast of
    // the unscaled defect); a start also clears the result array
    __global__ __launch_bounds__(256) void to_inner_kernel(const Slot * T, int L, int64_t n)
    {
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl) || sl.s->open == 0)
            return;
        const double inv = sl.s->inv;
        const bool clear = sl.s->do_start != 0;
        for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        {
            sl.v[L][F_RHS][i] = (float)(sl.od[i] * inv);
            if (clear)
                sl.result[i] = 0.0;
        }
    }

    // Opens the inner solve: its initial defect d is rhs itself when
    // initial_zero (multigrid.hh:1626-1628), else rhs - A x.
    template <int INITIAL_ZERO, int MK>
    __global__ __launch_bounds__(256) void open_inner_kernel(LevelArgs a, const Slot * T, int L, int fx)
    {
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl) || sl.s->open == 0)
            return;
        const int64_t n = a.n;
        const int64_t lane0 = (int64_t)vb.x * blockDim.x + threadIdx.x;
        const int64_t lanes = (int64_t)gridDim.x * blockDim.x;
        const float * rhs = sl.v[L][F_RHS];
        float * d = sl.v[L][F_D];
        if (INITIAL_ZERO)
        {
            typedef k::Vec16<float>::type F4;
            for (int64_t i0 = lane0 * 4; i0 < n; i0 += lanes * 4)
            {
                if (i0 + 4 <= n)
                    *(F4 *)(d + i0) = *(const F4 *)(rhs + i0);
                else
                    for (int64_t i = i0; i < n; ++i)
                        d[i] = rhs[i];
            }
        }
        else if (kBands<MK>)
        {
            for (int64_t lane = lane0; lane * 4 < n; lane += lanes)
                k::q1_body_v<float, 1, kOrd<MK>>(a.B, sl.v[L][fx], d, rhs, nullptr, n, a.m, lane);
        }
        else
        {
            const float * x = sl.v[L][fx];
            for (int64_t i = lane0; i < n; i += lanes)
                d[i] = rhs[i] - ell_row<float>(a.E, x, i);
        }
        if (vb.x == 0 && threadIdx.x == 0)
        {
            sl.s->run = 1;
            sl.s->need_outer = 0;
        }
    }

    // |d| of the finest FP32 defect.  MODE 0: start of an inner solve (gate
    // `open`); MODE 1: end of a V-cycle (gate `run`): count it, stop on
    // |d| <= tol |d_0| or after n_max_iter cycles.
    template <int MODE>
    __global__ __launch_bounds__(256) void inner_norm_kernel(const Slot * T, int L, int64_t n)
    {
        const VB vb = vblock();
        const Slot & sl = T[vb.z];
        if (! live(sl))
            return;
        Scalars * s = sl.s;
        if (MODE == 0 ? s->open == 0 : s->run == 0)
            return;
        const float * __restrict__ d = sl.v[L][F_D];
        double acc = 0.0;
        typedef k::Vec16<float>::type F4;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        const int64_t tid = (int64_t)vb.x * blockDim.x + threadIdx.x;
        const int64_t n4 = n >> 2;
        const F4 * d4 = reinterpret_cast<const F4 *>(d);
        int64_t i = tid;
        for (; i + 3 * stride < n4; i += 4 * stride)
        {
            F4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                v[u] = __builtin_nontemporal_load(d4 + i + u * stride);
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                acc += (double)v[u].x * (double)v[u].x;
                acc += (double)v[u].y * (double)v[u].y;
                acc += (double)v[u].z * (double)v[u].z;
                acc += (double)v[u].w * (double)v[u].w;
            }
        }
        for (; i < n4; i += stride)
        {
            const F4 v = __builtin_nontemporal_load(d4 + i);
            acc += (double)v.x * (double)v.x;
            acc += (double)v.y * (double)v.y;
            acc += (double)v.z * (double)v.z;
            acc += (double)v.w * (double)v.w;
        }
        for (int64_t t = 4 * n4 + tid; t < n; t += stride)
        {
            const double v = (double)d[t];
            acc += v * v;
        }
        k::put_partials2(acc, 0.0, sl.partials, vb.x);
    }

    // the inner stopping test from inner_norm_kernel's partials (nb blocks
    // per slot); one block per slot (blockIdx.z)
    template <int MODE>
    __global__ __launch_bounds__(256) void inner_norm_finish_kernel(const Slot * T, unsigned nb, float tol,
            int convergence_check, uint64_t n_max_iter)
    {
        const Slot & sl = T[blockIdx.z];
        if (! live(sl))
            return;
        Scalars * s = sl.s;
        if (MODE == 0 ? s->open == 0 : s->run == 0)
            return;
        double ss, unused;
        k::sum_partials2(sl.partials, nb, ss, unused);
        if (threadIdx.x != 0)
            return;
        {
            const float nrm = sqrtf((float)ss);
            if (MODE == 0)
            {
                s->inner_initial = nrm;
                s->inner_defect = nrm;
                if (convergence_check && nrm <= tol)
                {
                    // nothing to do; the count keeps the previous solve's value
                    s->run = 0;
                    s->need_outer = 1;
                }
                else
                    s->inner_last = 0;
            }
            else
            {
                s->inner_defect = nrm;
                s->inner_last += 1;
                if (nrm <= s->inner_initial * tol || s->inner_last >= n_max_iter)
                {
                    s->run = 0;
                    s->need_outer = 1;
                }
            }
        }
    }

    // one-level hierarchy: the cycle loop is left after the coarse solve
    __global__ void one_level_end_kernel(const Slot * T, unsigned nslots)
    {
        const unsigned z = blockIdx.x * blockDim.x + threadIdx.x;
        if (z >= nslots)
            return;
        Scalars * s = T[z].s;
        if (s == nullptr || s->run == 0)
            return;
        s->inner_last += 1;
        s->run = 0;
        s->need_outer = 1;
    }

    // --------------------------------------------------------------- V-cycle
#define HS_GATE_RUN                                     \
    const Slot & sl = T[vb.z];                          \
    if (! live(sl) || sl.s->run == 0)                   \
        return;

    // x (+)= dinv * d: the single first sweep when n_pre_smooth == 1
    template <int ACC>
    __global__ __launch_bounds__(256) void v_zero1_kernel(LevelArgs a, const Slot * T, int L, int fsrc, int fout)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
        if (i >= a.n)
            return;
        const float * d = sl.v[L][fsrc];
        float * out = sl.v[L][fout];
        const float c = a.dinv[i] * d[i];
        out[i] = ACC ? out[i] + c : c;
    }

    // The row kernels below cover 4 rows per lane: on Q1 bands 4 consecutive
    // rows as 16-byte vectors (kernels_la.hh / kernels_mg.hh bodies), on ELL
    // one row per pass of a grid-stride loop (lanes on consecutive rows, so
    // every slot load of a wave is one coalesced access).

    // the first two sweeps from zero (c1 = dinv*d, c2 = c1 + dinv*(d - A c1));
    // ACC: out = (field facc) + c2 (facc may be fout), else out = c2
    template <int ACC, int MK>
    __global__ __launch_bounds__(256) void v_zero2_kernel(LevelArgs a, const Slot * T, int L, int fsrc, int fout,
                                                          int facc)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const int64_t lane = (int64_t)vb.x * blockDim.x + threadIdx.x;
        const float * d = sl.v[L][fsrc];
        float * out = sl.v[L][fout];
        const float * accin = sl.v[L][facc];
        if (kBands<MK>)
        {
            k::jacobi_zero2_body<float, ACC, kOrd<MK>>(a.B, d, a.dinv, out, a.n, a.m, lane, accin);
            return;
        }
        for (int64_t i = lane; i < a.n; i += (int64_t)gridDim.x * blockDim.x)
        {
            const float s = ell_row_g<float>(a.E, i, [&](uint32_t c) { return a.dinv[c] * d[c]; });
            const float di = a.dinv[i];
            const float c1 = di * d[i];
            const float t = d[i] - s;
            const float c2 = c1 + t * di;
            out[i] = ACC ? accin[i] + c2 : c2;
        }
    }

    // workgroups (over all slots) of the per-tick gate kernels' capped grids
    constexpr unsigned kGateBlocks = 4096;
    constexpr int kPostLanes = 256;
    constexpr int kPostW = kPostLanes - 6;
    // slots per workgroup of the one-pass kernels when a tick runs several:
    // they share each tile's band and dinv loads
    constexpr int kPostSlots = 2;
    template <int N> using IntC = std::integral_constant<int, N>;

    // Row i of A x from three rows of x in LDS (up, mid, dn: rows r - 1, r,
    // r + 1 of the tile's column span, j the lane's column), b[0..8] the row's
    // nine bands (LL .. UU), in q1_row's order.  GUARD: test each neighbour's
    // index against [0, n) as q1_row does; without it every neighbour must
    // lie in the matrix (rows 1 .. m - 2 of the grid).  Indices fit in 32 bits
    // (levels up to 16385^2).
    template <int MK, bool GUARD, int NB>
    __device__ __forceinline__ float post_row(const float (&b)[NB], const float * __restrict__ up,
                                              const float * __restrict__ mid, const float * __restrict__ dn, int j,
                                              int i, int n, int m)
    {
        if (kOrd<MK> == 1)
        {
            float s = 0.0f;
            if (! GUARD || i - m - 1 >= 0) s += b[0] * up[j - 1];
            if (! GUARD || i - m >= 0)     s += b[1] * up[j];
            if (! GUARD || i - m + 1 >= 0) s += b[2] * up[j + 1];
            if (! GUARD || i - 1 >= 0)     s += b[3] * mid[j - 1];
            s += b[4] * mid[j];
            if (! GUARD || i + 1 < n)      s += b[5] * mid[j + 1];
            if (! GUARD || i + m - 1 < n)  s += b[6] * dn[j - 1];
            if (! GUARD || i + m < n)      s += b[7] * dn[j];
            if (! GUARD || i + m + 1 < n)  s += b[8] * dn[j + 1];
            return s;
        }
        float s = b[4] * mid[j];
        if (! GUARD || i - m - 1 >= 0) s += b[0] * up[j - 1];
        if (! GUARD || i - m >= 0)     s += b[1] * up[j];
        if (! GUARD || i - m + 1 >= 0) s += b[2] * up[j + 1];
        if (! GUARD || i - 1 >= 0)     s += b[3] * mid[j - 1];
        if (! GUARD || i + 1 < n)      s += b[5] * mid[j + 1];
        if (! GUARD || i + m - 1 < n)  s += b[6] * dn[j - 1];
        if (! GUARD || i + m < n)      s += b[7] * dn[j];
        if (! GUARD || i + m + 1 < n)  s += b[8] * dn[j + 1];
        return s;
    }

    // Tiles of a level of side m with th rows each (per slot).
    inline unsigned post_tiles(int64_t m, int th)
    {
        return (unsigned)(((m + kPostW - 1) / kPostW) * ((m + th - 1) / th));
    }

    // Tile rows of the one-pass kernels: a tile's steps are sequential (th + 6
    // of them, each waiting on memory), so the height is the even th <= 128
    // with the fewest rounds of steps, ceil(tiles / resident) * (th + 6),
    // where `resident` tiles fit on the device at once (the larger th on a
    // tie: fewer halo rows).
    inline int post_rows(int64_t m, unsigned slots, unsigned resident)
    {
        int best = 2;
        uint64_t best_cost = ~(uint64_t)0;
        for (int th = 2; th <= 128; th += 2)
        {
            const uint64_t tiles = (uint64_t)post_tiles(m, th) * slots;
            const uint64_t rounds = (tiles + resident - 1) / resident;
            const uint64_t cost = rounds * (uint64_t)((th + 7) & ~1);
            if (cost <= best_cost)
            {
                best_cost = cost;
                best = th;
            }
            if (th >= m)
                break;
        }
        return best;
    }

    // Workgroups of kernel `fn` (256 lanes) resident on the current device at
    // once: CUs x the occupancy its registers and LDS allow (cached).
    inline unsigned post_resident(const void * fn)
    {
        static std::mutex mu;
        static std::map<std::pair<int, const void *>, unsigned> cache;
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(std::make_pair(dev, fn));
        if (it != cache.end())
            return it->second;
        int cus = 0, per_cu = 0;
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPostLanes, 0));
        const unsigned r = (unsigned)std::max(1, cus * std::max(1, per_cu));
        cache.emplace(std::make_pair(dev, fn), r);
        return r;
    }

    // ------------------------------------------- post-smoothing in one pass
    // Two damped-Jacobi sweeps x1 = x + dinv (rhs - A x), x2 = x1 + dinv (rhs
    // - A x1) and, with DEFECT, d = rhs - A x2, in one pass over Q1 bands
    // (MK 0 / 2).  Every value is the one the separate sweep and defect
    // kernels form: each row sums in q1_row's order over the same linear
    // neighbours i +- 1, i +- m, i +- m +- 1 with the same index guards.
    //
    // A tile is kPostW consecutive columns of th grid rows (the last tiles of
    // a level are narrower / shorter).  Its kPostLanes lanes each own one column
    // offset c = c0 - 3 + lane of every row; a row's linear span [R m + c0 -
    // 3, R m + c0 + W + 3) runs past the grid row's ends into the
    // neighbouring rows, exactly as the linear index does.  The tile marches
    // through its rows, down (even row blocks) or up (odd ones): march
    // position p is grid row R0 + p or R0 + H - 1 - p.  Step s takes x at
    // position s - 3 and the bands, dinv and rhs at s - 4 (loaded two steps
    // ahead, in registers), then forms x1 at position s - 4, x2 at s - 5 and
    // d at s - 6 from three-row rings of x, x1, x2 in LDS.  Each row's bands
    // are read from HBM once per tile and kept in registers for its three
    // stages.  A node's value does not depend on the direction: its sum
    // reads the rows above and below from the ring either way.
    //
    // x2 goes to field fout, which must differ from fx: neighbouring tiles
    // read x around their own rows while this tile writes.
    //
    // The steps synchronise on LDS only (lds_barrier: no wait for global
    // loads), so the loads of the next two steps stay in flight across them;
    // no lane reads a global value another lane of the tile wrote.
    //
    // PROL: x is first corrected by the prolongation of level L-1's field
    // fcoarse, x + alpha P(coarse) (Prolongation<NONE> then ScaledSum with
    // the fixed correction factor), formed per loaded node, so the corrected
    // x is never stored.
    template <int MK, int DEFECT, int PROL, int K, bool CB = false>
    __global__ __launch_bounds__(kPostLanes) void v_post2_kernel(LevelArgs a, const Slot * T, int L, int fx, int fout,
                                                          int fcoarse, float alpha, k::Mask mk, int th, int nslots)
    {
        const VB vb = vblock();
        // the workgroup's K slots vb.z K + q: they share the tile's band and
        // dinv loads; a slot that is not running borrows the arrays of one
        // that is for its loads and stores nothing
        bool on[K];
        int pick = -1;
#pragma unroll
        for (int q = 0; q < K; ++q)
        {
            const int z = (int)vb.z * K + q;
            on[q] = z < nslots && live(T[z]) && T[z].s->run != 0;
            if (on[q] && pick < 0)
                pick = q;
        }
        if (pick < 0)
            return;
        // CB: the bands and dinv come as value codes (one 64-bit word per
        // node) and are decoded through their tables, held in LDS
        __shared__ float ctab[CB ? 160 : 1];
        if constexpr (CB)
        {
            if (threadIdx.x < 160)
                ctab[threadIdx.x] = a.ctab[threadIdx.x];
            __syncthreads();
        }
        const gptr<const uint64_t> codes = as_global<const uint64_t>(a.codes);
        auto band_in = [&](const auto & bq, int k) -> float {
            if constexpr (CB)
                return ctab[k * 16 + (int)((bq >> (4 * k)) & 15u)];
            else
                return bq[k];
        };
        using BandIn = std::conditional_t<CB, uint64_t, float[10]>;
        __shared__ float xr[K][3][kPostLanes];
        __shared__ float x1r[K][3][kPostLanes];
        __shared__ float x2r[K][DEFECT ? 3 : 1][kPostLanes];
        const int n = (int)a.n, m = (int)a.m;
        // tiles in row blocks, column strips fastest: horizontal neighbours
        // march through the same rows at the same time on one XCD
        // (k::xcd_block), so the cache lines they share (a row chunk's first
        // and last, which the grid's odd side leaves unaligned) come from HBM
        // once
        const unsigned t = gridDim.z == 1 ? k::xcd_block(vb.x, gridDim.x) : vb.x;
        const int nst = (m + kPostW - 1) / kPostW;
        const int c0 = (int)(t % nst) * kPostW;
        const int R0 = (int)(t / nst) * th;
        const int W = min(kPostW, m - c0);
        const int H = min(th, m - R0);
        // even row blocks march down, odd ones up, so the halo rows two
        // vertical neighbours share are read by both at about the same time
        // (both at their first steps or both at their last); march position
        // p is grid row Rs + dir p
        const int dir = ((t / nst) & 1) ? -1 : 1;
        const int Rs = dir > 0 ? R0 : R0 + H - 1;
        const int j = threadIdx.x;
        const int c = c0 - 3 + j;
        const bool in_x = j < W + 6;
        const bool in_1 = j >= 1 && j < W + 5;
        const bool in_2 = j >= 2 && j < W + 4;
        const bool own = j >= 3 && j < W + 3;
        gptr<const float> x[K], rhs[K], xc[K];
        gptr<float> out[K], dout[K];
#pragma unroll
        for (int q = 0; q < K; ++q)
        {
            const Slot & sl = T[(int)vb.z * K + (on[q] ? q : pick)];
            x[q] = as_global<const float>(sl.v[L][fx]);
            rhs[q] = as_global<const float>(sl.v[L][F_RHS]);
            out[q] = as_global<float>(sl.v[L][fout]);
            dout[q] = as_global<float>(sl.v[L][F_D]);
            xc[q] = as_global<const float>(PROL ? sl.v[L - 1][fcoarse] : sl.v[L][fx]);
        }
        const gptr<const float> dinv = as_global<const float>(a.dinv);
        // the nine bands from one base: band k of node i at element i + k bs
        // (9 bs 4 < 2^32, checked on the host)
        const gptr<const float> bands = as_global<const float>((const float *)a.B.b[0]);
        const uint32_t bs = (uint32_t)a.bs;
        const int Nc = (m + 1) / 2;

        // grid position (R, S) of the node at march position s - 3, column
        // offset c (c lies in [-3, 2m)), clamped to the grid
        auto grid_pos = [&](int s, int & R, int & S) {
            R = Rs + dir * (s - 3);
            S = c;
            if (S < 0)
            {
                S += m;
                R -= 1;
            }
            else if (S >= m)
            {
                S -= m;
                R += 1;
            }
            R = min(max(R, 0), m - 1);
            S = min(max(S, 0), m - 1);
        };
        // The loads of step s: each slot's x at march position s - 3 (with
        // PROL the four coarse values around it) and rhs at position s - 4,
        // and the bands and dinv at s - 4 (shared).  Every lane loads, at
        // indices clamped to the arrays, with no branch between the loads;
        // step s decides which values count.
        auto load = [&](int s, float (&xq)[K], float (&cq)[K][4], float (&rq)[K], BandIn & bq) {
            const int ix = (Rs + dir * (s - 3)) * m + c;
            const uint32_t ib = (uint32_t)min(max(ix - dir * m, 0), n - 1);
#pragma unroll
            for (int q = 0; q < K; ++q)
            {
                xq[q] = ld32(x[q], (uint32_t)min(max(ix, 0), n - 1));
                if (PROL)
                {
                    int R, S;
                    grid_pos(s, R, S);
                    const int ra = (R >> 1) * Nc, rb = ((R + 1) >> 1) * Nc;
                    const int jj = S >> 1, jj1 = min(jj + 1, Nc - 1);
                    cq[q][0] = ld32(xc[q], (uint32_t)(ra + jj));
                    cq[q][1] = ld32(xc[q], (uint32_t)(ra + jj1));
                    cq[q][2] = ld32(xc[q], (uint32_t)(rb + jj));
                    cq[q][3] = ld32(xc[q], (uint32_t)(rb + jj1));
                }
                rq[q] = ld32(rhs[q], ib);
            }
            if constexpr (CB)
                bq = ld32(codes, ib);
            else
            {
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    bq[k] = ld32(bands, ib + (uint32_t)k * bs);
                bq[9] = ld32(dinv, ib);
            }
        };
        // H + 6 steps, rounded up to an even count so that the loop body is
        // two whole steps with no branch around their loads (the extra step's
        // rows lie outside the tile: nothing of it is stored)
        const int steps = (H + 7) & ~1;
        // A tile whose rows R0 - 4 .. R0 + H + 3 lie in the grid has every
        // node the stored values depend on, and every neighbour of those, in
        // the matrix: no index tests
        const bool interior = R0 >= 4 && R0 + H + 4 <= m;
        auto march = [&](auto interior_tag) {
            constexpr bool G = ! decltype(interior_tag)::value;
            // stage registers: march position s - 4 (stage 1), s - 5 (stage
            // 2), s - 6 (stage 3): bands and dinv (b), each slot's rhs (r)
            float b1[10] = {}, b2[10] = {}, b3[10] = {};
            float r1[K] = {}, r2[K] = {}, r3[K] = {};
            // one step; its loads, then those of step s + 2
            auto step = [&](int s, float (&xq)[K], float (&cq)[K][4], float (&rq)[K], BandIn & bq) {
                // ring rows: positions p, p - 1, p - 2; the row above a stage's
                // row is the older one marching down, the newer one marching up
                const int cur = s % 3, prv = (s + 2) % 3, old = (s + 1) % 3;
                const int above = dir > 0 ? old : cur, below = dir > 0 ? cur : old;
                {
                    const int ix = (Rs + dir * (s - 3)) * m + c;
                    const bool vx = in_x && (! G || (ix >= 0 && ix < n));
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
                        float xv = 0.0f;
                        if (vx)
                        {
                            xv = xq[q];
                            if (PROL)
                            {
                                int R, S;
                                grid_pos(s, R, S);
                                xv = xv + alpha * k::prolongation_node_v<float>(cq[q], m, mk, R, S);
                            }
                        }
                        xr[q][cur][j] = xv;
                    }
                    const int ib = ix - dir * m;
                    const bool vb1 = in_1 && s >= 2 && (! G || (ib >= 0 && ib < n));
#pragma unroll
                    for (int k = 0; k < 10; ++k)
                    {
                        b3[k] = b2[k];
                        b2[k] = b1[k];
                        b1[k] = vb1 ? band_in(bq, k) : 0.0f;
                    }
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
                        r3[q] = r2[q];
                        r2[q] = r1[q];
                        r1[q] = vb1 ? rq[q] : 0.0f;
                    }
                }
                load(s + 2, xq, cq, rq, bq);
                lds_barrier();
                // stage 1: x1 at position s - 4
                if (s >= 2)
                {
                    const int i = (Rs + dir * (s - 4)) * m + c;
                    const bool v1 = in_1 && (! G || (i >= 0 && i < n));
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
                        float v = 0.0f;
                        if (v1)
                        {
                            const float sum = post_row<MK, G>(b1, xr[q][above], xr[q][prv], xr[q][below], j, i, n, m);
                            const float tt = r1[q] - sum;
                            v = xr[q][prv][j] + tt * b1[9];
                        }
                        x1r[q][cur][j] = v;
                    }
                }
                lds_barrier();
                // stage 2: x2 at position s - 5
                if (s >= 4)
                {
                    const int r2w = Rs + dir * (s - 5);
                    const int i = r2w * m + c;
                    const bool v2 = in_2 && (! G || (i >= 0 && i < n));
                    const bool st2 = own && r2w >= R0 && r2w < R0 + H;
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
                        float v = 0.0f;
                        if (v2)
                        {
                            const float sum = post_row<MK, G>(b2, x1r[q][above], x1r[q][prv], x1r[q][below], j, i, n, m);
                            const float tt = r2[q] - sum;
                            v = x1r[q][prv][j] + tt * b2[9];
                            if (st2 && on[q])
                                st32(out[q], (uint32_t)i, v);
                        }
                        if (DEFECT)
                            x2r[q][DEFECT ? cur : 0][j] = v;
                    }
                }
                if (DEFECT)
                {
                    lds_barrier();
                    // stage 3: d at position s - 6
                    const int r3w = Rs + dir * (s - 6);
                    const int i = r3w * m + c;
                    if (s >= 6 && own && r3w >= R0 && r3w < R0 + H && (! G || (i >= 0 && i < n)))
                    {
#pragma unroll
                        for (int q = 0; q < K; ++q)
                        {
                            const float sum = post_row<MK, G>(b3, x2r[q][DEFECT ? above : 0], x2r[q][DEFECT ? prv : 0],
                                                              x2r[q][DEFECT ? below : 0], j, i, n, m);
                            if (on[q])
                                st32(dout[q], (uint32_t)i, r3[q] - sum);
                        }
                    }
                }
            };
            float xA[K], xB[K], cA[K][4] = {}, cB[K][4] = {}, rA[K], rB[K];
            BandIn bA, bB;
            load(0, xA, cA, rA, bA);
            load(1, xB, cB, rB, bB);
            for (int s = 0; s < steps; s += 2)
            {
                step(s, xA, cA, rA, bA);
                step(s + 1, xB, cB, rB, bB);
            }
        };
        if (interior)
            march(std::true_type{});
        else
            march(std::false_type{});
    }

    // ---------------------- pre-smoothing, defect and restriction in one pass
    // The first two damped-Jacobi sweeps from zero (c1 = dinv d, c2 = c1 +
    // dinv (d - A c1)), x = c2 (ACC: x = (field facc) + c2), the defect d' =
    // rhs - A x and the coarse right-hand side Restriction<NONE>(d') with the
    // Dirichlet mask, in one pass over Q1 bands (MK 0 / 2).  Every value is
    // the one the separate kernels form (q1_row's order and guards,
    // restrict_node's addition sequence); d' itself is never stored.
    //
    // Tiles and march directions as v_post2_kernel, with an even th, so a
    // tile's first row is even and it owns the coarse rows r with 2r in its
    // rows.  Step s takes march position s - 3 (d, dinv, rhs, the bands and
    // facc, loaded two steps ahead; c1 is formed at once), forms x at s - 4
    // and d' at s - 5, and, when that d' row is 2r + 1 (marching down) or 2r
    // - 1 (up), coarse row r from the three d' rows in LDS (lanes on even
    // fine columns).  The lanes cover fine columns c0 - 3
    // .. c0 + W + 2 in linear index order, so the stencil steps are the
    // linear index's; the restriction reads grid columns 2s - 1 .. 2s + 1 of
    // owned columns 2s, which all lie in that span.
    template <int MK, int ACC, int K, bool CB = false>
    __global__ __launch_bounds__(kPostLanes) void v_pre2r_kernel(LevelArgs a, const Slot * T, int L, int fd, int facc,
                                                          int64_t Nc64, k::Mask mk, int th, int nslots)
    {
        const VB vb = vblock();
        // the workgroup's K slots, as in v_post2_kernel
        bool on[K];
        int pick = -1;
#pragma unroll
        for (int q = 0; q < K; ++q)
        {
            const int z = (int)vb.z * K + q;
            on[q] = z < nslots && live(T[z]) && T[z].s->run != 0;
            if (on[q] && pick < 0)
                pick = q;
        }
        if (pick < 0)
            return;
        // CB: value-coded bands and dinv, as in v_post2_kernel
        __shared__ float ctab[CB ? 160 : 1];
        if constexpr (CB)
        {
            if (threadIdx.x < 160)
                ctab[threadIdx.x] = a.ctab[threadIdx.x];
            __syncthreads();
        }
        const gptr<const uint64_t> codes = as_global<const uint64_t>(a.codes);
        auto band_in = [&](const auto & bq, int k) -> float {
            if constexpr (CB)
                return ctab[k * 16 + (int)((bq >> (4 * k)) & 15u)];
            else
                return bq[k];
        };
        using BandIn = std::conditional_t<CB, uint64_t, float[10]>;
        __shared__ float c1r[K][3][kPostLanes];
        __shared__ float xr[K][3][kPostLanes];
        __shared__ float dr[K][3][kPostLanes];
        const int n = (int)a.n, m = (int)a.m, Nc = (int)Nc64;
        const unsigned t = gridDim.z == 1 ? k::xcd_block(vb.x, gridDim.x) : vb.x;
        const int nst = (m + kPostW - 1) / kPostW;
        const int c0 = (int)(t % nst) * kPostW;
        const int R0 = (int)(t / nst) * th;
        const int W = min(kPostW, m - c0);
        const int H = min(th, m - R0);
        // march direction as in v_post2_kernel: position p is row Rs + dir p
        const int dir = ((t / nst) & 1) ? -1 : 1;
        const int Rs = dir > 0 ? R0 : R0 + H - 1;
        const int j = threadIdx.x;
        const int c = c0 - 3 + j;
        const bool in_x = j < W + 6;
        const bool in_1 = j >= 1 && j < W + 5;
        const bool in_2 = j >= 2 && j < W + 4;
        const bool own = j >= 3 && j < W + 3;
        const bool own_even = own && ((c & 1) == 0);
        gptr<const float> d[K], rhs[K], acc[K];
        gptr<float> xo[K], co[K];
#pragma unroll
        for (int q = 0; q < K; ++q)
        {
            const Slot & sl = T[(int)vb.z * K + (on[q] ? q : pick)];
            d[q] = as_global<const float>(sl.v[L][fd]);
            rhs[q] = as_global<const float>(sl.v[L][F_RHS]);
            acc[q] = as_global<const float>(sl.v[L][facc]);
            xo[q] = as_global<float>(sl.v[L][F_X]);
            co[q] = as_global<float>(sl.v[L - 1][F_RHS]);
        }
        const gptr<const float> dinv = as_global<const float>(a.dinv);
        // bands as in v_post2_kernel
        const gptr<const float> bands = as_global<const float>((const float *)a.B.b[0]);
        const uint32_t bs = (uint32_t)a.bs;

        // march position s - 3 for step s: the 9 bands and dinv (shared) and
        // each slot's d, rhs and facc value (pq[q][0..2]); every lane loads,
        // at an index clamped to the arrays, with no branch between the
        // loads; step s decides which values count
        auto load = [&](int s, BandIn & bq, float (&pq)[K][3]) {
            const uint32_t i = (uint32_t)min(max((Rs + dir * (s - 3)) * m + c, 0), n - 1);
            if constexpr (CB)
                bq = ld32(codes, i);
            else
            {
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    bq[k] = ld32(bands, i + (uint32_t)k * bs);
                bq[9] = ld32(dinv, i);
            }
#pragma unroll
            for (int q = 0; q < K; ++q)
            {
                pq[q][0] = ld32(d[q], i);
                pq[q][1] = ld32(rhs[q], i);
                pq[q][2] = ACC ? ld32(acc[q], i) : 0.0f;
            }
        };
        // H + 6 steps, rounded up to an even count (as v_post2_kernel)
        const int steps = (H + 7) & ~1;
        // rows R0 - 4 .. R0 + H + 3 in the grid: no index tests (as v_post2_kernel)
        const bool interior = R0 >= 4 && R0 + H + 4 <= m;
        auto march = [&](auto interior_tag) {
            constexpr bool G = ! decltype(interior_tag)::value;
            // stage registers: position s - 3 (its c1 only, yet), s - 4 (x:
            // bands, dinv, d, facc), s - 5 (d': bands, rhs); bands and dinv
            // (bc, ba, bb) shared, d, rhs, facc per slot (pc, pa, pb)
            float bc[10] = {}, ba[10] = {}, bb[10] = {};
            float pc[K][3] = {}, pa[K][3] = {}, pb[K][3] = {};
            auto step = [&](int s, BandIn & bq, float (&pq)[K][3]) {
                // ring rows: positions p, p - 1, p - 2 (above / below as in v_post2_kernel)
                const int cur = s % 3, prv = (s + 2) % 3, old = (s + 1) % 3;
                const int above = dir > 0 ? old : cur, below = dir > 0 ? cur : old;
                {
                    const int i = (Rs + dir * (s - 3)) * m + c;
                    const bool valid = in_x && (! G || (i >= 0 && i < n));
#pragma unroll
                    for (int k = 0; k < 10; ++k)
                    {
                        bb[k] = ba[k];
                        ba[k] = bc[k];
                        bc[k] = valid ? band_in(bq, k) : 0.0f;
                    }
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                        {
                            pb[q][k] = pa[q][k];
                            pa[q][k] = pc[q][k];
                            pc[q][k] = valid ? pq[q][k] : 0.0f;
                        }
                        c1r[q][cur][j] = bc[9] * pc[q][0];
                    }
                }
                load(s + 2, bq, pq);
                lds_barrier();
                // x at position s - 4
                if (s >= 2)
                {
                    const int rho = Rs + dir * (s - 4);
                    const int i = rho * m + c;
                    const bool v1 = in_1 && (! G || (i >= 0 && i < n));
                    const bool stx = own && rho >= R0 && rho < R0 + H;
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
                        float v = 0.0f;
                        if (v1)
                        {
                            const float sum = post_row<MK, G>(ba, c1r[q][above], c1r[q][prv], c1r[q][below], j, i, n, m);
                            const float tt = pa[q][0] - sum;
                            const float c2 = c1r[q][prv][j] + tt * ba[9];
                            v = ACC ? pa[q][2] + c2 : c2;
                            if (stx && on[q])
                                st32(xo[q], (uint32_t)i, v);
                        }
                        xr[q][cur][j] = v;
                    }
                }
                lds_barrier();
                // d' at position s - 5
                if (s >= 4)
                {
                    const int rho = Rs + dir * (s - 5);
                    const int i = rho * m + c;
                    const bool v2 = in_2 && (! G || (i >= 0 && i < n));
#pragma unroll
                    for (int q = 0; q < K; ++q)
                    {
                        float v = 0.0f;
                        if (v2)
                        {
                            const float sum = post_row<MK, G>(bb, xr[q][above], xr[q][prv], xr[q][below], j, i, n, m);
                            v = pb[q][1] - sum;
                        }
                        dr[q][cur][j] = v;
                    }
                    lds_barrier();
                    // coarse row r when this d' row is 2r + dir: rows 2r - 1,
                    // 2r, 2r + 1 are then in the ring
                    if (((rho - 1) & 1) == 0)
                    {
                        const int r = (rho - dir) >> 1;
                        if (own_even && 2 * r >= R0 && 2 * r < R0 + H && c < m)
                        {
                            const int sc = c >> 1;
#pragma unroll
                            for (int q = 0; q < K; ++q)
                            {
                                const float (*ring)[kPostLanes] = dr[q];
                                const float v = k::restrict_node<float>(
                                    [&](int ddr, int dds) {
                                        return ring[dir * ddr < 0 ? old : (ddr == 0 ? prv : cur)][j + dds];
                                    },
                                    r, sc, Nc, mk);
                                if (on[q])
                                    st32(co[q], (uint32_t)(r * Nc + sc), v);
                            }
                        }
                    }
                }
            };
            BandIn bA, bB;
            float pA[K][3], pB[K][3];
            load(0, bA, pA);
            load(1, bB, pB);
            for (int s = 0; s < steps; s += 2)
            {
                step(s, bA, pA);
                step(s + 1, bB, pB);
            }
        };
        if (interior)
            march(std::true_type{});
        else
            march(std::false_type{});
    }

    // y = x + dinv * (rhs - A x)
    template <int MK>
    __global__ __launch_bounds__(256) void v_sweep_kernel(LevelArgs a, const Slot * T, int L, int fx, int fy, int frhs)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const int64_t lane = (int64_t)vb.x * blockDim.x + threadIdx.x;
        const float * x = sl.v[L][fx];
        float * y = sl.v[L][fy];
        const float * rhs = sl.v[L][frhs];
        if (kBands<MK>)
        {
            k::q1_body_v<float, 2, kOrd<MK>>(a.B, x, y, rhs, a.dinv, a.n, a.m, lane);
            return;
        }
        for (int64_t i = lane; i < a.n; i += (int64_t)gridDim.x * blockDim.x)
        {
            const float t = rhs[i] - ell_row<float>(a.E, x, i);
            y[i] = x[i] + t * a.dinv[i];
        }
    }

    // d = rhs - A x
    template <int MK>
    __global__ __launch_bounds__(256) void v_defect_kernel(LevelArgs a, const Slot * T, int L)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const int64_t lane = (int64_t)vb.x * blockDim.x + threadIdx.x;
        const float * x = sl.v[L][F_X];
        float * d = sl.v[L][F_D];
        const float * rhs = sl.v[L][F_RHS];
        if (kBands<MK>)
        {
            k::q1_body_v<float, 1, kOrd<MK>>(a.B, x, d, rhs, nullptr, a.n, a.m, lane);
            return;
        }
        for (int64_t i = lane; i < a.n; i += (int64_t)gridDim.x * blockDim.x)
            d[i] = rhs[i] - ell_row<float>(a.E, x, i);
    }

    // Product<PROLMAT>: level Ldst's fdst = E * (level Lsrc's fsrc)
    __global__ __launch_bounds__(256) void v_apply_ell_kernel(EllMat32 E, const Slot * T, int Lsrc, int fsrc, int Ldst,
                                                              int fdst)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const float * x = sl.v[Lsrc][fsrc];
        float * y = sl.v[Ldst][fdst];
        const int64_t lane = (int64_t)vb.x * blockDim.x + threadIdx.x;
        for (int64_t i = lane; i < (int64_t)E.rows; i += (int64_t)gridDim.x * blockDim.x)
            y[i] = ell_row<float>(E, x, i);
    }

    // x += alpha * c (ScaledSum with the constant correction factor)
    __global__ __launch_bounds__(256) void v_axpy_const_kernel(const Slot * T, int L, int64_t n, float alpha)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        float * x = sl.v[L][F_X];
        x[i] = x[i] + alpha * sl.v[L][F_C][i];
    }

    // dst = src (ADD 0) or dst += src (ADD 1), four entries per lane
    template <int ADD>
    __global__ __launch_bounds__(256) void v_move_kernel(const Slot * T, int L, int64_t n, int fsrc, int fdst)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        typedef k::Vec16<float>::type F4;
        const float * src = sl.v[L][fsrc];
        float * dst = sl.v[L][fdst];
        const int64_t i0 = ((int64_t)vb.x * blockDim.x + threadIdx.x) * 4;
        if (i0 + 4 <= n)
        {
            const F4 s = *(const F4 *)(src + i0);
            if (ADD)
            {
                F4 d = *(const F4 *)(dst + i0);
                d.x = d.x + s.x;
                d.y = d.y + s.y;
                d.z = d.z + s.z;
                d.w = d.w + s.w;
                *(F4 *)(dst + i0) = d;
            }
            else
                *(F4 *)(dst + i0) = s;
        }
        else
            for (int64_t i = i0; i < n; ++i)
                dst[i] = ADD ? dst[i] + src[i] : src[i];
    }

    // rhs of level L-1 = restriction of d of level L
    __global__ __launch_bounds__(256) void v_restrict_kernel(const Slot * T, int L, int64_t Nc, k::Mask mk)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        k::restriction_body<float>(sl.v[L - 1][F_RHS], sl.v[L][F_D], Nc, mk, vb.y,
                                   (int64_t)vb.x * blockDim.x + threadIdx.x);
    }

    // ConjugateGradients<NONE> (konv_rad form) on the coarsest level
    template <int MK>
    __global__ __launch_bounds__(256) void v_coarse_kernel(LevelArgs a, const Slot * T, int L, int frhs,
                                                           double konv_rad, uint64_t max_iters)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        LevelRowOp<MK> A{a};
        k::cg_konv_body_op<float>(A, sl.v[L][frhs], sl.v[L][F_X], sl.cg_work, a.n, konv_rad, max_iters,
                                  &sl.s->cg_iters);
    }

    // v_coarse_kernel with the vectors and (Q1 bands) the level's nine bands
    // copied to LDS, for coarsest levels up to 17^2: the CG's sequential
    // sums then wait on LDS, not on a memory round trip per step.  It runs
    // k::cg_konv_body_op itself on the copies, so every operation and its
    // order are unchanged and so are the bits.  One wave per slot.
    constexpr int kCoarseLdsMax = 289;
    template <int MK>
    __global__ __launch_bounds__(64) void v_coarse_lds_kernel(LevelArgs a, const Slot * T, int L, int frhs,
                                                              double konv_rad, uint64_t max_iters)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        __shared__ float vec[6 * kCoarseLdsMax];                  // rhs, x, then the 4n of work
        __shared__ float bnd[kBands<MK> ? 9 * kCoarseLdsMax : 1];
        const int n = (int)a.n;
        float * rhs = vec;
        float * x = vec + kCoarseLdsMax;
        float * work = vec + 2 * kCoarseLdsMax;
        const float * rg = sl.v[L][frhs];
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            rhs[i] = rg[i];
        if constexpr (kBands<MK>)
        {
            for (int k = 0; k < 9; ++k)
            {
                const float * b = (const float *)a.B.b[k];
                for (int i = threadIdx.x; i < n; i += blockDim.x)
                    bnd[k * kCoarseLdsMax + i] = b[i];
                a.B.b[k] = bnd + k * kCoarseLdsMax;
            }
        }
        __syncthreads();
        LevelRowOp<MK> A{a};
        k::cg_konv_body_op<float>(A, rhs, x, work, a.n, konv_rad, max_iters, &sl.s->cg_iters);
        __syncthreads();
        float * xg = sl.v[L][F_X];
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            xg[i] = x[i];
    }

    // ------------------------------------------------ the bottom levels in LDS
    // Levels minl .. top (top <= 5, below the finest) of a V-cycle on Q1
    // bands, in one workgroup per slot with every vector and band of those
    // levels in LDS: the same operations, per node, as the one-pass kernels
    // and the coarse solve run there (two sweeps from zero, defect,
    // restriction down to minl, the CG on minl, then prolongation with the
    // correction factor and two sweeps back up to top), one level pass per
    // barrier.  It reads the rhs of level top (written by the finest-but-
    // one-pass kernel above it) and leaves top's smoothed iterate in its
    // F_TMP, where the one-pass post-smoother above reads it.  It replaces
    // two launches per level and the coarse solve's, each a chain of
    // dependent memory round trips.  The coarse CG runs on the first wave
    // alone, whose sequential sums read their terms from its lanes.
    constexpr int kBottomMax = 5;
    struct BottomArgs
    {
        LevelArgs lv[kBottomMax + 1];
        int off[kBottomMax + 1];   // float offset of each level's block in the dynamic LDS
        int minl, top;
        float alpha;
        k::Mask mk;
        double konv_rad;
        uint64_t max_iters;
    };

    // Row i of A x on a level held in LDS (band k at b + k n): q1_row's
    // order and index guards
    template <int ORD>
    __device__ __forceinline__ float lds_row(const float * b, const float * x, int i, int n, int m)
    {
        if (ORD == 1)
        {
            float s = 0.0f;
            if (i - m - 1 >= 0) s += b[i] * x[i - m - 1];
            if (i - m >= 0)     s += b[n + i] * x[i - m];
            if (i - m + 1 >= 0) s += b[2 * n + i] * x[i - m + 1];
            if (i - 1 >= 0)     s += b[3 * n + i] * x[i - 1];
            s += b[4 * n + i] * x[i];
            if (i + 1 < n)      s += b[5 * n + i] * x[i + 1];
            if (i + m - 1 < n)  s += b[6 * n + i] * x[i + m - 1];
            if (i + m < n)      s += b[7 * n + i] * x[i + m];
            if (i + m + 1 < n)  s += b[8 * n + i] * x[i + m + 1];
            return s;
        }
        float s = b[4 * n + i] * x[i];
        if (i - m - 1 >= 0) s += b[i] * x[i - m - 1];
        if (i - m >= 0)     s += b[n + i] * x[i - m];
        if (i - m + 1 >= 0) s += b[2 * n + i] * x[i - m + 1];
        if (i - 1 >= 0)     s += b[3 * n + i] * x[i - 1];
        if (i + 1 < n)      s += b[5 * n + i] * x[i + 1];
        if (i + m - 1 < n)  s += b[6 * n + i] * x[i + m - 1];
        if (i + m < n)      s += b[7 * n + i] * x[i + m];
        if (i + m + 1 < n)  s += b[8 * n + i] * x[i + m + 1];
        return s;
    }

    // A level held in LDS as the operator of the one-workgroup CG
    template <int ORD> struct LdsRowOp
    {
        const float * b;
        int n, m;
        __device__ __forceinline__ float row(const float * __restrict__ x, int64_t i) const
        {
            return lds_row<ORD>(b, x, (int)i, n, m);
        }
    };

    // Block of level L in LDS: bands (9n), dinv, rhs, x, then tmp (levels
    // above minl) or the CG's 4n of work (minl).
    inline int bottom_floats(int n, bool coarsest) { return (coarsest ? 16 : 13) * n; }

    template <int MK>
    __global__ __launch_bounds__(1024) void v_bottom_kernel(BottomArgs ba, const Slot * T)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        extern __shared__ float lds[];
        const int tid = threadIdx.x, nt = blockDim.x;
        // bands and dinv of every level: each lane issues the loads of two
        // rows of all ten arrays before it stores any, so a level costs one
        // or two memory round trips, not one per array and row
        for (int L = ba.minl; L <= ba.top; ++L)
        {
            const int n = (int)ba.lv[L].n;
            float * b = lds + ba.off[L];
            const float * src[10];
            for (int k = 0; k < 9; ++k)
                src[k] = (const float *)ba.lv[L].B.b[k];
            src[9] = ba.lv[L].dinv;
            for (int i0 = tid; i0 < n; i0 += 2 * nt)
            {
                float v[2][10];
#pragma unroll
                for (int h = 0; h < 2; ++h)
                {
                    const int i = min(i0 + h * nt, n - 1);
#pragma unroll
                    for (int k = 0; k < 10; ++k)
                        v[h][k] = src[k][i];
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
                {
                    const int i = i0 + h * nt;
                    if (i < n)
                    {
#pragma unroll
                        for (int k = 0; k < 10; ++k)
                            b[k * n + i] = v[h][k];
                    }
                }
            }
        }
        {
            const int n = (int)ba.lv[ba.top].n;
            const float * g = sl.v[ba.top][F_RHS];
            float * rhs = lds + ba.off[ba.top] + 10 * n;
            for (int i = tid; i < n; i += nt)
                rhs[i] = g[i];
        }
        __syncthreads();
        // down: x = c1 + dinv (rhs - A c1) with c1 = dinv rhs, d' = rhs - A x,
        // rhs of L - 1 = R d'
        for (int L = ba.top; L > ba.minl; --L)
        {
            const int n = (int)ba.lv[L].n, m = (int)ba.lv[L].m;
            const float * b = lds + ba.off[L];
            const float * dinv = b + 9 * n;
            const float * rhs = b + 10 * n;
            float * x = lds + ba.off[L] + 11 * n;
            float * tmp = lds + ba.off[L] + 12 * n;
            for (int i = tid; i < n; i += nt)
                tmp[i] = dinv[i] * rhs[i];
            __syncthreads();
            for (int i = tid; i < n; i += nt)
            {
                const float t = rhs[i] - lds_row<kOrd<MK>>(b, tmp, i, n, m);
                x[i] = tmp[i] + t * dinv[i];
            }
            __syncthreads();
            for (int i = tid; i < n; i += nt)
                tmp[i] = rhs[i] - lds_row<kOrd<MK>>(b, x, i, n, m);
            __syncthreads();
            const int Nc = (int)ba.lv[L - 1].m;
            float * crhs = lds + ba.off[L - 1] + 10 * Nc * Nc;
            for (int q = tid; q < Nc * Nc; q += nt)
            {
                const int r = q / Nc, sc = q - r * Nc;
                crhs[q] = k::restrict_node<float>(
                    [&](int dr, int ds) {
                        const int idx = (2 * r + dr) * m + 2 * sc + ds;
                        return tmp[min(max(idx, 0), n - 1)];
                    },
                    r, sc, Nc, ba.mk);
            }
            __syncthreads();
        }
        // the coarse solve, as v_coarse_lds_kernel, on the first wave
        {
            const int L = ba.minl;
            const int n = (int)ba.lv[L].n;
            float * b = lds + ba.off[L];
            const LdsRowOp<kOrd<MK>> A{b, n, (int)ba.lv[L].m};
            if (threadIdx.x < 64)
                k::cg_konv_body_op<float, LdsRowOp<kOrd<MK>>, 64>(A, b + 10 * n, b + 11 * n, b + 12 * n, (int64_t)n,
                                                                 ba.konv_rad, ba.max_iters, &sl.s->cg_iters);
            __syncthreads();
        }
        // up: x_in = x + alpha P (coarse iterate), then two sweeps into tmp
        for (int L = ba.minl + 1; L <= ba.top; ++L)
        {
            const int n = (int)ba.lv[L].n, m = (int)ba.lv[L].m;
            const int Nc = (int)ba.lv[L - 1].m;
            const float * cf = lds + ba.off[L - 1] + (L - 1 == ba.minl ? 11 : 12) * Nc * Nc;
            const float * b = lds + ba.off[L];
            const float * dinv = b + 9 * n;
            const float * rhs = b + 10 * n;
            float * x = lds + ba.off[L] + 11 * n;
            float * tmp = lds + ba.off[L] + 12 * n;
            for (int i = tid; i < n; i += nt)
            {
                const int R = i / m, S = i - R * m;
                tmp[i] = x[i] + ba.alpha * k::prolongation_node<float>(cf, m, ba.mk, R, S);
            }
            __syncthreads();
            for (int i = tid; i < n; i += nt)
            {
                const float t = rhs[i] - lds_row<kOrd<MK>>(b, tmp, i, n, m);
                x[i] = tmp[i] + t * dinv[i];
            }
            __syncthreads();
            for (int i = tid; i < n; i += nt)
            {
                const float t = rhs[i] - lds_row<kOrd<MK>>(b, x, i, n, m);
                tmp[i] = x[i] + t * dinv[i];
            }
            __syncthreads();
        }
        {
            const int n = (int)ba.lv[ba.top].n;
            const float * tmp = lds + ba.off[ba.top] + 12 * n;
            float * g = sl.v[ba.top][F_TMP];
            for (int i = tid; i < n; i += nt)
                g[i] = tmp[i];
        }
    }

    // level L's fout (+)= alpha * P (x of level L-1)
    template <bool AXPY>
    __global__ __launch_bounds__(256) void v_prolong_kernel(const Slot * T, int L, int Nf, float alpha, k::Mask mk,
                                                            int fout, int fcoarse)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        k::prolongation2_body<float, AXPY>(sl.v[L][fout], sl.v[L - 1][fcoarse], Nf, alpha, mk, vb.y,
                                           vb.x * blockDim.x + threadIdx.x);
    }

    // adapt_correction_factor == 0: alpha = (d.c) / ((A c).c)
    template <int MK>
    __global__ __launch_bounds__(256) void v_adapt_alpha_kernel(LevelArgs a, const Slot * T, int L)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const float * d = sl.v[L][F_D];
        const float * c = sl.v[L][F_C];
        double a1 = 0.0, a2 = 0.0;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x; i < a.n; i += stride)
        {
            const float t = level_row<MK>(a, c, i);
            a1 += (double)d[i] * (double)c[i];
            a2 += (double)t * (double)c[i];
        }
        k::put_partials2(a1, a2, sl.partials, vb.x);
    }

    // alpha from v_adapt_alpha_kernel's partials; one block per slot
    __global__ __launch_bounds__(256) void v_adapt_alpha_finish_kernel(const Slot * T, unsigned nb)
    {
        const Slot & sl = T[blockIdx.z];
        if (! live(sl) || sl.s->run == 0)
            return;
        double d1, d2;
        k::sum_partials2(sl.partials, nb, d1, d2);
        if (threadIdx.x == 0)
            sl.s->alpha = (float)d1 / (float)d2;
    }

    __global__ __launch_bounds__(256) void v_axpy_alpha_kernel(const Slot * T, int L, int64_t n)
    {
        const VB vb = vblock();
        HS_GATE_RUN
        const int64_t i = (int64_t)vb.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const float alpha = sl.s->alpha;
        float * x = sl.v[L][F_X];
        const float * c = sl.v[L][F_C];
        x[i] = x[i] + alpha * c[i];
    }
#undef HS_GATE_RUN

    __global__ void report_kernel(const Slot * T, unsigned nslots, Report * out)
    {
        const unsigned z = blockIdx.x * blockDim.x + threadIdx.x;
        if (z >= nslots)
            return;
        const Scalars * s = T[z].s;
        Report r;
        if (s == nullptr)
        {
            r.phase = PH_IDLE;
            r.status = 0;
            r.outer_iters = r.inner_total = r.cg_iters = r.ticks = 0;
            r.initial_defect = r.def_norm = 0.0;
        }
        else
        {
            r.phase = s->phase;
            r.status = s->status;
            r.outer_iters = s->outer_iters;
            r.inner_total = s->inner_total;
            r.cg_iters = s->cg_iters;
            r.ticks = s->ticks;
            r.initial_defect = s->initial_defect;
            r.def_norm = s->def_norm;
        }
        out[z] = r;
    }

    // ----------------------------------------------------------------- setup
    // b_i = -1 + 2 u_i, u_i = (bits_i >> 11) 2^-53 with SplitMix64 counters;
    // 0 on the boundary nodes of the square grid
    __global__ __launch_bounds__(256) void rhs_uniform_kernel(double * b, int64_t n, int64_t sq, uint64_t key)
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const int64_t r = i / sq, c = i - r * sq;
        if (r == 0 || r == sq - 1 || c == 0 || c == sq - 1)
        {
            b[i] = 0.0;
            return;
        }
        const uint64_t bits = mix64(key + 0x9E3779B97F4A7C15ull * ((uint64_t)i + 1));
        const double u = (double)(bits >> 11) * 0x1.0p-53;
        b[i] = -1.0 + 2.0 * u;
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

    inline uint64_t align256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }
    inline unsigned lanes4(uint64_t n) { return grid_for((n + 3) / 4, kBlock); }
}  // namespace

void * Hierarchy::alloc(size_t b)
{
    void * ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, b ? b : 1);
    if (e != hipSuccess)
        throw Error(HSOLVE_ERR_ALLOC, "MG hierarchy allocation of " + std::to_string(b) + " bytes failed");
    allocations.push_back(ptr);
    bytes += b;
    return ptr;
}

Hierarchy::~Hierarchy()
{
    for (void * a : allocations)
        (void)hipFree(a);
}

Params make_params(const hsolve_mg_params & p)
{
    Params q;
    q.min_level = p.min_level;
    q.n_max_iter = p.n_max_iter;
    q.initial_zero = p.initial_zero;
    q.convergence_check = p.convergence_check;
    q.tolerance = (float)p.tolerance;
    q.n_pre_smooth = p.n_pre_smooth;
    q.adapt = (float)p.adapt_correction_factor;
    for (int i = 0; i < 8; ++i)
        q.mask.dir[i] = (p.macro_border_mask[i] == 2) ? 1 : 0;
    q.max_outer = p.max_outer;
    return q;
}

namespace
{
    // The occupancy of every one-pass kernel, looked up at set-up, never
    // under stream capture.
    // Dynamic LDS v_bottom_kernel may take on the current device (cached per
    // device): the device's per-workgroup limit, once the kernels are allowed
    // more than 64 KiB, else 60 KiB (4 KiB of either left for its static LDS)
    inline unsigned bottom_lds_limit()
    {
        static std::mutex mu;
        static std::map<int, unsigned> cache;
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(dev);
        if (it != cache.end())
            return it->second;
        int v = 0;
        HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // (4 KiB left for the kernel's static LDS)
        const int want = v - 4096;
        unsigned r = (unsigned)std::min(std::max(want, 0), 65536 - 4096);
        if (want > 65536 &&
            hipFuncSetAttribute((const void *)v_bottom_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
                hipSuccess &&
            hipFuncSetAttribute((const void *)v_bottom_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
                hipSuccess)
            r = (unsigned)want;
        (void)hipGetLastError();
        cache.emplace(dev, r);
        return r;
    }

    void warm_post_kernels()
    {
#define HS_K(KS)                                                                                               \
    (const void *)v_pre2r_kernel<0, 0, KS>, (const void *)v_pre2r_kernel<0, 1, KS>,                            \
        (const void *)v_pre2r_kernel<2, 0, KS>, (const void *)v_pre2r_kernel<2, 1, KS>,                        \
        (const void *)v_post2_kernel<0, 0, 0, KS>, (const void *)v_post2_kernel<0, 0, 1, KS>,                  \
        (const void *)v_post2_kernel<0, 1, 0, KS>, (const void *)v_post2_kernel<0, 1, 1, KS>,                  \
        (const void *)v_post2_kernel<2, 0, 0, KS>, (const void *)v_post2_kernel<2, 0, 1, KS>,                  \
        (const void *)v_post2_kernel<2, 1, 0, KS>, (const void *)v_post2_kernel<2, 1, 1, KS>,                  \
        (const void *)v_pre2r_kernel<0, 0, KS, true>, (const void *)v_pre2r_kernel<0, 1, KS, true>,            \
        (const void *)v_post2_kernel<0, 0, 0, KS, true>, (const void *)v_post2_kernel<0, 0, 1, KS, true>,      \
        (const void *)v_post2_kernel<0, 1, 0, KS, true>, (const void *)v_post2_kernel<0, 1, 1, KS, true>,      \
        (const void *)v_pre2r_kernel<2, 0, KS, true>, (const void *)v_pre2r_kernel<2, 1, KS, true>,            \
        (const void *)v_post2_kernel<2, 0, 0, KS, true>, (const void *)v_post2_kernel<2, 0, 1, KS, true>,      \
        (const void *)v_post2_kernel<2, 1, 0, KS, true>, (const void *)v_post2_kernel<2, 1, 1, KS, true>
        const void * fns[] = {HS_K(1), HS_K(kPostSlots)};
#undef HS_K
        for (const void * f : fns)
            (void)post_resident(f);
        (void)bottom_lds_limit();
    }
}

void hierarchy_alloc(Hierarchy & h, const Params & p, uint64_t max_level)
{
    warm_post_kernels();
    require(p.min_level >= 1 && p.min_level <= max_level && max_level < (uint64_t)kMaxLevels, HSOLVE_ERR_INVALID,
            "mg: need 1 <= min_level <= max_level <= 14");
    require(p.n_pre_smooth >= 1, HSOLVE_ERR_INVALID, "mg: n_pre_smooth must be >= 1");
    require(p.n_max_iter >= 1, HSOLVE_ERR_INVALID, "mg: n_max_iter must be >= 1");
    require(size_of(p.min_level) <= 16641, HSOLVE_ERR_INVALID, "mg: coarsest level above 129^2");
    h.p = p;
    h.max_level = max_level;
    for (uint64_t L = p.min_level; L <= max_level; ++L)
    {
        HierLevel & l = h.lv[L];
        l.side = side_of(L);
        l.n = size_of(L);
        l.bs = band_stride(l.n);
        l.bands = (float *)h.alloc(sizeof(float) * 9 * l.bs);
        l.dinv = (float *)h.alloc(sizeof(float) * l.bs);
    }
}

void hierarchy_alloc64(Hierarchy & h, uint64_t L)
{
    require(L >= h.p.min_level && L <= h.max_level, HSOLVE_ERR_INVALID, "mg: FP64 level outside the hierarchy");
    if (h.bands64[L])
        return;
    h.bs64[L] = band_stride(size_of(L));
    h.bands64[L] = (double *)h.alloc(sizeof(double) * 9 * h.bs64[L]);
}

void hierarchy_dinv(Hierarchy & h, uint64_t L, hipStream_t st)
{
    HierLevel & l = h.lv[L];
    hipLaunchKernelGGL(dinv_kernel, dim3(grid_for(l.n, kBlock)), dim3(kBlock), 0, st,
                       (const float *)(l.bands + 4 * l.bs), l.dinv, (int64_t)l.n);
    HIP_CHECK(hipGetLastError());
}

void hierarchy_fill_poisson(Hierarchy & h, hipStream_t st)
{
    for (uint64_t L = h.p.min_level; L <= h.max_level; ++L)
    {
        HierLevel & l = h.lv[L];
        hipLaunchKernelGGL((k::fill_poisson_kernel<float>), dim3(grid_for(l.n, kBlock)), dim3(kBlock), 0, st, l.B(),
                           (int64_t)l.n, l.side);
        hierarchy_dinv(h, L, st);
        if (h.bands64[L])
            hipLaunchKernelGGL((k::fill_poisson_kernel<double>), dim3(grid_for(l.n, kBlock)), dim3(kBlock), 0, st,
                               h.B64(L), (int64_t)l.n, l.side);
    }
    HIP_CHECK(hipGetLastError());
    hierarchy_codes(h);
}

// Value-indexed storage of a level's bands (Q1, or a 9-point structured ELL
// matrix kept as bands) and dinv (the value-index
// compression of Kourtis, Goumas and Koziris, "Optimizing sparse
// matrix-vector multiplication using index and value compression", 2008):
// when each of the ten arrays holds at most 16 distinct values (as bit
// patterns) over the level's n nodes, node i's ten 4-bit indices go into one
// 64-bit word and the values into a 16-entry table per array.  The one-pass
// kernels then read 8 bytes per node instead of 40 and look the values up
// in LDS: the same floats, so the same bits.  Levels whose arrays hold more
// values keep the plain bands only.
void hierarchy_codes(Hierarchy & h)
{
    if (h.kind == 1)    // ELL matrices: no bands
        return;
    HIP_CHECK(hipDeviceSynchronize());
    for (uint64_t L = h.p.min_level; L <= h.max_level; ++L)
    {
        HierLevel & l = h.lv[L];
        l.codes = nullptr;
        l.ctab = nullptr;
        if (l.n < 4225)     // the one-pass kernels run on levels of 65^2 and up
            continue;
        const uint64_t n = l.n;
        std::vector<uint32_t> v(n);
        std::vector<uint64_t> code(n, 0);
        uint32_t tab[160] = {};
        bool ok = true;
        for (int k = 0; k < 10 && ok; ++k)
        {
            const float * src = k < 9 ? l.bands + (uint64_t)k * l.bs : l.dinv;
            HIP_CHECK(hipMemcpy(v.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
            int cnt = 0;
            uint32_t last = 0;
            uint64_t last_t = 0;
            for (uint64_t i = 0; i < n; ++i)
            {
                const uint32_t bits = v[i];
                uint64_t t = last_t;
                if (cnt == 0 || bits != last)
                {
                    int f = -1;
                    for (int q = 0; q < cnt; ++q)
                        if (tab[16 * k + q] == bits)
                        {
                            f = q;
                            break;
                        }
                    if (f < 0)
                    {
                        if (cnt == 16)
                        {
                            ok = false;
                            break;
                        }
                        f = cnt;
                        tab[16 * k + cnt++] = bits;
                    }
                    t = (uint64_t)f;
                    last = bits;
                    last_t = t;
                }
                code[i] |= t << (4 * k);
            }
        }
        if (! ok)
            continue;
        l.codes = (uint64_t *)h.alloc(sizeof(uint64_t) * n);
        l.ctab = (float *)h.alloc(sizeof(tab));
        HIP_CHECK(hipMemcpy(l.codes, code.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(l.ctab, tab, sizeof(tab), hipMemcpyHostToDevice));
    }
    // the FP64 bands of the finest levels solved on (the outer defect, whose
    // coded path forms Product<Q1>'s order: Q1 hierarchies only)
    for (uint64_t L = h.p.min_level; h.kind == 0 && L <= h.max_level; ++L)
    {
        h.codes64[L] = nullptr;
        h.ctab64[L] = nullptr;
        if (! h.bands64[L] || size_of(L) < 4225)
            continue;
        const uint64_t n = size_of(L);
        std::vector<uint64_t> v(n), code(n, 0);
        uint64_t tab[144] = {};
        bool ok = true;
        for (int k = 0; k < 9 && ok; ++k)
        {
            HIP_CHECK(hipMemcpy(v.data(), h.bands64[L] + (uint64_t)k * h.bs64[L], sizeof(double) * n,
                                hipMemcpyDeviceToHost));
            int cnt = 0;
            uint64_t last = 0, last_t = 0;
            for (uint64_t i = 0; i < n; ++i)
            {
                const uint64_t bits = v[i];
                uint64_t t = last_t;
                if (cnt == 0 || bits != last)
                {
                    int f = -1;
                    for (int q = 0; q < cnt; ++q)
                        if (tab[16 * k + q] == bits)
                        {
                            f = q;
                            break;
                        }
                    if (f < 0)
                    {
                        if (cnt == 16)
                        {
                            ok = false;
                            break;
                        }
                        f = cnt;
                        tab[16 * k + cnt++] = bits;
                    }
                    t = (uint64_t)f;
                    last = bits;
                    last_t = t;
                }
                code[i] |= t << (4 * k);
            }
        }
        if (! ok)
            continue;
        h.codes64[L] = (uint64_t *)h.alloc(sizeof(uint64_t) * n);
        h.ctab64[L] = (double *)h.alloc(sizeof(tab));
        HIP_CHECK(hipMemcpy(h.codes64[L], code.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(h.ctab64[L], tab, sizeof(tab), hipMemcpyHostToDevice));
    }
}

namespace
{
    // HONEI-layout host ELL (threads == 1) -> engine ELL in device memory
    template <typename DT>
    EllMat<DT> upload_ell(Hierarchy & h, const EllHost & e)
    {
        require(e.Aj && e.Ax && e.Arl, HSOLVE_ERR_INVALID, "mg ELL level: null array");
        require(e.rows < (1ull << 32) && e.cols < (1ull << 32), HSOLVE_ERR_SIZE, "mg ELL level: too many rows");
        require(e.stride >= e.rows, HSOLVE_ERR_INVALID, "mg ELL level: stride below rows");
        const uint64_t size = e.stride * e.ncpr;
        std::vector<uint32_t> col(size), len(e.rows);
        for (uint64_t i = 0; i < size; ++i)
        {
            require(e.Aj[i] < (e.cols ? e.cols : 1), HSOLVE_ERR_INVALID, "mg ELL level: column out of range");
            col[i] = (uint32_t)e.Aj[i];
        }
        for (uint64_t r = 0; r < e.rows; ++r)
        {
            require(e.Arl[r] <= e.ncpr, HSOLVE_ERR_INVALID, "mg ELL level: row length above num_cols_per_row");
            len[r] = (uint32_t)e.Arl[r];
        }
        EllMat<DT> m;
        uint32_t * c = (uint32_t *)h.alloc(sizeof(uint32_t) * (size ? size : 1));
        DT * v = (DT *)h.alloc(sizeof(DT) * (size ? size : 1));
        uint32_t * l = (uint32_t *)h.alloc(sizeof(uint32_t) * (e.rows ? e.rows : 1));
        HIP_CHECK(hipMemcpy(c, col.data(), sizeof(uint32_t) * size, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(v, e.Ax, sizeof(DT) * size, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(l, len.data(), sizeof(uint32_t) * e.rows, hipMemcpyHostToDevice));
        m.col = c;
        m.val = v;
        m.len = l;
        m.stride = e.stride;
        m.rows = e.rows;
        m.cols = e.cols;
        m.ncpr = (uint32_t)e.ncpr;
        return m;
    }
}

namespace
{
    // neighbour position (dy + 1) * 3 + (dx + 1) of column c in row r of a
    // side-M grid, or -1 when c is not one of the 9 neighbours
    int q1_position(uint64_t r, uint64_t c, uint64_t M)
    {
        const int64_t dy = (int64_t)(c / M) - (int64_t)(r / M), dx = (int64_t)(c % M) - (int64_t)(r % M);
        if (dy < -1 || dy > 1 || dx < -1 || dx > 1)
            return -1;
        return (int)((dy + 1) * 3 + (dx + 1));
    }

    // 9-point structure of a host ELL on the (2^L+1)^2 grid: every stored
    // column of a row is one of its neighbours, in ascending order
    bool ell_is_q1(const EllHost & e, uint64_t L)
    {
        const uint64_t M = (uint64_t)side_of(L);
        if (e.rows != size_of(L) || e.cols != e.rows || e.ncpr > 9 || ! e.Aj || ! e.Ax || ! e.Arl)
            return false;
        for (uint64_t r = 0; r < e.rows; ++r)
        {
            if (e.Arl[r] > e.ncpr)
                return false;
            int prev = -1;
            for (uint64_t k = 0; k < e.Arl[r]; ++k)
            {
                const int q = q1_position(r, e.Aj[r + k * e.stride], M);
                if (q <= prev)
                    return false;
                prev = q;
            }
        }
        return true;
    }

    // the matrix as 9 bands LL .. UU (band b of row r at b * bs + r), zero
    // where the row has no entry
    template <typename DT>
    std::vector<DT> ell_bands(const EllHost & e, uint64_t L, uint64_t bs)
    {
        const uint64_t M = (uint64_t)side_of(L);
        std::vector<DT> out(9 * bs, DT(0));
        const DT * ax = (const DT *)e.Ax;
        for (uint64_t r = 0; r < e.rows; ++r)
            for (uint64_t k = 0; k < e.Arl[r]; ++k)
                out[(uint64_t)q1_position(r, e.Aj[r + k * e.stride], M) * bs + r] = ax[r + k * e.stride];
        return out;
    }
}

void hierarchy_alloc_ell(Hierarchy & h, const Params & p, uint64_t max_level, const EllHost * a32,
                         const EllHost & A64, const float * const * dinv32, const EllHost * prol32,
                         const EllHost * res32)
{
    require(p.min_level >= 1 && p.min_level <= max_level && max_level < (uint64_t)kMaxLevels, HSOLVE_ERR_INVALID,
            "mg: need 1 <= min_level <= max_level <= 14");
    require(p.n_pre_smooth >= 1, HSOLVE_ERR_INVALID, "mg: n_pre_smooth must be >= 1");
    require(p.n_max_iter >= 1, HSOLVE_ERR_INVALID, "mg: n_max_iter must be >= 1");
    require(a32 != nullptr, HSOLVE_ERR_INVALID, "mg: null level matrices");
    require((prol32 == nullptr) == (res32 == nullptr), HSOLVE_ERR_INVALID, "mg: give both prol and res or neither");
    // 9-point structured levels with structured transfers: kept as bands
    // (kind 2, summed in the ELL rows' order), read with 16-byte vectors
    bool banded = prol32 == nullptr && ell_is_q1(A64, max_level) && size_of(p.min_level) <= 16641;
    for (uint64_t L = p.min_level; banded && L <= max_level; ++L)
        banded = ell_is_q1(a32[L], L);
    if (banded)
    {
        hierarchy_alloc(h, p, max_level);
        hierarchy_alloc64(h, max_level);
        h.kind = 2;
        h.prolmat = 0;
        for (uint64_t L = p.min_level; L <= max_level; ++L)
        {
            HierLevel & l = h.lv[L];
            const std::vector<float> b = ell_bands<float>(a32[L], L, l.bs);
            HIP_CHECK(hipMemcpy(l.bands, b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice));
            std::vector<float> dv(l.n);
            if (dinv32 && dinv32[L])
                std::memcpy(dv.data(), dinv32[L], sizeof(float) * l.n);
            else
                for (uint64_t r = 0; r < l.n; ++r)
                {
                    // diags_inverted as the reference tests build it
                    const float dd = b[4 * l.bs + r];
                    const float inv = dd != 0.0f ? 1.0f / dd : 0.0f;
                    dv[r] = (float)((double)inv * 0.7);
                }
            HIP_CHECK(hipMemcpy(l.dinv, dv.data(), sizeof(float) * l.n, hipMemcpyHostToDevice));
        }
        const std::vector<double> b64 = ell_bands<double>(A64, max_level, h.bs64[max_level]);
        HIP_CHECK(hipMemcpy(h.bands64[max_level], b64.data(), sizeof(double) * b64.size(), hipMemcpyHostToDevice));
        hierarchy_codes(h);
        return;
    }
    h.p = p;
    h.kind = 1;
    h.prolmat = prol32 ? 1 : 0;
    h.max_level = max_level;
    for (uint64_t L = p.min_level; L <= max_level; ++L)
    {
        const EllHost & e = a32[L];
        require(e.rows == e.cols, HSOLVE_ERR_SIZE, "mg: level matrix is not square");
        HierLevel & l = h.lv[L];
        l.n = e.rows;
        l.side = (int64_t)std::sqrt((double)e.rows);
        if (! h.prolmat)
            require(e.rows == size_of(L), HSOLVE_ERR_SIZE,
                    "mg: structured transfers need level L of size (2^L+1)^2");
        l.ell = upload_ell<float>(h, e);
        l.dinv = (float *)h.alloc(sizeof(float) * (l.n ? l.n : 1));
        std::vector<float> dv(l.n);
        if (dinv32 && dinv32[L])
            std::memcpy(dv.data(), dinv32[L], sizeof(float) * l.n);
        else
        {
            // diags_inverted as the reference tests build it: ElementInverse, then
            // Scale by 0.7 in double
            const float * ax = (const float *)e.Ax;
            for (uint64_t r = 0; r < l.n; ++r)
            {
                float dd = 0.0f;
                for (uint64_t k = 0; k < e.Arl[r]; ++k)
                    if (e.Aj[r + k * e.stride] == r)
                    {
                        dd = ax[r + k * e.stride];
                        break;
                    }
                const float inv = dd != 0.0f ? 1.0f / dd : 0.0f;
                dv[r] = (float)((double)inv * 0.7);
            }
        }
        HIP_CHECK(hipMemcpy(l.dinv, dv.data(), sizeof(float) * l.n, hipMemcpyHostToDevice));
        if (h.prolmat && L > p.min_level)
        {
            require(prol32[L].rows == l.n && prol32[L].cols == a32[L - 1].rows, HSOLVE_ERR_SIZE,
                    "mg: prolongation matrix shape");
            require(res32[L].rows == a32[L - 1].rows && res32[L].cols == l.n, HSOLVE_ERR_SIZE,
                    "mg: restriction matrix shape");
            l.prol = upload_ell<float>(h, prol32[L]);
            l.res = upload_ell<float>(h, res32[L]);
        }
    }
    require(h.lv[p.min_level].n <= 1048576, HSOLVE_ERR_INVALID, "mg: coarsest level above 2^20 rows");
    require(A64.rows == h.lv[max_level].n && A64.cols == A64.rows, HSOLVE_ERR_SIZE,
            "mg: FP64 matrix size differs from the finest level");
    h.ell64[max_level] = upload_ell<double>(h, A64);
}

uint64_t slot_bytes(const Hierarchy & h, uint64_t L, bool own_b, bool own_result)
{
    const uint64_t n = h.lv[L].n;
    uint64_t b = align256(sizeof(Scalars)) + align256(sizeof(double) * 2 * kNormBlocks);
    b += align256(sizeof(double) * n);                                  // od
    if (own_b)
        b += align256(sizeof(double) * n);
    if (own_result)
        b += align256(sizeof(double) * n);
    b += align256(sizeof(float) * 4 * h.lv[h.p.min_level].n);           // coarse CG
    for (uint64_t l = h.p.min_level; l <= L; ++l)
        b += 5 * align256(sizeof(float) * h.lv[l].n);
    return b;
}

Slot slot_carve(const Hierarchy & h, uint64_t L, char * base, bool own_b, bool own_result)
{
    Slot sl;
    std::memset(&sl, 0, sizeof(sl));
    const uint64_t n = h.lv[L].n;
    char * c = base;
    auto take = [&](uint64_t bytes) {
        char * r = c;
        c += align256(bytes);
        return r;
    };
    sl.s = (Scalars *)take(sizeof(Scalars));
    sl.partials = (double *)take(sizeof(double) * 2 * kNormBlocks);
    sl.od = (double *)take(sizeof(double) * n);
    if (own_b)
        sl.b = (const double *)take(sizeof(double) * n);
    if (own_result)
        sl.result = (double *)take(sizeof(double) * n);
    sl.cg_work = (float *)take(sizeof(float) * 4 * h.lv[h.p.min_level].n);
    for (uint64_t l = h.p.min_level; l <= L; ++l)
        for (int f = 0; f < 5; ++f)
            sl.v[l][f] = (float *)take(sizeof(float) * h.lv[l].n);
    return sl;
}

Scalars start_scalars(bool zero_guess)
{
    Scalars s;
    std::memset(&s, 0, sizeof(s));
    s.phase = PH_START;
    s.outer_active = 1;
    s.zero_guess = zero_guess ? 1 : 0;
    return s;
}

void enqueue_rhs_uniform(double * b, uint64_t L, uint64_t seed, uint64_t stream, hipStream_t st)
{
    const uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ull * (stream + 1));
    const uint64_t n = size_of(L);
    hipLaunchKernelGGL(rhs_uniform_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, b, (int64_t)n,
                       side_of(L), key);
    HIP_CHECK(hipGetLastError());
}

void enqueue_report(const Slot * table, unsigned nslots, Report * reports, hipStream_t st)
{
    hipLaunchKernelGGL(report_kernel, dim3(grid_for(nslots, 64)), dim3(64), 0, st, table, nslots, reports);
    HIP_CHECK(hipGetLastError());
}

namespace
{
template <int MK>
void tick_impl(const Hierarchy & h, uint64_t maxl, const Slot * T, unsigned S, hipStream_t st)
{
    const Params & p = h.p;
    const uint64_t minl = p.min_level;
    require(maxl >= minl && maxl <= h.max_level && (MK == 1 ? h.ell64[maxl].val != nullptr : h.bands64[maxl] != nullptr),
            HSOLVE_ERR_INVALID,
            "mg tick: level not prepared");
    const HierLevel & top = h.lv[maxl];
    const int64_t n = (int64_t)top.n;
    const dim3 blk(kBlock);
    auto largs = [&](uint64_t L) {
        const HierLevel & l = h.lv[L];
        LevelArgs a;
        a.B = l.B();
        a.E = l.ell;
        a.dinv = l.dinv;
        a.n = (int64_t)l.n;
        a.m = l.side;
        a.bs = (int64_t)l.bs;
        a.codes = h.use_codes ? l.codes : nullptr;
        a.ctab = h.use_codes ? l.ctab : nullptr;
        return a;
    };
    auto norm_grid = [&](uint64_t nn) { return dim3(std::min<unsigned>(lanes4(nn), kNormBlocks), 1, S); };
    // row kernels: 4 rows per lane on Q1 bands, one row per lane on ELL
    auto rgrid = [&](uint64_t nn) { return kBands<MK> ? lanes4(nn) : grid_for(nn, kBlock); };
    // Post-smoothing in one pass (v_post2_kernel) on Q1 bands with two
    // sweeps: a level's smoothed iterate then lands in F_TMP, so the
    // prolongation out of a level above the coarsest reads F_TMP, and the
    // finest level's iterate lives in F_TMP between V-cycles (the start
    // writes it there, the outer step and the next cycle's first sweeps
    // read it there).
    // (the one-pass kernels address a level's nine bands by 32-bit offsets
    // from the first: levels up to 8193^2)
    const bool post2 = kBands<MK> && p.n_pre_smooth == 2 && maxl > minl && 36 * h.lv[maxl].bs < (1ull << 32);
    const int fx_top = post2 ? F_TMP : F_X;
    // Pre-smoothing, defect and restriction in one pass (v_pre2r_kernel)
    // when the transfers are the structured ones and the correction factor
    // is fixed (the adaptive one reads the level's defect again).
    const bool pre2r = post2 && p.adapt != 0.0f && ! h.prolmat;
    // With the one-pass kernels, levels minl .. btop (btop <= 5, below the
    // finest, as many as fit the LDS of one workgroup) run in one
    // v_bottom_kernel launch (one workgroup per slot); btop = 0: none.
    int btop = 0;
    BottomArgs ba{};
    size_t bbytes = 0;
    if (pre2r)
    {
        for (int top = (int)std::min<uint64_t>(kBottomMax, maxl - 1); top > (int)minl; --top)
        {
            size_t fl = 0;
            for (int L = (int)minl; L <= top; ++L)
                fl += (size_t)bottom_floats((int)h.lv[L].n, L == (int)minl);
            if (fl * sizeof(float) <= bottom_lds_limit())
            {
                btop = top;
                bbytes = fl * sizeof(float);
                break;
            }
        }
        if (btop)
        {
            int off = 0;
            for (int L = (int)minl; L <= btop; ++L)
            {
                ba.lv[L] = largs(L);
                ba.off[L] = off;
                off += bottom_floats((int)h.lv[L].n, L == (int)minl);
            }
            ba.minl = (int)minl;
            ba.top = btop;
            ba.alpha = p.adapt;
            ba.mk = p.mask;
            ba.konv_rad = (double)1.19209290e-07F;
            ba.max_iters = (uint64_t)(100 * h.lv[minl].n + 1000);
        }
    }

    // ---- gates, start of new solves, FP64 outer step, opening of inner solves
    hipLaunchKernelGGL(plan_kernel, dim3(grid_for(S, 64)), dim3(64), 0, st, T, S);
    // the per-tick kernels that most ticks skip (start, outer step, opening
    // of an inner solve) loop over their rows on a capped grid, so a skipped
    // launch costs a few thousand empty workgroups, not one per 256 rows
    auto ggrid = [&](uint64_t blocks) {
        return dim3((unsigned)std::min<uint64_t>(blocks, std::max<unsigned>(64u, kGateBlocks / S)), 1, S);
    };
    hipLaunchKernelGGL(start_init_kernel, ggrid(grid_for(top.n, kBlock)), blk, 0, st, T, (int)maxl, n,
                       top.side, fx_top);
    if (maxl > minl)
    {
        LevelSizes sz;
        for (int l = 0; l < kMaxLevels; ++l)
            sz.n[l] = (int64_t)h.lv[l].n;
        hipLaunchKernelGGL(start_zero_kernel, ggrid(grid_for(h.lv[maxl - 1].n, kBlock)), blk, 0, st, T,
                           (int)minl, (int)maxl, sz);
    }
    hipLaunchKernelGGL(outer_accumulate_kernel, ggrid(grid_for(top.n, kBlock)), blk, 0, st, T, (int)maxl, n,
                       fx_top);
    const unsigned od_blocks = std::min<unsigned>(grid_for((top.n + 1) / 2, kBlock), kNormBlocks);
    bool od_coded = false;
    if constexpr (MK == 0)
    {
        if (h.use_codes && h.codes64[maxl])
        {
            hipLaunchKernelGGL((outer_defect_kernel<MK, true>), dim3(od_blocks, 1, S), blk, 0, st, h.B64(maxl),
                               h.ell64[maxl], T, n, top.side, (const uint64_t *)h.codes64[maxl],
                               (const double *)h.ctab64[maxl]);
            od_coded = true;
        }
    }
    if (! od_coded)
        hipLaunchKernelGGL((outer_defect_kernel<MK>), dim3(od_blocks, 1, S), blk, 0, st,
                           kBands<MK> ? h.B64(maxl) : k::Bands{}, h.ell64[maxl], T, n, top.side,
                           (const uint64_t *)nullptr, (const double *)nullptr);
    hipLaunchKernelGGL(outer_defect_finish_kernel, dim3(1, 1, S), blk, 0, st, T, od_blocks, (double)p.tolerance,
                       p.max_outer);
    hipLaunchKernelGGL(to_inner_kernel, ggrid(grid_for(top.n, kBlock)), blk, 0, st, T, (int)maxl, n);
    if (p.initial_zero)
        hipLaunchKernelGGL((open_inner_kernel<1, MK>), ggrid(lanes4(top.n)), blk, 0, st, largs(maxl), T,
                           (int)maxl, fx_top);
    else
        hipLaunchKernelGGL((open_inner_kernel<0, MK>), ggrid(rgrid(top.n)), blk, 0, st, largs(maxl), T,
                           (int)maxl, fx_top);
    hipLaunchKernelGGL((inner_norm_kernel<0>), norm_grid(top.n), blk, 0, st, T, (int)maxl, n);
    hipLaunchKernelGGL((inner_norm_finish_kernel<0>), dim3(1, 1, S), blk, 0, st, T, norm_grid(top.n).x, p.tolerance,
                       p.convergence_check, p.n_max_iter);

    // ---- one V-cycle (multigrid.hh:1660-1905)
    const uint64_t npre = p.n_pre_smooth;
    auto sweeps = [&](uint64_t L, int & xc, int & xa, int frhs, uint64_t count) {
        const LevelArgs a = largs(L);
        for (uint64_t i = 0; i < count; ++i)
        {
            hipLaunchKernelGGL((v_sweep_kernel<MK>), dim3(rgrid(a.n), 1, S), blk, 0, st, a, T, (int)L, xc, xa, frhs);
            std::swap(xc, xa);
        }
    };
    auto move = [&](uint64_t L, int fsrc, int fdst, bool add) {
        const int64_t nn = (int64_t)h.lv[L].n;
        if (add)
            hipLaunchKernelGGL((v_move_kernel<1>), dim3(lanes4(nn), 1, S), blk, 0, st, T, (int)L, nn, fsrc, fdst);
        else
            hipLaunchKernelGGL((v_move_kernel<0>), dim3(lanes4(nn), 1, S), blk, 0, st, T, (int)L, nn, fsrc, fdst);
    };
    auto defect = [&](uint64_t L) {
        const LevelArgs a = largs(L);
        hipLaunchKernelGGL((v_defect_kernel<MK>), dim3(rgrid(a.n), 1, S), blk, 0, st, a, T, (int)L);
    };

    if (minl != maxl)
    {
        for (uint64_t L = maxl; L > minl; --L)
        {
            if (btop && L <= (uint64_t)btop)
                break;
            const LevelArgs a = largs(L);
            // the finest level's defect is d; coarser levels' defect is the
            // restricted one in rhs
            const int dsrc = (L == maxl) ? F_D : F_RHS;
            if constexpr (kBands<MK>)
            {
                if (pre2r)
                {
                    const int64_t Nc = h.lv[L - 1].side;
                    auto go = [&](auto acc_tag, auto k_tag) {
                        constexpr int ACC = decltype(acc_tag)::value, KS = decltype(k_tag)::value;
                        const unsigned groups = (S + KS - 1) / KS;
                        auto launch = [&](auto cb_tag) {
                            constexpr bool CB = decltype(cb_tag)::value;
                            const int th =
                                    post_rows(a.m, groups, post_resident((const void *)v_pre2r_kernel<MK, ACC, KS, CB>));
                            hipLaunchKernelGGL((v_pre2r_kernel<MK, ACC, KS, CB>), dim3(post_tiles(a.m, th), 1, groups),
                                               dim3(kPostLanes), 0, st, a, T, (int)L, dsrc, ACC ? fx_top : (int)F_X, Nc,
                                               p.mask, th, (int)S);
                        };
                        // value-coded bands (levels whose bands allow them)
                        if constexpr (MK != 1)
                        {
                            if (a.codes)
                                launch(std::true_type{});
                            else
                                launch(std::false_type{});
                        }
                        else
                            launch(std::false_type{});
                    };
                    if (L == maxl)
                        S > 1 ? go(IntC<1>{}, IntC<kPostSlots>{}) : go(IntC<1>{}, IntC<1>{});
                    else
                        S > 1 ? go(IntC<0>{}, IntC<kPostSlots>{}) : go(IntC<0>{}, IntC<1>{});
                    continue;
                }
            }
            if (L == maxl)
            {
                // c = dinv*d, (npre-1) sweeps on c with rhs d, x += c
                if (npre == 1)
                    hipLaunchKernelGGL((v_zero1_kernel<1>), dim3(grid_for(a.n, kBlock), 1, S), blk, 0, st, a, T,
                                       (int)L, dsrc, (int)F_X);
                else if (npre == 2)
                    hipLaunchKernelGGL((v_zero2_kernel<1, MK>), dim3(rgrid(a.n), 1, S), blk, 0, st, a, T, (int)L, dsrc,
                                       (int)F_X, fx_top);
                else
                {
                    hipLaunchKernelGGL((v_zero2_kernel<0, MK>), dim3(rgrid(a.n), 1, S), blk, 0, st, a, T, (int)L, dsrc,
                                       (int)F_C, (int)F_C);
                    int cc = F_C, ca = F_TMP;
                    sweeps(L, cc, ca, dsrc, npre - 2);
                    move(L, cc, F_X, true);
                }
            }
            else
            {
                // x = dinv*d, (npre-1) sweeps on x with rhs d
                if (npre == 1)
                    hipLaunchKernelGGL((v_zero1_kernel<0>), dim3(grid_for(a.n, kBlock), 1, S), blk, 0, st, a, T,
                                       (int)L, dsrc, (int)F_X);
                else
                {
                    hipLaunchKernelGGL((v_zero2_kernel<0, MK>), dim3(rgrid(a.n), 1, S), blk, 0, st, a, T, (int)L, dsrc,
                                       (int)F_X, (int)F_X);
                    int xc = F_X, xa = F_TMP;
                    sweeps(L, xc, xa, dsrc, npre - 2);
                    if (xc != F_X)
                        move(L, xc, F_X, false);
                }
            }
            defect(L);
            if (h.prolmat)
                hipLaunchKernelGGL(v_apply_ell_kernel, dim3(grid_for(h.lv[L].res.rows, kBlock), 1, S), blk, 0, st, h.lv[L].res,
                                   T, (int)L, (int)F_D, (int)(L - 1), (int)F_RHS);
            else
            {
                const int64_t Nc = h.lv[L - 1].side;
                hipLaunchKernelGGL(v_restrict_kernel, dim3(grid_for(Nc, kBlock), (unsigned)Nc, S), blk, 0, st, T,
                                   (int)L, Nc, p.mask);
            }
        }
    }
    {
        // coarse grid solve: ConjugateGradients<NONE>(A, d, x, FLT_EPSILON);
        // with one level it sees the finest defect d
        const LevelArgs a = largs(minl);
        const int crhs = (minl == maxl) ? F_D : F_RHS;
        if (btop)
            hipLaunchKernelGGL((v_bottom_kernel<MK>), dim3(1, 1, S), dim3(1024), bbytes, st, ba, T);
        else if (a.n <= kCoarseLdsMax)
            hipLaunchKernelGGL((v_coarse_lds_kernel<MK>), dim3(1, 1, S), dim3(64), 0, st, a, T, (int)minl, crhs,
                               (double)1.19209290e-07F, (uint64_t)(100 * a.n + 1000));
        else
            hipLaunchKernelGGL((v_coarse_kernel<MK>), dim3(1, 1, S), blk, 0, st, a, T, (int)minl, crhs,
                               (double)1.19209290e-07F, (uint64_t)(100 * a.n + 1000));
        if (minl == maxl)
        {
            defect(minl);
            hipLaunchKernelGGL(one_level_end_kernel, dim3(grid_for(S, 64)), dim3(64), 0, st, T, S);
            HIP_CHECK(hipGetLastError());
            return;
        }
    }
    // v_post2_kernel on level L (DEFECT on the finest), kPostSlots slots per
    // workgroup when the tick runs several
    auto post2_launch = [&](uint64_t L, const LevelArgs & a, int fcoarse, float alpha, auto prol_tag) {
        constexpr int PROL = decltype(prol_tag)::value;
        auto go = [&](auto def_tag, auto k_tag) {
            constexpr int DEF = decltype(def_tag)::value, KS = decltype(k_tag)::value;
            const unsigned groups = (S + KS - 1) / KS;
            auto launch = [&](auto cb_tag) {
                constexpr bool CB = decltype(cb_tag)::value;
                const int th = post_rows(a.m, groups, post_resident((const void *)v_post2_kernel<MK, DEF, PROL, KS, CB>));
                hipLaunchKernelGGL((v_post2_kernel<MK, DEF, PROL, KS, CB>), dim3(post_tiles(a.m, th), 1, groups),
                                   dim3(kPostLanes), 0, st, a, T, (int)L, (int)F_X, (int)F_TMP, fcoarse, alpha, p.mask,
                                   th, (int)S);
            };
            if constexpr (MK != 1)
            {
                if (a.codes)
                    launch(std::true_type{});
                else
                    launch(std::false_type{});
            }
            else
                launch(std::false_type{});
        };
        if (L == maxl)
            S > 1 ? go(IntC<1>{}, IntC<kPostSlots>{}) : go(IntC<1>{}, IntC<1>{});
        else
            S > 1 ? go(IntC<0>{}, IntC<kPostSlots>{}) : go(IntC<0>{}, IntC<1>{});
    };
    for (uint64_t L = minl + 1; L <= maxl; ++L)
    {
        if (btop && L <= (uint64_t)btop)
            continue;
        const LevelArgs a = largs(L);
        const int Nf = (int)h.lv[L].side;
        const int Nc = (int)h.lv[L - 1].side;
        const dim3 pg(grid_for((uint64_t)Nc, kBlock), (unsigned)Nf, S);
        // the coarse level's iterate: the coarse solve's F_X, or F_TMP
        // after that level's one-pass post-smoothing
        const int fcoarse = (post2 && L - 1 > minl) ? F_TMP : F_X;
        if constexpr (kBands<MK>)
        {
            if (pre2r)
            {
                // prolongation, correction and post-smoothing in one pass
                post2_launch(L, a, fcoarse, p.adapt, IntC<1>{});
                continue;
            }
        }
        if (h.prolmat)
        {
            // c = P x_coarse, then x += alpha c
            hipLaunchKernelGGL(v_apply_ell_kernel, dim3(grid_for(h.lv[L].prol.rows, kBlock), 1, S), blk, 0, st, h.lv[L].prol, T,
                               (int)(L - 1), fcoarse, (int)L, (int)F_C);
            if (p.adapt != 0.0f)
                hipLaunchKernelGGL(v_axpy_const_kernel, dim3(grid_for(a.n, kBlock), 1, S), blk, 0, st, T, (int)L, a.n,
                                   p.adapt);
        }
        else if (p.adapt != 0.0f)
            hipLaunchKernelGGL((v_prolong_kernel<true>), pg, blk, 0, st, T, (int)L, Nf, p.adapt, p.mask, (int)F_X,
                               fcoarse);
        else
            hipLaunchKernelGGL((v_prolong_kernel<false>), pg, blk, 0, st, T, (int)L, Nf, 0.0f, p.mask, (int)F_C,
                               fcoarse);
        if (p.adapt == 0.0f)
        {
            const unsigned ab = std::min<unsigned>(grid_for(a.n, kBlock), kNormBlocks);
            hipLaunchKernelGGL((v_adapt_alpha_kernel<MK>), dim3(ab, 1, S), blk, 0, st, a, T, (int)L);
            hipLaunchKernelGGL(v_adapt_alpha_finish_kernel, dim3(1, 1, S), blk, 0, st, T, ab);
            hipLaunchKernelGGL(v_axpy_alpha_kernel, dim3(grid_for(a.n, kBlock), 1, S), blk, 0, st, T, (int)L, a.n);
        }
        // post-smoothing: n_pre_smooth sweeps (multigrid.hh:1851)
        if constexpr (kBands<MK>)
        {
            if (post2)
            {
                post2_launch(L, a, fcoarse, 0.0f, IntC<0>{});
                continue;
            }
        }
        int xc = F_X, xa = F_TMP;
        sweeps(L, xc, xa, F_RHS, npre);
        if (xc != F_X)
            move(L, xc, F_X, false);
        if (L == maxl)
            defect(L);
    }
    hipLaunchKernelGGL((inner_norm_kernel<1>), norm_grid(top.n), blk, 0, st, T, (int)maxl, n);
    hipLaunchKernelGGL((inner_norm_finish_kernel<1>), dim3(1, 1, S), blk, 0, st, T, norm_grid(top.n).x, p.tolerance,
                       p.convergence_check, p.n_max_iter);
    HIP_CHECK(hipGetLastError());
}
}  // namespace

void enqueue_tick(const Hierarchy & h, uint64_t maxl, const Slot * T, unsigned S, hipStream_t st)
{
    if (h.kind == 1)
        tick_impl<1>(h, maxl, T, S, st);
    else if (h.kind == 2)
        tick_impl<2>(h, maxl, T, S, st);
    else
        tick_impl<0>(h, maxl, T, S, st);
}

}  // namespace mg
}  // namespace hsolve
