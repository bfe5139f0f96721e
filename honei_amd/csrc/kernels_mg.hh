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

    // The same predicate for node (r, s) of the side-M grid, without 64-bit
    // division (the transfer kernels run on a 2-D grid of rows x columns).
    __device__ __forceinline__ bool dirichlet_rs(int r, int s, int M, const Mask & mk)
    {
        // corners first (0: r = 0, s = 0; 1: r = 0, s = M-1; 2: r = M-1, s =
        // M-1; 3: r = M-1, s = 0), then the edges r = 0, s = M-1, r = M-1,
        // s = 0 in that priority; as selects, so lanes do not diverge
        const bool r0 = r == 0, rM = r == M - 1, s0 = s == 0, sM = s == M - 1;
        const bool corner = (r0 || rM) && (s0 || sM);
        const bool c = r0 ? (s0 ? mk.dir[0] : mk.dir[1]) : (sM ? mk.dir[2] : mk.dir[3]);
        const bool e = r0 ? mk.dir[4] : sM ? mk.dir[5] : rM ? mk.dir[6] : s0 ? mk.dir[7] : false;
        return corner ? c : e;
    }

    // (DT)(acc + w v) evaluated in double, as the reference's transfers do
    // (a DT value times a double weight, added to a DT accumulator).  Every
    // caller's weight is 0.5 or 0.25.
    template <typename DT>
    __device__ __forceinline__ DT add_d(DT acc, double w, DT v)
    {
        return (DT)((double)acc + w * (double)v);
    }

    // For float the same value in one fused multiply-add.  With w a power of
    // two, w v is exact in double, so the double sum is acc + w v rounded
    // once to 53 bits and then to 24.  If the addends' exponents differ by at
    // most 29 the sum fits 53 bits and that first rounding is exact; if they
    // differ by more, the smaller addend lies below the bit under acc's last
    // float bit, so the 53-bit result is no float rounding midpoint.  Either
    // way the two roundings give what rounding the exact sum once to float
    // gives, which is what fma(w, v, acc) returns (signed zeros and
    // overflow included).
    template <>
    __device__ __forceinline__ float add_d<float>(float acc, double w, float v)
    {
        return __builtin_fmaf((float)w, v, acc);
    }

    // Restriction<Tag, NONE> at coarse node (r, s) of a side-Nc grid, with
    // F(dr, ds) the fine value at (2r + dr, 2s + ds).  The reference visits
    // the node in its outer loop iterations i = r, r+1, s, s+1 (those in
    // [1, Nc-1], ascending, each distinct value once) and, inside one
    // iteration, runs the statements
    //   (a) r == i:     0.5 F(-1, 0)      (e) r == i-1, s <= Nc-2: 0.25 F(1, 1)
    //   (b) r == i-1:   0.5 F(1, 0)       (f) r == i-1, s >= 1:    0.25 F(1, -1)
    //   (c) s == i-1:   0.5 F(0, 1)       (g) r == i, s <= Nc-2:   0.25 F(-1, 1)
    //   (d) s == i:     0.5 F(0, -1)      (h) r == i, s >= 1:      0.25 F(-1, -1)
    // in that order.  Each statement's condition fixes its iteration (i = r
    // for a, g, h; r + 1 for b, e, f; s + 1 for c; s for d), so a statement
    // runs iff that i lies in [1, Nc-1] and its own test holds, and the
    // sequence is the statements sorted by (i, letter): a g h b e f with d
    // and c placed by how s compares to r:
    //   s < r-1:  d c a g h b e f        s = r+1:  a g h b d e f c
    //   s = r-1:  d a c g h b e f        s > r+1:  a g h b e f d c
    //   s = r:    a d g h b c e f
    // Below, twelve fixed slots in that order, each added when it applies
    // (a select, no branch), cover all five cases.
    template <typename DT, typename FV>
    __device__ __forceinline__ DT restrict_node(FV && F, int r, int s, int Nc, const Mask & mk)
    {
        const DT fa = F(-1, 0), fb = F(1, 0), fc = F(0, 1), fd = F(0, -1);
        const DT fe = F(1, 1), ff = F(1, -1), fg = F(-1, 1), fh = F(-1, -1);
        const bool r_lo = r >= 1, r_hi = r <= Nc - 2, s_lo = s >= 1, s_hi = s <= Nc - 2;
        DT v = F(0, 0);
        auto put = [&](bool on, double w, DT f) {
            const DT t = add_d(v, w, f);
            v = on ? t : v;
        };
        put(s < r && s_lo, 0.5, fd);                                             // d      (s < r)
        put(s < r - 1 && s_hi, 0.5, fc);                                         // c      (s < r-1)
        put(r_lo, 0.5, fa);                                                      // a
        put((s == r - 1 && s_hi) || (s == r && s_lo), 0.5, s == r - 1 ? fc : fd); // c | d  (s = r-1 | r)
        put(r_lo && s_hi, 0.25, fg);                                             // g
        put(r_lo && s_lo, 0.25, fh);                                             // h
        put(r_hi, 0.5, fb);                                                      // b
        put((s == r && s_hi) || (s == r + 1 && s_lo), 0.5, s == r ? fc : fd);     // c | d  (s = r | r+1)
        put(r_hi && s_hi, 0.25, fe);                                             // e
        put(r_hi && s_lo, 0.25, ff);                                             // f
        put((s > r + 1 && s_lo) || (s == r + 1 && s_hi), 0.5, s > r + 1 ? fd : fc); // d | c  (s > r+1 | r+1)
        put(s > r + 1 && s_hi, 0.5, fc);                                         // c      (s > r+1)
        if (dirichlet_rs(r, s, Nc, mk))
            v = DT(0);
        return v;
    }

    // Restriction<Tag, NONE>: one lane per coarse node (r, s).  The nine
    // fine values around (2r, 2s) are loaded up front (indices clamped to the
    // grid; the clamped ones are never among those restrict_node adds).
    template <typename DT>
    __device__ __forceinline__ void restriction_body(DT * __restrict__ coarse, const DT * __restrict__ fine,
                                                     int64_t Nc, const Mask & mk, int64_t r, int64_t s)
    {
        if (s >= Nc)
            return;
        const int64_t Nf = 2 * Nc - 1;
        DT fv[3][3];
#pragma unroll
        for (int dr = 0; dr < 3; ++dr)
        {
            const int64_t row = min(max(2 * r + dr - 1, (int64_t)0), Nf - 1);
#pragma unroll
            for (int ds = 0; ds < 3; ++ds)
            {
                const int64_t col = min(max(2 * s + ds - 1, (int64_t)0), Nf - 1);
                fv[dr][ds] = fine[row * Nf + col];
            }
        }
        coarse[r * Nc + s] = restrict_node<DT>([&](int dr, int ds) { return fv[dr + 1][ds + 1]; }, (int)r, (int)s,
                                               (int)Nc, mk);
    }

    // coarse row blockIdx.y, columns along x
    template <typename DT>
    __global__ __launch_bounds__(256) void restriction_kernel(DT * __restrict__ coarse,
            const DT * __restrict__ fine, int64_t Nc, Mask mk, const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        restriction_body<DT>(coarse, fine, Nc, mk, blockIdx.y, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    }

    // Prolongation<Tag, NONE> (prolongation.hh:43-139) on a side-Nf grid.
    // AXPY == false: fine = P coarse.
    // AXPY == true:  fine += alpha * (P coarse) (Prolongation followed by
    //                ScaledSum(x, c, alpha): x = x + alpha*c).
    // One lane per pair of fine columns (2j, 2j+1) of fine row R =
    // blockIdx.y: the row parity is uniform over the block and the pair
    // shares its coarse loads, so no lane diverges.  Per fine node, in the
    // reference's addition order (accumulated in double from 0):
    //   (2i, 2j)      c[i][j]                      (copied)
    //   (2i, 2j+1)    0.5 c[i][j], 0.5 c[i][j+1]
    //   (2i-1, 2j)    0.5 c[i-1][j], 0.5 c[i][j]
    //   (2i-1, 2j+1)  0.25 c[i-1][j], c[i-1][j+1], c[i][j], c[i][j+1]
    // then the Dirichlet mask of the fine grid zeroes boundary nodes.
    template <typename DT, bool AXPY>
    __device__ __forceinline__ void prolongation2_body(DT * __restrict__ fine, const DT * __restrict__ coarse,
                                                       int Nf, DT alpha, const Mask & mk, int R, int j)
    {
        const int Nc = (Nf + 1) / 2;
        if (j >= Nc)
            return;
        const bool has1 = j + 1 < Nc;
        DT v0, v1 = DT(0);
        if ((R & 1) == 0)
        {
            const DT * c = coarse + (int64_t)(R >> 1) * Nc;
            const DT a = c[j];
            v0 = a;
            if (has1)
            {
                v1 = add_d(DT(0), 0.5, a);
                v1 = add_d(v1, 0.5, c[j + 1]);
            }
        }
        else
        {
            const int i = (R + 1) >> 1;
            const DT * cm = coarse + (int64_t)(i - 1) * Nc;
            const DT * cp = coarse + (int64_t)i * Nc;
            const DT a = cm[j], b = cp[j];
            v0 = add_d(DT(0), 0.5, a);
            v0 = add_d(v0, 0.5, b);
            if (has1)
            {
                v1 = add_d(DT(0), 0.25, a);
                v1 = add_d(v1, 0.25, cm[j + 1]);
                v1 = add_d(v1, 0.25, b);
                v1 = add_d(v1, 0.25, cp[j + 1]);
            }
        }
        const int S0 = 2 * j;
        if (dirichlet_rs(R, S0, Nf, mk))
            v0 = DT(0);
        DT * f = fine + (int64_t)R * Nf + S0;
        if (AXPY)
            f[0] = f[0] + alpha * v0;
        else
            f[0] = v0;
        if (has1)
        {
            if (dirichlet_rs(R, S0 + 1, Nf, mk))
                v1 = DT(0);
            if (AXPY)
                f[1] = f[1] + alpha * v1;
            else
                f[1] = v1;
        }
    }

    // Prolongation<Tag, NONE> at one fine node (R, S) of a side-Nf grid: the
    // same values as prolongation2_body (the reference's addition order in
    // double from 0, the fine grid's Dirichlet mask applied).
    template <typename DT>
    __device__ __forceinline__ DT prolongation_node(const DT * __restrict__ coarse, int Nf, const Mask & mk, int R,
                                                    int S)
    {
        const int Nc = (Nf + 1) / 2;
        const int jj = S >> 1;
        DT v;
        if ((R & 1) == 0)
        {
            const DT * c = coarse + (int64_t)(R >> 1) * Nc;
            if ((S & 1) == 0)
                v = c[jj];
            else
            {
                v = add_d(DT(0), 0.5, c[jj]);
                v = add_d(v, 0.5, c[jj + 1]);
            }
        }
        else
        {
            const int i = (R + 1) >> 1;
            const DT * cm = coarse + (int64_t)(i - 1) * Nc;
            const DT * cp = coarse + (int64_t)i * Nc;
            if ((S & 1) == 0)
            {
                v = add_d(DT(0), 0.5, cm[jj]);
                v = add_d(v, 0.5, cp[jj]);
            }
            else
            {
                v = add_d(DT(0), 0.25, cm[jj]);
                v = add_d(v, 0.25, cm[jj + 1]);
                v = add_d(v, 0.25, cp[jj]);
                v = add_d(v, 0.25, cp[jj + 1]);
            }
        }
        if (dirichlet_rs(R, S, Nf, mk))
            v = DT(0);
        return v;
    }

    // prolongation_node from values already in registers: c[0..3] = coarse
    // (R >> 1, S >> 1), (R >> 1, S >> 1 + 1), ((R + 1) >> 1, S >> 1),
    // ((R + 1) >> 1, S >> 1 + 1) (the second column clamped to the grid; the
    // values a node does not use may be anything).  Same additions, same bits.
    template <typename DT>
    __device__ __forceinline__ DT prolongation_node_v(const DT (&c)[4], int Nf, const Mask & mk, int R, int S)
    {
        // the cases of prolongation_node as one chain with selects (lanes of
        // both column parities share a wave): the weight is 0.5 with one
        // odd coordinate and 0.25 with two, the second term is c[1] for an
        // odd column and c[2] for an odd row
        const bool ro = (R & 1) != 0, so = (S & 1) != 0;
        const DT w = (ro && so) ? DT(0.25) : DT(0.5);
        const DT v1 = add_d(DT(0), (double)w, c[0]);
        const DT v2 = add_d(v1, (double)w, so ? c[1] : c[2]);
        const DT v4 = add_d(add_d(v2, (double)w, c[2]), (double)w, c[3]);
        DT v = (ro && so) ? v4 : ((ro || so) ? v2 : c[0]);
        if (dirichlet_rs(R, S, Nf, mk))
            v = DT(0);
        return v;
    }

    template <typename DT, bool AXPY>
    __global__ __launch_bounds__(256) void prolongation2_kernel(DT * __restrict__ fine,
            const DT * __restrict__ coarse, int Nf, DT alpha, Mask mk, const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        prolongation2_body<DT, AXPY>(fine, coarse, Nf, alpha, mk, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
    }

    // First two damped-Jacobi sweeps from a zero start in one pass:
    //   c1 = dinv * d                         (Jacobi MG overload (1), ElementProduct)
    //   c2 = c1 + (d - A c1) * dinv           (one sweep with rhs d)
    // with c1 recomputed at the neighbours instead of stored.  ACC selects
    // the output: 0 -> out = c2, 1 -> out = out + c2 (Sum<Tag>(x, c)).
    // V = 16 / sizeof(DT) rows per lane; bands and out must be 16-byte
    // aligned.  Lanes whose rows touch either end of the vector take the
    // guarded per-row form; every row sums in q1_row's order.
    // ORD as q1_row: 0 Product<Q1>'s order, 1 ascending-column (ELL) order.
    // With ACC the added-to values come from accin (out itself, or another
    // array: out = accin + c2).
    template <typename DT, int ACC, int ORD = 0>
    __device__ __forceinline__ void jacobi_zero2_body(const Bands & B, const DT * __restrict__ d,
            const DT * __restrict__ dinv, DT * out, int64_t n, int64_t m, int64_t lane, const DT * accin)
    {
        constexpr int V = Vec16<DT>::V;
        typedef typename Vec16<DT>::type VT;
        const int64_t i0 = lane * V;
        if (i0 >= n)
            return;
        const DT * const * b = reinterpret_cast<const DT * const *>(B.b);
        if (i0 - m - 1 < 0 || i0 + V + m >= n)
        {
            for (int v = 0; v < V && i0 + v < n; ++v)
            {
                const int64_t i = i0 + v;
#define HS_C1(j) (dinv[(j)] * d[(j)])
                DT s = ORD == 1 ? DT(0) : ld_stream(b[4] + i) * HS_C1(i);
                if (i - m - 1 >= 0) s += ld_stream(b[0] + i) * HS_C1(i - m - 1);
                if (i - m >= 0)     s += ld_stream(b[1] + i) * HS_C1(i - m);
                if (i - m + 1 >= 0) s += ld_stream(b[2] + i) * HS_C1(i - m + 1);
                if (i - 1 >= 0)     s += ld_stream(b[3] + i) * HS_C1(i - 1);
                if (ORD == 1)       s += ld_stream(b[4] + i) * HS_C1(i);
                if (i + 1 < n)      s += ld_stream(b[5] + i) * HS_C1(i + 1);
                if (i + m - 1 < n)  s += ld_stream(b[6] + i) * HS_C1(i + m - 1);
                if (i + m < n)      s += ld_stream(b[7] + i) * HS_C1(i + m);
                if (i + m + 1 < n)  s += ld_stream(b[8] + i) * HS_C1(i + m + 1);
#undef HS_C1
                const DT di = dinv[i];
                const DT c1 = di * d[i];
                const DT t = d[i] - s;
                const DT c2 = c1 + t * di;
                out[i] = ACC ? accin[i] + c2 : c2;
            }
            return;
        }
        VT bv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k)
            bv[k] = __builtin_nontemporal_load((const VT *)(b[k] + i0));
        DT cm[V + 2], cc[V + 2], cp[V + 2], dc[V + 2], ic[V + 2];
#pragma unroll
        for (int jj = 0; jj < V + 2; ++jj)
        {
            cm[jj] = dinv[i0 - m - 1 + jj] * d[i0 - m - 1 + jj];
            ic[jj] = dinv[i0 - 1 + jj];
            dc[jj] = d[i0 - 1 + jj];
            cc[jj] = ic[jj] * dc[jj];
            cp[jj] = dinv[i0 + m - 1 + jj] * d[i0 + m - 1 + jj];
        }
        VT o;
        if (ACC)
            o = *(const VT *)(accin + i0);
#pragma unroll
        for (int v = 0; v < V; ++v)
        {
            DT s = ORD == 1 ? DT(0) : bv[4][v] * cc[v + 1];
            s += bv[0][v] * cm[v];
            s += bv[1][v] * cm[v + 1];
            s += bv[2][v] * cm[v + 2];
            s += bv[3][v] * cc[v];
            if (ORD == 1)
                s += bv[4][v] * cc[v + 1];
            s += bv[5][v] * cc[v + 2];
            s += bv[6][v] * cp[v];
            s += bv[7][v] * cp[v + 1];
            s += bv[8][v] * cp[v + 2];
            const DT di = ic[v + 1];
            const DT c1 = di * dc[v + 1];
            const DT t = dc[v + 1] - s;
            const DT c2 = c1 + t * di;
            o[v] = ACC ? o[v] + c2 : c2;
        }
        *(VT *)(out + i0) = o;
    }

    template <typename DT, int ACC>
    __global__ __launch_bounds__(256) void jacobi_zero2_q1_v(Bands B, const DT * __restrict__ d,
            const DT * __restrict__ dinv, DT * __restrict__ out, int64_t n, int64_t m,
            const int * __restrict__ active)
    {
        if (active && *active == 0)
            return;
        jacobi_zero2_body<DT, ACC>(B, d, dinv, out, n, m, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, out);
    }

    // FillMatrix<Tag, POISSON, DIRICHLET_0>::value (honei/math/fill_matrix.hh:37-76)
    // on the nine bands of an M x M grid (N = M * M rows): boundary rows are
    // identity rows (DD 1, every other band 0), interior rows the Q1
    // stencil DD 8/3, off-diagonals -1/3 (each formed in double, stored as DT).
    template <typename DT>
    __global__ void fill_poisson_kernel(Bands B, int64_t N, int64_t M)
    {
        const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= N)
            return;
        const int64_t i = k + 1;
        const bool bnd = (i <= M || i > N - M || (i % M) == 1 || (i % M) == 0);
        const DT off = bnd ? DT(0.0) : DT(-1.0 / 3.0);
#pragma unroll
        for (int b = 0; b < 9; ++b)
            ((DT *)B.b[b])[k] = (b == 4) ? (bnd ? DT(1.0) : DT(8.0 / 3.0)) : off;
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

    // The threads that run a one-workgroup CG: NT = 0, the whole workgroup
    // (barriers); NT = 64, the first wave alone (the others wait elsewhere),
    // whose "barrier" waits for the wave's own memory operations to finish.
    template <int NT> struct Team
    {
        __device__ static int stride() { return NT ? NT : (int)blockDim.x; }
        __device__ static bool one_wave() { return NT == 64 || blockDim.x == 64; }
        __device__ static void sync()
        {
            if (NT)
                asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
            else
                __syncthreads();
        }
    };

    // v as held by lane j of the wave (j uniform)
    __device__ __forceinline__ float lane_value(float v, int j)
    {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
    }
    __device__ __forceinline__ double lane_value(double v, int j)
    {
        const long long u = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), j);
        const int hi = __builtin_amdgcn_readlane((int)(u >> 32), j);
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }

    // Sequential sum of a[i]*b[i], as DotProduct<tags::CPU>: by one thread,
    // or, on one wave, from the lanes' products (below).
    template <typename DT, int NT = 0>
    __device__ __forceinline__ void seq_dot(const DT * a, const DT * b, int64_t n, DT * dst)
    {
        typedef Team<NT> T;
        if (T::one_wave())
        {
            // one wave: the lanes form 64 products at once and the sum adds
            // them in index order, each read from its lane (the same
            // products and the same order of additions as the loop below,
            // without a memory round trip per term)
            // (terms past n, and lanes past the last term, add +0, which
            // leaves r unchanged: r starts at +0 and a sum that starts at +0
            // is never -0, the one value +0 would change)
            DT r = DT(0);
            for (int64_t c = 0; c < n; c += 64)
            {
                const int64_t i = c + threadIdx.x;
                const DT p = i < n ? a[i] * b[i] : DT(0);
                const int cnt = (int)(n - c < 64 ? n - c : 64);
                for (int j0 = 0; j0 < cnt; j0 += 8)
                {
                    DT q[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        q[u] = lane_value(p, j0 + u);
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        r += q[u];
                }
            }
            if (threadIdx.x == 0)
                *dst = r;
            T::sync();
            return;
        }
        if (threadIdx.x == 0)
        {
            DT r = DT(0);
            for (int64_t i = 0; i < n; ++i)
                r += a[i] * b[i];
            *dst = r;
        }
        T::sync();
    }

    template <typename DT> struct Eps;
    template <> struct Eps<float> { static constexpr float v = 1.19209290e-07F; };
    template <> struct Eps<double> { static constexpr double v = 2.2204460492503131e-16; };

    // A row of a Q1 banded matrix for the generic one-workgroup kernels.
    template <typename DT, int ORD = 0> struct Q1RowOp
    {
        Bands B;
        int64_t n, m;
        __device__ __forceinline__ DT row(const DT * __restrict__ x, int64_t i) const
        {
            return q1_row<DT, ORD>(B, x, i, n, m);
        }
    };

    // ConjugateGradients<Tag, NONE>::value(A, rhs, x, konv_rad) in one
    // workgroup (Q1: conjugate_gradients.hh:490-526, kernel :108-151; ELL:
    // :588-620, kernel :153-200 -- the same steps).  A.row(x, i) is row i of
    // A x.  Vectors g, u, energy, new_gradient live in `work` (4n);
    // reductions are the sequential CPU sums, so the iteration is
    // bit-identical to tags::CPU.
    template <typename DT, typename Op, int NT = 0>
    __device__ __forceinline__ void cg_konv_body_op(const Op & A, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, double konv_rad, uint64_t max_iters,
            uint64_t * iters_out);

    template <typename DT>
    __device__ __forceinline__ void cg_konv_body(const Bands & B, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, int64_t m, double konv_rad,
            uint64_t max_iters, uint64_t * iters_out)
    {
        Q1RowOp<DT> A{B, n, m};
        cg_konv_body_op<DT>(A, rhs, x, work, n, konv_rad, max_iters, iters_out);
    }

    // One step of the "version 1" cg_kernel (conjugate_gradients.hh:108-151
    // Q1, :153-200 ELL) on gradient g, iterate x and direction u, with the
    // work vectors en (energy) and ng (new gradient); sh holds at least 3
    // scalars.  Called by every thread of the team (Team<NT>).
    template <typename DT, typename Op, int NT = 0>
    __device__ __forceinline__ void cg_step_v1(const Op & A, const DT * __restrict__ rhs, DT * __restrict__ g,
            DT * __restrict__ x, DT * __restrict__ u, DT * __restrict__ en, DT * __restrict__ ng, int64_t n, DT * sh)
    {
        const DT eps = Eps<DT>::v;
        seq_dot<DT, NT>(g, g, n, &sh[0]);
        DT upper = sh[0];
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
            en[i] = A.row(u, i);
        Team<NT>::sync();
        seq_dot<DT, NT>(en, u, n, &sh[1]);
        DT lower = sh[1];
        const DT f1 = (fabs(lower) >= eps) ? DT(upper / lower) : DT(upper / eps);
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
        {
            const DT t = u[i] * f1;
            en[i] = t + x[i];
        }
        Team<NT>::sync();
        lower = upper;
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
        {
            const DT p = A.row(en, i);
            ng[i] = p - rhs[i];
        }
        Team<NT>::sync();
        seq_dot<DT, NT>(ng, ng, n, &sh[2]);
        upper = sh[2];
        const DT f2 = (fabs(lower) >= eps) ? DT(upper / lower) : DT(upper / eps);
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
        {
            const DT t = u[i] * f2;
            u[i] = t - ng[i];
            g[i] = ng[i];
            x[i] = en[i];
        }
        Team<NT>::sync();
    }

    // x = 0, g = A x - rhs, u = g * -1 (the set-up of both fixed forms)
    template <typename DT, typename Op, int NT = 0>
    __device__ __forceinline__ void cg_v1_start(const Op & A, const DT * __restrict__ rhs, DT * __restrict__ g,
            DT * __restrict__ x, DT * __restrict__ u, int64_t n)
    {
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
            x[i] = DT(0);
        Team<NT>::sync();
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
        {
            const DT p = A.row(x, i);
            g[i] = p - rhs[i];
        }
        Team<NT>::sync();
        for (int64_t i = threadIdx.x; i < n; i += Team<NT>::stride())
            u[i] = g[i] * DT(-1.);
    }

    template <typename DT, typename Op, int NT>
    __device__ __forceinline__ void cg_konv_body_op(const Op & A, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, double konv_rad, uint64_t max_iters,
            uint64_t * iters_out)
    {
        __shared__ DT sh[4];
        DT * g = work;
        DT * u = work + n;
        DT * en = work + 2 * n;
        DT * ng = work + 3 * n;
        cg_v1_start<DT, Op, NT>(A, rhs, g, x, u, n);
        seq_dot<DT, NT>(x, x, n, &sh[3]);
        DT norm_x = sh[3];
        DT norm_x_last = norm_x - DT(1);
        uint64_t it = 0;
        while ((double)(norm_x - norm_x_last) > konv_rad && it < max_iters)
        {
            cg_step_v1<DT, Op, NT>(A, rhs, g, x, u, en, ng, n, sh);
            ++it;
            norm_x_last = norm_x;
            seq_dot<DT, NT>(x, x, n, &sh[3]);
            norm_x = sh[3];
            Team<NT>::sync();
        }
        if (iters_out && threadIdx.x == 0)
            *iters_out = it;
    }

    // ConjugateGradients<Tag, NONE>::value(Q1, rhs, x, unsigned long iters)
    // (conjugate_gradients.hh:526-569) in one workgroup: x = 0, then up to
    // `iters` cg_kernel steps, each followed by the squared norm of the true
    // defect rhs - A x; stop when it is below 1e-8 times the squared norm of
    // the initial gradient (the product formed in double).  Vectors g, u,
    // energy, new_gradient and the defect live in `work` (5n); sequential
    // sums, so the iterates are bit-identical to tags::CPU.  *used_out = the
    // steps taken.
    template <typename DT, typename Op>
    __device__ __forceinline__ void cg_iters_body_op(const Op & A, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, uint64_t iters, uint64_t * used_out)
    {
        __shared__ DT sh[5];
        DT * g = work;
        DT * u = work + n;
        DT * en = work + 2 * n;
        DT * ng = work + 3 * n;
        DT * r = work + 4 * n;
        cg_v1_start<DT>(A, rhs, g, x, u, n);
        seq_dot<DT>(g, g, n, &sh[4]);
        const double initial = (double)sh[4];
        uint64_t i = 0;
        while (i < iters)
        {
            ++i;
            cg_step_v1<DT>(A, rhs, g, x, u, en, ng, n, sh);
            for (int64_t j = threadIdx.x; j < n; j += blockDim.x)
                r[j] = rhs[j] - A.row(x, j);
            __syncthreads();
            seq_dot<DT>(r, r, n, &sh[3]);
            const double current = (double)sh[3];
            __syncthreads();
            if (current < initial * 1e-08)
                break;
        }
        if (used_out && threadIdx.x == 0)
            *used_out = i;
    }

    template <typename DT>
    __global__ __launch_bounds__(256) void cg_iters_q1_kernel(Bands B, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, int64_t m, uint64_t iters, uint64_t * used_out)
    {
        Q1RowOp<DT> A{B, n, m};
        cg_iters_body_op<DT>(A, rhs, x, work, n, iters, used_out);
    }

    template <typename DT>
    __global__ __launch_bounds__(256) void cg_konv_q1_kernel(Bands B, const DT * __restrict__ rhs,
            DT * __restrict__ x, DT * __restrict__ work, int64_t n, int64_t m, double konv_rad,
            uint64_t max_iters, uint64_t * iters_out, const int * active)
    {
        if (active && *active == 0)
            return;
        cg_konv_body<DT>(B, rhs, x, work, n, m, konv_rad, max_iters, iters_out);
    }
}
}
