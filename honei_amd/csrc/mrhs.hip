// mrhs.hip -- many right-hand sides on one ELL matrix.
//
// A *column-blocked multivector* of R <= 64 columns keeps row i's R entries
// contiguous at [i * R, i * R + R): a wavefront owns one row at a time with
// lane j on column j, so every gather of x[col] is one 8 R-byte coalesced
// load and every matrix entry is read once for all columns.
//
//   hsolve_ell_spmm    Y = A X or Y = RHS - A X (Product / Defect,
//                      product.hh:148-206, defect.hh:357-384, per column)
//   hsolve_cg_ell_multi CGSolver<Tag, NONE | VAR> (cg.hh:63-219) on every
//                      column at once: each column follows the reference's
//                      scalar recurrences and stopping test step for step
//                      (its own alpha, lambda, iteration count and
//                      convergence flag); converged columns stop loading.
//   hsolve_rhs_uniform_multi  seeded U[-1, 1] right-hand sides with zeros on
//                      the Dirichlet nodes of a macro border mask.
//
// Work split: rows are cut into tiles of 64; each block takes a contiguous
// range of tiles (in XCD-contiguous block order, so the rows +-side away
// that a Q1 row gathers stay in one XCD's L2) and its 4 waves take every
// 4th tile of the range.  A wave loads a tile's ELL slots with one
// coalesced load per slot (lane l on row base + l) and hands them to the
// row loop by readlane.  Per-column reductions accumulate per lane in
// double in the wave's row order, then over the 4 waves, then over the
// blocks in block order (a one-block finish launch), so results are
// deterministic.
#include "ell_handle.hh"
#include "kernels_mg.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hsolve
{
namespace
{
    constexpr int kMaxCols = 64;           // columns of a multivector (one per lane)
    constexpr int kTileK = 16;             // most ELL slots per row a tile keeps in registers
    constexpr unsigned kMvBlocks = 2048;   // grid cap of the multivector kernels
    constexpr int kFinThreads = 1024;      // threads of the one-block finish launches

    template <typename DT, typename IT, typename LT> struct MvEll
    {
        const IT * Aj;
        const DT * Ax;
        const LT * Arl;
        uint64_t stride;
        int64_t n;
        uint32_t ncpr;
        int64_t q1_side;       // hsolve_ell_s::q1_side
    };

    __device__ __forceinline__ uint32_t rlane(uint32_t v, int l)
    {
        return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
    }
    __device__ __forceinline__ float rlane(float v, int l)
    {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    }
    __device__ __forceinline__ double rlane(double v, int l)
    {
        const long long b = __double_as_longlong(v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
        return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }

    // The tiles of the calling wave: block range in XCD-contiguous order,
    // every 4th tile of it.
    struct TileRange
    {
        int64_t first, end, step;
    };
    __device__ __forceinline__ TileRange wave_tiles(int64_t tiles)
    {
        const uint32_t b = k::xcd_block(blockIdx.x, gridDim.x);
        const int64_t per = (tiles + gridDim.x - 1) / gridDim.x;
        const int64_t lo = (int64_t)b * per;
        const int64_t hi = lo + per < tiles ? lo + per : tiles;
        const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        return TileRange{lo + wid, hi, (int64_t)(blockDim.x >> 6)};
    }

    // Calls f(row, sum_j) for every row of the tile at `base`, sum_j = row
    // of A times column j of X (lane j), the row's slots summed from zero in
    // ascending slot order (product.hh:192-202), and own_j = OWN[row][j]
    // (0 when OWN is null).  K >= num_cols_per_row: the tile keeps K slots
    // per row in registers (unused slots point at the row itself), and each
    // row issues all K gathers and the OWN load before the first add, so a
    // row costs one memory round trip.
    template <int K, typename DT, typename IT, typename LT, typename F>
    __device__ __forceinline__ void tile_rows(const MvEll<DT, IT, LT> & A, const DT * __restrict__ X,
                                              const DT * __restrict__ OWN, unsigned R, int64_t base, bool on, F && f)
    {
        const int lane = threadIdx.x & 63;
        const int64_t myrow = base + lane;
        const int rows = (int)(A.n - base < 64 ? A.n - base : 64);
        uint32_t len = 0;
        if (myrow < A.n)
            len = (uint32_t)A.Arl[myrow];
        const uint32_t self = myrow < A.n ? (uint32_t)myrow : 0u;
        // every stored slot (k < num_cols_per_row) of the tile's rows in one
        // batch of coalesced loads, then the row lengths applied
        IT rc[K];
        DT ra[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
        {
            rc[kk] = IT(0);
            ra[kk] = DT(0);
            if ((uint32_t)kk < A.ncpr)
            {
                const uint64_t at = (uint64_t)kk * A.stride + (uint64_t)self;
                rc[kk] = k::ld_stream(A.Aj + at);
                ra[kk] = k::ld_stream(A.Ax + at);
            }
        }
        uint32_t c[K];
        DT a[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
        {
            const bool in = (uint32_t)kk < len;
            c[kk] = in ? (uint32_t)k::col_of<IT>(rc[kk], (uint64_t)myrow) : self;
            a[kk] = in ? ra[kk] : DT(0);
        }
        for (int r = 0; r < rows; ++r)
        {
            DT xv[K];
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
            {
                const uint32_t col = rlane(c[kk], r);
                xv[kk] = on ? X[(uint64_t)col * R + lane] : DT(0);
            }
            const DT ov = (on && OWN) ? OWN[(uint64_t)(base + r) * R + lane] : DT(0);
            const uint32_t lr = rlane(len, r);
            DT sum = DT(0);
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
            {
                if ((uint32_t)kk < lr)
                    sum += rlane(a[kk], r) * xv[kk];
            }
            f(base + r, sum, ov);
        }
    }

    // The same for Q1-structured matrices (every stored column of row
    // (y, x) is a neighbour (y + dy, x + dx), |dy|, |dx| <= 1, ascending, on
    // a side-M grid).  The tile's slots are first sorted into their 9
    // neighbour positions (value, presence bit) per lane.  The wave then
    // walks the rows in order keeping the 3 x 3 window of X rows around the
    // current row in registers: one step right along a grid line drops the
    // window's left column and gathers the 3 entries of its new right column,
    // so a row costs 3 gathers instead of 9; 4 rows of one grid line issue
    // their 12 gathers together.  Present entries are summed in
    // position order, which is the slot order, so the result is the same
    // bits as tile_rows.  Window entries off the grid are read at a clamped
    // (valid) address and never summed.
    template <typename DT, typename IT, typename LT, typename F>
    __device__ __forceinline__ void tile_rows_q1(const MvEll<DT, IT, LT> & A, const DT * __restrict__ X,
                                                 const DT * __restrict__ OWN, unsigned R, int64_t base, bool on,
                                                 F && f)
    {
        const int lane = threadIdx.x & 63;
        const int64_t n = A.n, M = A.q1_side;
        const int64_t myrow = base + lane;
        const int rows = (int)(n - base < 64 ? n - base : 64);
        const uint32_t self = myrow < n ? (uint32_t)myrow : 0u;
        const uint32_t len = myrow < n ? (uint32_t)A.Arl[myrow] : 0u;
        IT rc[9];
        DT ra[9];
#pragma unroll
        for (int kk = 0; kk < 9; ++kk)
        {
            rc[kk] = IT(0);
            ra[kk] = DT(0);
            if ((uint32_t)kk < A.ncpr)
            {
                const uint64_t at = (uint64_t)kk * A.stride + (uint64_t)self;
                rc[kk] = k::ld_stream(A.Aj + at);
                ra[kk] = k::ld_stream(A.Ax + at);
            }
        }
        DT ap[9];
        uint32_t mask = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q)
            ap[q] = DT(0);
#pragma unroll
        for (int kk = 0; kk < 9; ++kk)
        {
            if ((uint32_t)kk < len)
            {
                const int64_t d = (int64_t)k::col_of<IT>(rc[kk], (uint64_t)myrow) - myrow;
                const int p = d < -1 ? (int)(d + M) + 1 : (d > 1 ? (int)(d - M) + 7 : (int)d + 4);
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    ap[q] = (p == q) ? ra[kk] : ap[q];
                mask |= 1u << p;
            }
        }
        auto at = [&](int64_t c) { return X[(uint64_t)(c < 0 ? 0 : (c >= n ? n - 1 : c)) * R + lane]; };
        auto row_sum = [&](int r, const DT (&w)[9]) {
            const uint32_t mr = rlane(mask, r);
            DT sum = DT(0);
#pragma unroll
            for (int q = 0; q < 9; ++q)
            {
                if ((mr >> q) & 1u)
                    sum += rlane(ap[q], r) * w[q];
            }
            return sum;
        };
        // c0[b], c1[b]: window columns x - 1 and x of the next row, band b
        // (grid line y - 1, y, y + 1); valid after a row of the same line
        DT c0[3], c1[3];
        bool have = false;
        int64_t x = base % M;
        int r = 0;
        while (r < rows)
        {
            const int64_t i = base + r;
            if (x == 0 || ! have)
            {
#pragma unroll
                for (int b = 0; b < 3; ++b)
                {
                    c0[b] = on ? at(i + (b - 1) * M - 1) : DT(0);
                    c1[b] = on ? at(i + (b - 1) * M) : DT(0);
                }
            }
            if (r + 4 <= rows && x + 3 < M)
            {
                // 4 rows of one grid line: their 12 new gathers in flight at once
                DT v[3][4], ov[4];
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        v[b][u] = on ? at(i + (b - 1) * M + 1 + u) : DT(0);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    ov[u] = (on && OWN && OWN != X) ? OWN[(uint64_t)(i + u) * R + lane] : DT(0);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    DT w[9];
#pragma unroll
                    for (int b = 0; b < 3; ++b)
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                        {
                            const int k = u + d;       // column x - 1 + k of band b
                            w[b * 3 + d] = k == 0 ? c0[b] : (k == 1 ? c1[b] : v[b][k - 2]);
                        }
                    f(i + u, row_sum(r + u, w), OWN == X ? w[4] : ov[u]);
                }
#pragma unroll
                for (int b = 0; b < 3; ++b)
                {
                    c0[b] = v[b][2];
                    c1[b] = v[b][3];
                }
                r += 4;
                x += 4;
            }
            else
            {
                DT w[9];
#pragma unroll
                for (int b = 0; b < 3; ++b)
                {
                    w[b * 3] = c0[b];
                    w[b * 3 + 1] = c1[b];
                    w[b * 3 + 2] = on ? at(i + (b - 1) * M + 1) : DT(0);
                }
                const DT ov = OWN == X ? w[4] : ((on && OWN) ? OWN[(uint64_t)i * R + lane] : DT(0));
                f(i, row_sum(r, w), ov);
#pragma unroll
                for (int b = 0; b < 3; ++b)
                {
                    c0[b] = w[b * 3 + 1];
                    c1[b] = w[b * 3 + 2];
                }
                r += 1;
                x += 1;
            }
            if (x >= M)
                x -= M;
            have = true;
        }
    }

    // The same for matrices with more than kTileK slots per row: the row's
    // slots are read with wave-uniform addresses.
    template <typename DT, typename IT, typename LT, typename F>
    __device__ __forceinline__ void tile_rows_long(const MvEll<DT, IT, LT> & A, const DT * __restrict__ X,
                                                   const DT * __restrict__ OWN, unsigned R, int64_t base, bool on,
                                                   F && f)
    {
        const int lane = threadIdx.x & 63;
        const int rows = (int)(A.n - base < 64 ? A.n - base : 64);
        for (int r = 0; r < rows; ++r)
        {
            const int64_t row = base + r;
            const uint32_t len = (uint32_t)A.Arl[row];
            DT sum = DT(0);
            for (uint32_t kk = 0; kk < len; ++kk)
            {
                const uint64_t at = (uint64_t)kk * A.stride + (uint64_t)row;
                const uint64_t col = k::col_of<IT>(A.Aj[at], (uint64_t)row);
                const DT val = A.Ax[at];
                if (on)
                    sum += val * X[col * R + lane];
            }
            f(row, sum, (on && OWN) ? OWN[(uint64_t)row * R + lane] : DT(0));
        }
    }

    // K: slots per row kept in registers (9 or 16), 0 for longer rows, -1
    // for Q1-structured matrices
    template <typename DT, typename IT, typename LT, int K, typename F>
    __device__ __forceinline__ void for_rows_spmv(const MvEll<DT, IT, LT> & A, const DT * __restrict__ X,
                                                  const DT * __restrict__ OWN, unsigned R, bool on, F && f)
    {
        const TileRange t = wave_tiles((A.n + 63) / 64);
        for (int64_t tile = t.first; tile < t.end; tile += t.step)
        {
            if constexpr (K == 0)
                tile_rows_long(A, X, OWN, R, tile * 64, on, f);
            else if constexpr (K < 0)
                tile_rows_q1(A, X, OWN, R, tile * 64, on, f);
            else
                tile_rows<K>(A, X, OWN, R, tile * 64, on, f);
        }
    }

    // f(row0, cnt) for the wave's rows in groups of 4 (cnt < 4 only at the
    // end of the matrix), in ascending row order: the callers issue the loads
    // of a group before its first store.
    template <typename F>
    __device__ __forceinline__ void for_rows4(int64_t n, F && f)
    {
        const TileRange t = wave_tiles((n + 63) / 64);
        for (int64_t tile = t.first; tile < t.end; tile += t.step)
        {
            const int64_t base = tile * 64;
            const int rows = (int)(n - base < 64 ? n - base : 64);
            for (int r = 0; r < rows; r += 4)
                f(base + r, rows - r < 4 ? rows - r : 4);
        }
    }

    // Y = A X (MODE 0) or Y = RHS - A X (MODE 1), every column
    template <typename DT, typename IT, typename LT, int K, int MODE>
    __global__ __launch_bounds__(256) void spmm_kernel(MvEll<DT, IT, LT> A, const DT * __restrict__ X,
                                                       DT * __restrict__ Y, const DT * __restrict__ RHS, unsigned R)
    {
        const unsigned lane = threadIdx.x & 63;
        const bool on = lane < R;
        for_rows_spmv<DT, IT, LT, K>(A, X, MODE == 0 ? nullptr : RHS, R, on, [&](int64_t row, DT sum, DT rhs) {
            if (! on)
                return;
            const uint64_t e = (uint64_t)row * R + lane;
            Y[e] = MODE == 0 ? sum : rhs - sum;
        });
    }

    // ------------------------------------------------------------ CG state
    template <typename DT> struct MState
    {
        DT alpha[kMaxCols], alpha_old[kMaxCols], alpha_new[kMaxCols], lambda[kMaxCols], talpha[kMaxCols];
        DT initial_defect[kMaxCols], current_defect[kMaxCols];
        uint64_t iters[kMaxCols];
        int run[kMaxCols], converged[kMaxCols];
        DT eps_relative;
        uint64_t max_iters;
        int active;          // columns still iterating
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

    // Per-column reductions in two launches.  Each block stores, per column,
    // the sum of its 4 waves' lane sums in wave order (mput, no fence and no
    // atomic); a following one-block launch sums those stores per column
    // (msum: wave w over blocks w, w + W, ..., then the waves in order) and
    // applies the scalar step on lane j < R of its first wave.  The kernel
    // boundary orders the stores.
    template <int V>
    __device__ __forceinline__ void mput(const double (&acc)[V], double * partials)
    {
        __shared__ double lds[4][V][64];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int v = 0; v < V; ++v)
            lds[wid][v][lane] = acc[v];
        __syncthreads();
        if (wid == 0)
        {
#pragma unroll
            for (int v = 0; v < V; ++v)
            {
                double s = lds[0][v][lane];
                s += lds[1][v][lane];
                s += lds[2][v][lane];
                s += lds[3][v][lane];
                partials[((uint64_t)blockIdx.x * V + v) * 64 + lane] = s;
            }
        }
    }

    // in a one-block launch of kFinThreads threads (kFinThreads / 64 waves)
    template <int V, typename F>
    __device__ __forceinline__ void msum(const double * partials, unsigned nb, unsigned R, F && fin)
    {
        constexpr int W = kFinThreads / 64;
        __shared__ double lds[W][V][64];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        double t[V];
#pragma unroll
        for (int v = 0; v < V; ++v)
            t[v] = 0.0;
#pragma unroll 8
        for (unsigned b = (unsigned)wid; b < nb; b += W)
        {
#pragma unroll
            for (int v = 0; v < V; ++v)
                t[v] += partials[((uint64_t)b * V + v) * 64 + lane];
        }
#pragma unroll
        for (int v = 0; v < V; ++v)
            lds[wid][v][lane] = t[v];
        __syncthreads();
        if (wid == 0)
        {
            double tot[V];
#pragma unroll
            for (int v = 0; v < V; ++v)
            {
                double s = lds[0][v][lane];
                for (int w = 1; w < W; ++w)
                    s += lds[w][v][lane];
                tot[v] = s;
            }
            if ((unsigned)lane < R)
                fin(lane, tot);
        }
    }

    // R = B - A X; P = Pc * R (VAR) or R (NONE); per column r.r, r.p
    template <typename DT, typename IT, typename LT, int K, bool VAR>
    __global__ __launch_bounds__(256) void mcg_init_kernel(MvEll<DT, IT, LT> A, const DT * __restrict__ Bv,
            const DT * __restrict__ X, DT * __restrict__ Rv, DT * __restrict__ Pv, const DT * __restrict__ Pc,
            unsigned R, double * partials)
    {
        const unsigned lane = threadIdx.x & 63;
        const bool on = lane < R;
        double acc[2] = {0.0, 0.0};
        for_rows_spmv<DT, IT, LT, K>(A, X, Bv, R, on, [&](int64_t row, DT sum, DT bi) {
            if (! on)
                return;
            const uint64_t e = (uint64_t)row * R + lane;
            const DT ri = bi - sum;
            Rv[e] = ri;
            const DT pi = VAR ? Pc[row] * ri : ri;
            Pv[e] = pi;
            acc[0] += (double)ri * (double)ri;
            acc[1] += (double)ri * (double)pi;
        });
        mput<2>(acc, partials);
    }

    template <typename DT, bool VAR>
    __global__ __launch_bounds__(kFinThreads) void mcg_init_finish(const double * partials, unsigned nb, unsigned R,
            MState<DT> * s)
    {
        msum<2>(partials, nb, R, [&](int j, const double (&t)[2]) {
            s->iters[j] = 0;
            s->converged[j] = 0;
            s->run[j] = (s->max_iters > 0) ? 1 : 0;
            if (VAR)
            {
                s->initial_defect[j] = (DT)sqrt((DT)t[0]);
                s->alpha_new[j] = (DT)t[1];
            }
            else
            {
                s->alpha[j] = (DT)t[0];
                s->initial_defect[j] = (DT)sqrt((DT)t[0]);
            }
            s->current_defect[j] = s->initial_defect[j];
            if (j == 0)
                s->active = (s->max_iters > 0) ? (int)R : 0;
        });
    }

    // V = A P; per column temp = v.p and lambda
    template <typename DT, typename IT, typename LT, int K>
    __global__ __launch_bounds__(256) void mcg_spmv_kernel(MvEll<DT, IT, LT> A, const DT * __restrict__ Pv,
            DT * __restrict__ Vv, unsigned R, double * partials, const MState<DT> * s)
    {
        if (s->active == 0)
            return;
        const unsigned lane = threadIdx.x & 63;
        const bool on = lane < R && s->run[lane] != 0;
        double acc[1] = {0.0};
        for_rows_spmv<DT, IT, LT, K>(A, Pv, Pv, R, on, [&](int64_t row, DT sum, DT pi) {
            if (! on)
                return;
            Vv[(uint64_t)row * R + lane] = sum;
            acc[0] += (double)sum * (double)pi;
        });
        mput<1>(acc, partials);
    }

    template <typename DT, bool VAR>
    __global__ __launch_bounds__(kFinThreads) void mcg_spmv_finish(const double * partials, unsigned nb, unsigned R,
            MState<DT> * s)
    {
        if (s->active == 0)
            return;
        msum<1>(partials, nb, R, [&](int j, const double (&t)[1]) {
            if (s->run[j] == 0)
                return;
            const DT temp = (DT)t[0];
            if (VAR)
                s->lambda[j] = s->alpha_new[j] / guard_signed<DT>(temp);
            else
                s->lambda[j] = s->alpha[j] / guard<DT>(temp);
        });
    }

    // X += lambda P; R += (-lambda) V; per column r.r (and r.(Pc r)); stopping test
    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void mcg_update_kernel(DT * __restrict__ X, DT * __restrict__ Rv,
            const DT * __restrict__ Pv, const DT * __restrict__ Vv, const DT * __restrict__ Pc, int64_t n,
            unsigned R, double * partials, const MState<DT> * s)
    {
        if (s->active == 0)
            return;
        const unsigned lane = threadIdx.x & 63;
        const bool on = lane < R && s->run[lane] != 0;
        const DT lambda = on ? s->lambda[lane] : DT(0);
        const DT mlambda = -lambda;
        double acc[2] = {0.0, 0.0};
        if (on)
        {
            for_rows4(n, [&](int64_t row0, int cnt) {
                DT xv[4], pv[4], rv[4], vv[4], pc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    const int64_t row = row0 + (u < cnt ? u : 0);
                    const uint64_t e = (uint64_t)row * R + lane;
                    xv[u] = X[e];
                    pv[u] = Pv[e];
                    rv[u] = Rv[e];
                    vv[u] = Vv[e];
                    pc[u] = VAR ? Pc[row] : DT(0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    if (u < cnt)
                    {
                        const uint64_t e = (uint64_t)(row0 + u) * R + lane;
                        X[e] = xv[u] + lambda * pv[u];
                        const DT ri = rv[u] + mlambda * vv[u];
                        Rv[e] = ri;
                        acc[0] += (double)ri * (double)ri;
                        if (VAR)
                            acc[1] += (double)ri * (double)(pc[u] * ri);
                    }
                }
            });
        }
        mput<2>(acc, partials);
    }

    template <typename DT, bool VAR>
    __global__ __launch_bounds__(kFinThreads) void mcg_update_finish(const double * partials, unsigned nb, unsigned R,
            MState<DT> * s)
    {
        if (s->active == 0)
            return;
        msum<2>(partials, nb, R, [&](int j, const double (&t)[2]) {
            if (s->run[j] == 0)
                return;
            const DT eps = s->eps_relative;
            bool stop = false;
            if (VAR)
            {
                ++s->iters[j];
                const DT cur = (DT)sqrt((DT)t[0]);
                s->current_defect[j] = cur;
                if (cur < eps * s->initial_defect[j] || cur < eps)
                {
                    s->converged[j] = 1;
                    stop = true;
                }
                else if (s->iters[j] == s->max_iters)
                    stop = true;
                else
                {
                    s->alpha[j] = s->alpha_new[j];
                    s->alpha_new[j] = (DT)t[1];
                    s->talpha[j] = s->alpha_new[j] / guard_signed<DT>(s->alpha[j]);
                }
            }
            else
            {
                s->alpha_old[j] = s->alpha[j];
                s->alpha[j] = (DT)t[0];
                s->talpha[j] = s->alpha[j] / guard<DT>(s->alpha_old[j]);
                ++s->iters[j];
                const DT cur = (DT)sqrt(s->alpha[j]);
                s->current_defect[j] = cur;
                if (cur < eps * s->initial_defect[j] || cur < eps)
                {
                    s->converged[j] = 1;
                    stop = true;
                }
                else if (s->iters[j] == s->max_iters)
                    stop = true;
            }
            if (stop)
            {
                s->run[j] = 0;
                atomicSub(&s->active, 1);   // lanes of one wave, block-local
            }
        });
    }

    // P = P * talpha + Pc * R (VAR) or P * talpha + R (NONE)
    template <typename DT, bool VAR>
    __global__ __launch_bounds__(256) void mcg_direction_kernel(DT * __restrict__ Pv, const DT * __restrict__ Rv,
            const DT * __restrict__ Pc, int64_t n, unsigned R, const MState<DT> * s)
    {
        if (s->active == 0)
            return;
        const unsigned lane = threadIdx.x & 63;
        if (! (lane < R && s->run[lane] != 0))
            return;
        const DT talpha = s->talpha[lane];
        for_rows4(n, [&](int64_t row0, int cnt) {
            DT pv[4], rv[4], pc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                const int64_t row = row0 + (u < cnt ? u : 0);
                const uint64_t e = (uint64_t)row * R + lane;
                pv[u] = Pv[e];
                rv[u] = Rv[e];
                pc[u] = VAR ? Pc[row] : DT(0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                if (u < cnt)
                {
                    const DT t = pv[u] * talpha;
                    Pv[(uint64_t)(row0 + u) * R + lane] = t + (VAR ? pc[u] * rv[u] : rv[u]);
                }
            }
        });
    }

    unsigned mv_grid(uint64_t n)
    {
        const uint64_t tiles = (n + 63) / 64;
        return (unsigned)std::min<uint64_t>(kMvBlocks, std::max<uint64_t>(1, (tiles + 3) / 4));
    }

    template <typename F>
    void with_mv(const hsolve_ell_s * m, F && f)
    {
        require(m->threads == 1, HSOLVE_ERR_INVALID, "multivector kernels need an ELL handle with threads == 1");
        require(m->rows < (1ull << 32) && m->cols < (1ull << 32), HSOLVE_ERR_SIZE, "multivector: rows >= 2^32");
        with_ell_types(m, [&](auto Aj, auto Arl) {
            using IT = std::remove_const_t<std::remove_pointer_t<decltype(Aj)>>;
            using LT = std::remove_const_t<std::remove_pointer_t<decltype(Arl)>>;
            auto run = [&](auto dt_tag) {
                using DT = decltype(dt_tag);
                MvEll<DT, IT, LT> A{Aj, (const DT *)m->Ax, Arl, m->stride, (int64_t)m->rows, (uint32_t)m->ncpr,
                                    (int64_t)m->q1_side};
                if (m->q1_side != 0 && m->ncpr <= 9 && (m->variant & HSOLVE_ELL_SET) == 0)
                    f(A, std::integral_constant<int, -1>());
                else if (m->ncpr <= 9)
                    f(A, std::integral_constant<int, 9>());
                else if (m->ncpr <= (uint64_t)kTileK)
                    f(A, std::integral_constant<int, kTileK>());
                else
                    f(A, std::integral_constant<int, 0>());
            };
            if (m->dtype)
                run(double(0));
            else
                run(float(0));
        });
    }

    template <typename DT, typename IT, typename LT, int K, bool VAR>
    void run_mcg(const MvEll<DT, IT, LT> & A, unsigned R, const DT * Bv, DT * X, const DT * Pc, uint64_t max_iters,
                 double eps, uint64_t * used, double * defect, int * converged, hipStream_t st)
    {
        const uint64_t n = (uint64_t)A.n;
        const size_t vb = ((sizeof(DT) * n * R + 255) / 256) * 256;
        const unsigned grid = mv_grid(n);
        const size_t pb = sizeof(double) * 2 * 64 * grid;
        void * block = nullptr;
        if (hipMalloc(&block, 3 * vb + pb + sizeof(MState<DT>) + 256) != hipSuccess)
            throw Error(HSOLVE_ERR_ALLOC, "cg_ell_multi: workspace allocation failed");
        struct Free
        {
            void * p;
            ~Free() { (void)hipFree(p); }
        } guard_block{block};
        char * c = (char *)block;
        DT * Rv = (DT *)c; c += vb;
        DT * Pv = (DT *)c; c += vb;
        DT * Vv = (DT *)c; c += vb;
        double * partials = (double *)c; c += pb;
        MState<DT> * s = (MState<DT> *)c;

        MState<DT> h;
        std::memset(&h, 0, sizeof(h));
        h.eps_relative = (DT)eps;
        h.max_iters = max_iters;
        HIP_CHECK(hipMemcpyAsync(s, &h, sizeof(h), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((mcg_init_kernel<DT, IT, LT, K, VAR>), dim3(grid), dim3(kBlock), 0, st, A, Bv,
                           (const DT *)X, Rv, Pv, Pc, R, partials);
        hipLaunchKernelGGL((mcg_init_finish<DT, VAR>), dim3(1), dim3(kFinThreads), 0, st, (const double *)partials, grid,
                           R, s);
        HIP_CHECK(hipGetLastError());
        uint64_t enqueued = 0, chunk = 8;
        while (true)
        {
            for (uint64_t i = 0; i < chunk && enqueued < max_iters; ++i, ++enqueued)
            {
                hipLaunchKernelGGL((mcg_spmv_kernel<DT, IT, LT, K>), dim3(grid), dim3(kBlock), 0, st, A,
                                   (const DT *)Pv, Vv, R, partials, (const MState<DT> *)s);
                hipLaunchKernelGGL((mcg_spmv_finish<DT, VAR>), dim3(1), dim3(kFinThreads), 0, st,
                                   (const double *)partials, grid, R, s);
                hipLaunchKernelGGL((mcg_update_kernel<DT, VAR>), dim3(grid), dim3(kBlock), 0, st, X, Rv,
                                   (const DT *)Pv, (const DT *)Vv, Pc, (int64_t)n, R, partials,
                                   (const MState<DT> *)s);
                hipLaunchKernelGGL((mcg_update_finish<DT, VAR>), dim3(1), dim3(kFinThreads), 0, st,
                                   (const double *)partials, grid, R, s);
                hipLaunchKernelGGL((mcg_direction_kernel<DT, VAR>), dim3(grid), dim3(kBlock), 0, st, Pv,
                                   (const DT *)Rv, Pc, (int64_t)n, R, (const MState<DT> *)s);
            }
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(&h, s, sizeof(h), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            if (h.active == 0 || enqueued >= max_iters)
                break;
            if (chunk < 64)
                chunk *= 2;
        }
        for (unsigned j = 0; j < R; ++j)
        {
            if (used)
                used[j] = h.iters[j];
            if (defect)
                defect[j] = (h.initial_defect[j] > 0) ? (double)h.current_defect[j] / (double)h.initial_defect[j]
                                                     : (double)h.current_defect[j];
            if (converged)
                converged[j] = h.converged[j];
        }
    }

    // b[i * rs + j * cs] = U[-1, 1] from (seed + j, stream) counter i, 0 on
    // the Dirichlet nodes of the side-M grid
    __global__ __launch_bounds__(256) void rhs_multi_kernel(double * b, int64_t n, int M, uint64_t rs, uint64_t cs,
                                                            uint64_t seed, uint64_t stream_id, k::Mask mk)
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const uint64_t j = blockIdx.y;
        const int r = (int)(i / M), c = (int)(i - (int64_t)r * M);
        double v = 0.0;
        if (! k::dirichlet_rs(r, c, M, mk))
        {
            const uint64_t key = mix64(seed + j + kGolden * (stream_id + 1));
            const uint64_t bits = mix64(key + kGolden * ((uint64_t)i + 1));
            const double u = (double)(bits >> 11) * 0x1.0p-53;
            v = -1.0 + 2.0 * u;
        }
        b[(uint64_t)i * rs + j * cs] = v;
    }
}  // namespace
}  // namespace hsolve

using namespace hsolve;

extern "C" {

int hsolve_ell_spmm(hsolve_ell m, uint64_t nrhs, const void * X, void * Y, const void * RHS, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && X && Y, HSOLVE_ERR_INVALID, "ell_spmm: null pointer");
    require(nrhs >= 1 && nrhs <= (uint64_t)kMaxCols, HSOLVE_ERR_INVALID, "ell_spmm: nrhs in [1, 64]");
    hipStream_t st = as_stream(stream);
    const unsigned R = (unsigned)nrhs;
    with_mv(m, [&](auto A, auto k_tag) {
        using MV = decltype(A);
        using DT = std::remove_const_t<std::remove_pointer_t<decltype(MV::Ax)>>;
        constexpr int K = decltype(k_tag)::value;
        using IT = std::remove_const_t<std::remove_pointer_t<decltype(MV::Aj)>>;
        using LT = std::remove_const_t<std::remove_pointer_t<decltype(MV::Arl)>>;
        const unsigned grid = mv_grid(m->rows);
        if (RHS)
            hipLaunchKernelGGL((spmm_kernel<DT, IT, LT, K, 1>), dim3(grid), dim3(kBlock), 0, st, A,
                               (const DT *)X, (DT *)Y, (const DT *)RHS, R);
        else
            hipLaunchKernelGGL((spmm_kernel<DT, IT, LT, K, 0>), dim3(grid), dim3(kBlock), 0, st, A,
                               (const DT *)X, (DT *)Y, (const DT *)nullptr, R);
        HIP_CHECK(hipGetLastError());
    });
    HSOLVE_API_END
}

int hsolve_cg_ell_multi(hsolve_ell m, uint64_t nrhs, const void * P, const void * B, void * X, uint64_t max_iters,
                        double eps, uint64_t * used_iters, double * defect_out, void * stream)
{
    HSOLVE_API_BEGIN
    require(m && B && X, HSOLVE_ERR_INVALID, "cg_ell_multi: null pointer");
    require(m->rows == m->cols, HSOLVE_ERR_SIZE, "cg_ell_multi: matrix is not square");
    require(nrhs >= 1 && nrhs <= (uint64_t)kMaxCols, HSOLVE_ERR_INVALID, "cg_ell_multi: nrhs in [1, 64]");
    hipStream_t st = as_stream(stream);
    const unsigned R = (unsigned)nrhs;
    std::vector<int> conv(R, 0);
    with_mv(m, [&](auto A, auto k_tag) {
        using MV = decltype(A);
        using DT = std::remove_const_t<std::remove_pointer_t<decltype(MV::Ax)>>;
        constexpr int K = decltype(k_tag)::value;
        using IT = std::remove_const_t<std::remove_pointer_t<decltype(MV::Aj)>>;
        using LT = std::remove_const_t<std::remove_pointer_t<decltype(MV::Arl)>>;
        if (P)
            run_mcg<DT, IT, LT, K, true>(A, R, (const DT *)B, (DT *)X, (const DT *)P, max_iters, eps, used_iters,
                                            defect_out, conv.data(), st);
        else
            run_mcg<DT, IT, LT, K, false>(A, R, (const DT *)B, (DT *)X, nullptr, max_iters, eps, used_iters,
                                             defect_out, conv.data(), st);
    });
    if (max_iters > 0)
        for (unsigned j = 0; j < R; ++j)
            if (! conv[j])
            {
                set_last_error("cg_ell_multi: max_iters reached on column " + std::to_string(j));
                return HSOLVE_ERR_NOT_CONVERGED;
            }
    HSOLVE_API_END
}

int hsolve_rhs_uniform_multi(void * b, uint64_t side, uint64_t count, uint64_t row_stride, uint64_t col_stride,
                             const uint64_t * macro_border_mask, uint64_t seed, uint64_t stream_id, void * stream)
{
    HSOLVE_API_BEGIN
    require(b != nullptr, HSOLVE_ERR_INVALID, "rhs_uniform_multi: null vector");
    require(side >= 2 && side <= 65536, HSOLVE_ERR_SIZE, "rhs_uniform_multi: side in [2, 65536]");
    require(count >= 1 && count <= 65535, HSOLVE_ERR_INVALID, "rhs_uniform_multi: count in [1, 65535]");
    k::Mask mk;
    for (int i = 0; i < 8; ++i)
        mk.dir[i] = macro_border_mask ? (macro_border_mask[i] == 2) : 1;
    const uint64_t n = side * side;
    hipLaunchKernelGGL(rhs_multi_kernel, dim3(grid_for(n, kBlock), (unsigned)count), dim3(kBlock), 0,
                       as_stream(stream), (double *)b, (int64_t)n, (int)side, row_stride, col_stride, seed,
                       stream_id, mk);
    HIP_CHECK(hipGetLastError());
    HSOLVE_API_END
}

}  // extern "C"
