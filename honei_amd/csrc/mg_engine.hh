// mg_engine.hh -- the FP32-multigrid / FP64-iterative-refinement engine
// behind Multigrid<Tag, OuterTag, NONE, JAC, CYCLE::V, MIXED>
// (honei/math/multigrid.hh:1616-2393) and behind the request server.
//
// A *hierarchy* holds what every solve on it shares: the FP32 Q1 bands and
// inverted diagonals of each level and the FP64 bands of each finest level
// that is solved on.  A *slot* holds one solve's iterates: per level x, tmp,
// c, d, rhs (FP32), the FP64 outer defect, the coarse-CG work vectors and a
// block of device scalars (flags, counts, norms).
//
// The host advances slots in *ticks*.  One tick, for every slot of one
// finest level in the same launches (blockIdx.z = slot index), runs
//   [start a new solve | close an inner solve with the FP64 outer step]
//   -> one V-cycle -> the inner stopping test.
// Every kernel first reads its slot's flags and returns when they say the
// step does not apply, so the launch sequence of a tick is fixed (and can be
// captured as a graph) while each slot follows its own solve.  The
// operations applied to one slot, and their order, are exactly those of the
// reference's solve whatever the other slots do, so a solve gives the same
// bits alone and batched with others.
#pragma once

#include "common.hh"
#include "kernels_mg.hh"

#include <vector>

namespace hsolve
{
namespace mg
{
    constexpr int kMaxLevels = 15;          // levels 0..14 (side 2^L + 1)
    constexpr unsigned kNormBlocks = 1024;  // reduction blocks per slot

    enum Phase : int
    {
        PH_IDLE = 0,
        PH_START = 1,   // admitted, starts at the next tick
        PH_SOLVE = 2,
        PH_DONE = 3
    };

    // Per-slot device scalars.
    struct Scalars
    {
        int phase;
        int outer_active;    // 1 while the outer (FP64) loop runs
        int run;             // 1 while an inner solve runs
        int need_outer;      // the inner solve ended, outer step pending
        int do_start;        // gates of the current tick (plan_kernel)
        int do_outer;
        int open;            // an inner solve opens in this tick
        int status;          // 1: outer step cap reached without convergence
        int zero_guess;      // start from x = 0 instead of the result array
        float alpha;         // correction factor when it is adapted
        float inner_initial, inner_defect;
        uint64_t outer_iters, inner_total, inner_last, cg_iters;
        uint64_t ticks;      // ticks that found the solve started or running
        double initial_defect, def_norm, scale, inv;
    };

    enum Field : int
    {
        F_X = 0,
        F_TMP = 1,
        F_C = 2,
        F_D = 3,
        F_RHS = 4
    };

    // One slot's pointers (a device-resident table entry; s == nullptr marks
    // a free table index).
    struct Slot
    {
        Scalars * s;
        double * partials;       // 2 * kNormBlocks
        const double * b;        // finest right-hand side (FP64)
        double * result;         // finest solution (FP64); holds the initial guess until the start
        double * od;             // outer defect (FP64)
        float * cg_work;         // 4 * n(min_level)
        float * v[kMaxLevels][5];
    };

    // What the host reads back after a tick.
    struct Report
    {
        int phase, status;
        uint64_t outer_iters, inner_total, cg_iters, ticks;
        double initial_defect, def_norm;
    };

    struct Params
    {
        uint64_t min_level;
        uint64_t n_max_iter;
        int initial_zero, convergence_check;
        float tolerance;
        uint64_t n_pre_smooth;
        float adapt;
        k::Mask mask;
        uint64_t max_outer;
    };

    // An ELL matrix owned by a hierarchy: HONEI's one-lane slot order (slot k
    // of row r at r + k * stride), 32-bit absolute columns, 32-bit row lengths.
    template <typename DT> struct EllMat
    {
        const uint32_t * col = nullptr;
        const DT * val = nullptr;
        const uint32_t * len = nullptr;
        uint64_t stride = 0, rows = 0, cols = 0;
        uint32_t ncpr = 0;
    };
    typedef EllMat<float> EllMat32;
    typedef EllMat<double> EllMat64;

    struct HierLevel
    {
        uint64_t n = 0;
        uint64_t bs = 0;          // band stride (n rounded up to 64)
        int64_t side = 0;
        float * bands = nullptr;  // 9 * bs (Q1 hierarchies)
        EllMat32 ell;             // the level matrix (ELL hierarchies)
        EllMat32 prol, res;       // PROLMAT transfers to and from level - 1
        float * dinv = nullptr;
        // value codes of the nine bands and dinv (hierarchies on bands, when
        // every one of the ten arrays holds at most 16 distinct values): node i's
        // ten 4-bit table indices in one word, array k in bits 4k .. 4k + 3,
        // and the tables, ctab[16 k + t]; null otherwise
        uint64_t * codes = nullptr;
        float * ctab = nullptr;
        k::Bands B() const
        {
            k::Bands r;
            for (int i = 0; i < 9; ++i)
                r.b[i] = bands + (uint64_t)i * bs;
            return r;
        }
    };

    // Shared data of every solve on levels min_level..max_level.
    struct Hierarchy
    {
        Params p;
        int kind = 0;                        // 0: Q1 bands, 1: ELL matrices, 2: bands holding
                                             // 9-point structured ELL matrices (ELL sum order)
        int prolmat = 0;                     // 1: transfers by the prol/res matrices
        uint64_t max_level = 0;
        HierLevel lv[kMaxLevels];
        double * bands64[kMaxLevels] = {};   // FP64 bands of each finest level solved on
        uint64_t bs64[kMaxLevels] = {};
        uint64_t * codes64[kMaxLevels] = {};  // value codes of the FP64 bands (as HierLevel::codes,
        double * ctab64[kMaxLevels] = {};     // nine arrays), or null
        EllMat64 ell64[kMaxLevels];          // FP64 finest ELL matrix (ELL hierarchies)
        std::vector<void *> allocations;
        uint64_t bytes = 0;
        bool use_codes = true;               // read the value codes where a level has them

        void * alloc(size_t b);
        k::Bands B64(uint64_t L) const
        {
            k::Bands r;
            for (int i = 0; i < 9; ++i)
                r.b[i] = bands64[L] + (uint64_t)i * bs64[L];
            return r;
        }
        ~Hierarchy();
    };

    __host__ __device__ inline uint64_t band_stride(uint64_t n) { return (n + 63) & ~(uint64_t)63; }
    __host__ __device__ inline int64_t side_of(uint64_t L) { return (int64_t)(1ull << L) + 1; }
    __host__ __device__ inline uint64_t size_of(uint64_t L) { return (uint64_t)(side_of(L) * side_of(L)); }

    Params make_params(const hsolve_mg_params & p);
    // allocates bands and dinv for levels p.min_level..max_level
    void hierarchy_alloc(Hierarchy & h, const Params & p, uint64_t max_level);
    // allocates FP64 bands for finest level L
    void hierarchy_alloc64(Hierarchy & h, uint64_t L);
    // FillMatrix<POISSON, DIRICHLET_0> on every level and the FP64 levels present
    void hierarchy_fill_poisson(Hierarchy & h, hipStream_t st);
    // dinv = (float)((double)(1/dd) * 0.7) on level L
    void hierarchy_dinv(Hierarchy & h, uint64_t L, hipStream_t st);
    // value codes (HierLevel::codes) of every level on bands the one-pass
    // kernels run on, where the bands and dinv allow them, and of the FP64
    // finest bands of a Q1 hierarchy; call once they are final
    void hierarchy_codes(Hierarchy & h);

    // An ELL matrix in HONEI's host layout (threads == 1: slot k of row r
    // at r + k * stride; uint64 columns and row lengths).
    struct EllHost
    {
        uint64_t rows, cols, stride, ncpr;
        const uint64_t * Aj;
        const void * Ax;
        const uint64_t * Arl;
    };
    // ELL hierarchy: a32[L] for L in [min_level, max_level], A64 the FP64
    // finest matrix, dinv32[L] (or null: (float)((double)(1/dd) * 0.7)),
    // prol32 / res32 [L] for L in (min_level, max_level] or null for the
    // structured transfers (then every level must be (2^L+1)^2).
    void hierarchy_alloc_ell(Hierarchy & h, const Params & p, uint64_t max_level, const EllHost * a32,
                             const EllHost & A64, const float * const * dinv32, const EllHost * prol32,
                             const EllHost * res32);

    // Bytes of one slot whose finest level is L (own_b / own_result: the
    // slot also holds the right-hand side / the result array).
    uint64_t slot_bytes(const Hierarchy & h, uint64_t L, bool own_b, bool own_result);
    // Carves a slot out of `base` (slot_bytes long, 256-byte aligned).
    Slot slot_carve(const Hierarchy & h, uint64_t L, char * base, bool own_b, bool own_result);

    // Host image of a slot's scalars at admission.
    Scalars start_scalars(bool zero_guess);

    // Enqueues one tick for table entries [0, nslots) of finest level L.
    void enqueue_tick(const Hierarchy & h, uint64_t L, const Slot * table, unsigned nslots, hipStream_t st);
    // Copies every slot's report into `reports` (device, nslots entries).
    void enqueue_report(const Slot * table, unsigned nslots, Report * reports, hipStream_t st);
    // b = U[-1, 1] from SplitMix64(seed, stream) with zero Dirichlet nodes
    void enqueue_rhs_uniform(double * b, uint64_t L, uint64_t seed, uint64_t stream, hipStream_t st);
}  // namespace mg
}  // namespace hsolve
