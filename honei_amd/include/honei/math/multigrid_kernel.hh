// honei/math/multigrid_kernel.hh -- Multigrid<…, JAC, CYCLE::V, FIXED> on Q1
// bands or ELL matrices in either precision, and Multigrid<…, MIXED> on ELL
// matrices (HONEI's honei/math/multigrid.hh:156-872 FIXED, 1936-2543 MIXED
// ELL).  MIXED on ELL runs on the device engine (hsolve_mg_create_ell) when
// every matrix has one lane per row; otherwise, and for FIXED, it is driven
// from the host: each step of the reference's _multigrid_kernel is one
// backend operation on device-resident data, and the host reads one norm per
// V-cycle.  (Multigrid<…, MIXED> on Q1 bands runs on the device engine, see
// solvers.hh.)
//
// Steps per V-cycle, as in the reference: per level down, c = dinv*d and
// n_pre_smooth-1 Jacobi sweeps (finest: x += c; coarser: x = c), defect,
// restriction into the coarser d and rhs; ConjugateGradients<NONE> (konv_rad
// = the precision's epsilon) on the coarsest level; per level up,
// prolongation, x += alpha*c, n_pre_smooth post-sweeps (the reference's
// post-smoother uses n_pre_smooth, multigrid.hh:1851), defect.
#pragma once

#include <honei/la/dense_vector.hh>
#include <honei/la/operations.hh>
#include <honei/math/methods.hh>
#include <honei/math/solvers.hh>

#include <cmath>
#include <limits>
#include <memory>
#include <type_traits>
#include <vector>

namespace honei
{
    namespace mg_kernel_detail
    {
        typedef tags::GPU::MI355X Tag;

        /// b's values on the first row, the last row and the first column of
        /// the square grid replace x's (multigrid.hh:2277-2298)
        template <typename DT> DenseVector<DT> inject_boundary(const DenseVector<DT> & x, const DenseVector<DT> & b)
        {
            DenseVector<DT> ig(x.copy());
            DT * g = ig.elements();
            const DT * r = b.read_elements();
            const unsigned long n = b.size();
            const unsigned long sq = (unsigned long)std::sqrt((double)n);
            for (unsigned long i(0) ; i < sq ; ++i)
                g[i] = r[i];
            for (unsigned long i(n - sq) ; i < n ; ++i)
                g[i] = r[i];
            for (unsigned long i(0) ; i < n ; i += sq)
                g[i] = r[i];
            return ig;
        }

        /// _multigrid_kernel: one inner solve on info.rhs[max_level] from
        /// info.x[max_level]; sets info.inner_iterations like the reference
        template <typename ProlType_, typename Prec_, typename Layout_>
        void kernel(MGInfo<Prec_, Layout_> & info)
        {
            const unsigned long maxl = info.max_level, minl = info.min_level;
            if (info.a.size() <= maxl || info.x.size() <= maxl || info.diags_inverted.size() <= maxl)
                throw InternalError("Multigrid: MGInfo holds fewer levels than max_level");
            if (info.initial_zero)
                copy<Tag>(info.rhs[maxl], info.d[maxl]);
            else
                Defect<Tag>::value(info.d[maxl], info.rhs[maxl], info.a[maxl], info.x[maxl]);
            Prec_ defect = info.is_smoother ? Prec_(1e8) : Norm<vnt_l_two, true, Tag>::value(info.d[maxl]);
            const Prec_ initial_defect = defect;
            if (info.convergence_check && defect <= info.tolerance)
                return;
            const Prec_ eps = std::numeric_limits<Prec_>::epsilon();
            unsigned long iter;
            for (iter = 1 ; iter <= info.n_max_iter ; ++iter)
            {
                unsigned long L = maxl;
                bool restriction_started = true;
                while (minl != maxl)
                {
                    if (restriction_started)
                    {
                        Jacobi<Tag>::value(info.a[L], info.d[L], info.c[L], Prec_(0.7), info.diags_inverted[L]);
                        Jacobi<Tag>::value(info.c[L], info.a[L], info.d[L], info.c[L], info.n_pre_smooth - 1,
                                           Prec_(0.7), info.diags_inverted[L]);
                        Sum<Tag>::value(info.x[L], info.c[L]);
                    }
                    else
                    {
                        Jacobi<Tag>::value(info.a[L], info.d[L], info.x[L], Prec_(0.7), info.diags_inverted[L]);
                        Jacobi<Tag>::value(info.x[L], info.a[L], info.d[L], info.x[L], info.n_pre_smooth - 1,
                                           Prec_(0.7), info.diags_inverted[L]);
                    }
                    Defect<Tag>::value(info.d[L], info.rhs[L], info.a[L], info.x[L]);
                    --L;
                    Restriction<Tag, ProlType_>::value(info.d[L], info.d[L + 1], info.macro_border_mask,
                                                       info.resmats[L + 1]);
                    copy<Tag>(info.d[L], info.rhs[L]);
                    if (L == minl)
                        break;
                    restriction_started = false;
                }
                ConjugateGradients<Tag, methods::NONE>::value(info.a[L], info.d[L], info.x[L], (double)eps);
                if (minl == maxl)
                {
                    Defect<Tag>::value(info.d[L], info.rhs[L], info.a[L], info.x[L]);
                    break;
                }
                while (true)
                {
                    ++L;
                    Prolongation<Tag, ProlType_>::value(info.c[L], info.x[L - 1], info.macro_border_mask,
                                                        info.prolmats[L]);
                    Prec_ alpha;
                    if (info.adapt_correction_factor == Prec_(0))
                    {
                        const Prec_ d1 = DotProduct<Tag>::value(info.d[L], info.c[L]);
                        DenseVector<Prec_> t(info.c[L].size(), Prec_(0));
                        Product<Tag>::value(t, info.a[L], info.c[L]);
                        const Prec_ d2 = DotProduct<Tag>::value(t, info.c[L]);
                        alpha = Prec_(d1 / d2);
                    }
                    else
                        alpha = info.adapt_correction_factor;
                    ScaledSum<Tag>::value(info.x[L], info.c[L], alpha);
                    Jacobi<Tag>::value(info.x[L], info.a[L], info.rhs[L], info.x[L], info.n_pre_smooth, Prec_(0.7),
                                       info.diags_inverted[L]);
                    Defect<Tag>::value(info.d[L], info.rhs[L], info.a[L], info.x[L]);
                    if (L == maxl)
                        break;
                }
                defect = Norm<vnt_l_two, true, Tag>::value(info.d[maxl]);
                if (defect <= initial_defect * info.tolerance)
                    break;
            }
            info.inner_iterations = (iter > info.n_max_iter) ? iter - 1 : iter;
        }

        /// Multigrid<..., FIXED>::value (multigrid.hh:790-872)
        template <typename ProlType_, typename Prec_, typename Layout_>
        void fixed(const Layout_ & system, const DenseVector<Prec_> & rhs, DenseVector<Prec_> & x,
                   MGInfo<Prec_, Layout_> & info)
        {
            mi355x::same_size(rhs.size(), system.rows());
            mi355x::same_size(x.size(), system.rows());
            info.x[info.max_level] = inject_boundary(x, rhs);
            kernel<ProlType_>(info);
            copy<Tag>(info.x[info.max_level], x);
        }

        /// The engine's view (hsolve_ell_host) of a one-lane-per-row SparseMatrixELL.
        template <typename DT> hsolve_ell_host ell_host(const SparseMatrixELL<DT> & A)
        {
            hsolve_ell_host e;
            e.rows = A.rows();
            e.cols = A.columns();
            e.stride = A.stride();
            e.num_cols_per_row = A.num_cols_per_row();
            e.Aj = (const uint64_t *)A.Aj().read_elements();
            e.Ax = A.Ax().read_elements();
            e.Arl = (const uint64_t *)A.Arl().read_elements();
            return e;
        }

        /// Does the device engine take this hierarchy: every matrix with one
        /// lane per row, and (structured transfers) level L of (2^L+1)^2 rows?
        template <typename ProlType_>
        bool engine_takes(const SparseMatrixELL<double> & system, const MGInfo<float, SparseMatrixELL<float>> & info)
        {
            const unsigned long minl = info.min_level, maxl = info.max_level;
            if (system.threads() != 1 || info.a.size() <= maxl || info.diags_inverted.size() <= maxl || minl < 1 ||
                minl > maxl || maxl > 14)
                return false;
            const bool prolmat = ! std::is_same<ProlType_, methods::NONE>::value;
            if (prolmat && (info.prolmats.size() <= maxl || info.resmats.size() <= maxl))
                return false;
            for (unsigned long L = minl; L <= maxl; ++L)
            {
                if (info.a[L].threads() != 1 || info.diags_inverted[L].size() != info.a[L].rows())
                    return false;
                const unsigned long side = (1ul << L) + 1;
                if (! prolmat && info.a[L].rows() != side * side)
                    return false;
                if (prolmat && L > minl && (info.prolmats[L].threads() != 1 || info.resmats[L].threads() != 1))
                    return false;
            }
            return info.a[maxl].rows() == system.rows();
        }

        /// Multigrid<..., MIXED>::value on ELL matrices on the device engine:
        /// the hierarchy (level matrices, inverted diagonals, transfer
        /// matrices, the FP64 finest matrix) is uploaded on first use and kept
        /// in info.device_hierarchy; each solve is then device resident, the
        /// host waiting once per chunk of ticks.
        template <typename ProlType_>
        void mixed_ell_engine(const SparseMatrixELL<double> & system, const DenseVector<double> & rhs,
                              DenseVector<double> & x, MGInfo<float, SparseMatrixELL<float>> & info)
        {
            using namespace mi355x;
            const unsigned long minl = info.min_level, maxl = info.max_level;
            if (! info.device_hierarchy)
            {
                hsolve_mg_params p = engine_params(info);
                const bool prolmat = ! std::is_same<ProlType_, methods::NONE>::value;
                std::vector<hsolve_ell_host> a32(maxl + 1), prol(maxl + 1), res(maxl + 1);
                std::vector<const float *> d32(maxl + 1, nullptr);
                for (unsigned long L = minl; L <= maxl; ++L)
                {
                    a32[L] = ell_host(info.a[L]);
                    d32[L] = info.diags_inverted[L].read_elements();
                    if (prolmat && L > minl)
                    {
                        prol[L] = ell_host(info.prolmats[L]);
                        res[L] = ell_host(info.resmats[L]);
                    }
                }
                const hsolve_ell_host a64 = ell_host(system);
                hsolve_mg h = nullptr;
                check(hsolve_mg_create_ell(&p, &a64, a32.data(), d32.data(), prolmat ? prol.data() : nullptr,
                                           prolmat ? res.data() : nullptr, 1, &h),
                      "Multigrid setup");
                info.device_hierarchy = std::shared_ptr<void>(h, [](void * q) { hsolve_mg_destroy((hsolve_mg)q); });
            }
            uint64_t outer = 0, inner = 0;
            double fd = 0, idf = 0;
            check(hsolve_mg_solve((hsolve_mg)info.device_hierarchy.get(), rd(rhs), rw(x), &outer, &inner, &fd, &idf,
                                  stream()),
                  "Multigrid");
            info.outer_iterations = outer;
            info.inner_iterations = inner;
        }

        /// Multigrid<..., MIXED>::value on ELL matrices (multigrid.hh:2400-2543):
        /// on the device engine when it takes the hierarchy, else driven from
        /// the host one backend operation per step
        template <typename ProlType_>
        void mixed_ell(const SparseMatrixELL<double> & system, const DenseVector<double> & rhs, DenseVector<double> & x,
                       MGInfo<float, SparseMatrixELL<float>> & info)
        {
            mi355x::same_size(rhs.size(), system.rows());
            mi355x::same_size(x.size(), system.rows());
            if (engine_takes<ProlType_>(system, info))
            {
                mixed_ell_engine<ProlType_>(system, rhs, x, info);
                return;
            }
            const unsigned long n = rhs.size(), maxl = info.max_level;
            DenseVector<double> initial_guess(inject_boundary(x, rhs));
            convert<Tag>(info.x[maxl], initial_guess);
            DenseVector<double> result(n, 0.0), product(n, 0.0), outer_defect(n), x_outer(n);
            Product<Tag>::value(product, system, initial_guess);
            copy<Tag>(rhs, outer_defect);
            Difference<Tag>::value(outer_defect, product);
            const double initial_defect = Norm<vnt_l_two, true, Tag>::value(outer_defect);
            const double outer_eps = (double)info.tolerance;
            double scale_factor(1.0);
            unsigned long inner_total(0), outer(0);
            while (true)
            {
                convert<Tag>(info.rhs[maxl], outer_defect);
                kernel<ProlType_>(info);
                ++outer;
                inner_total += info.inner_iterations;
                convert<Tag>(x_outer, info.x[maxl]);
                ScaledSum<Tag>::value(result, x_outer, scale_factor);
                Defect<Tag>::value(outer_defect, rhs, system, result);
                const double def_norm = Norm<vnt_l_two, true, Tag>::value(outer_defect);
                if (def_norm < outer_eps * initial_defect)
                    break;
                scale_factor = def_norm;
                Scale<Tag>::value(outer_defect, 1.0 / scale_factor);
            }
            copy<Tag>(result, x);
            info.outer_iterations = outer;
            info.inner_iterations = inner_total;
        }
    }

    /// Multigrid<..., JAC, CYCLE::V, FIXED> on Q1 bands or ELL, FP32 or FP64
    template <typename ProlType_>
    struct Multigrid<tags::GPU::MI355X, tags::GPU::MI355X, ProlType_, methods::JAC, methods::CYCLE::V, methods::FIXED>
    {
        template <typename Prec_, typename Layout_>
        static void value(const Layout_ & system, const DenseVector<Prec_> & right_hand_side, DenseVector<Prec_> & x,
                          unsigned long /*max_levels*/, Prec_ /*conv_rad*/, MGInfo<Prec_, Layout_> & info)
        {
            mg_kernel_detail::fixed<ProlType_>(system, right_hand_side, x, info);
        }
    };

    /// Multigrid<..., PROLMAT, JAC, CYCLE::V, MIXED> on ELL matrices
    template <>
    struct Multigrid<tags::GPU::MI355X, tags::GPU::MI355X, methods::PROLMAT, methods::JAC, methods::CYCLE::V,
                     methods::MIXED>
    {
        static void value(const SparseMatrixELL<double> & system, const DenseVector<double> & right_hand_side,
                          DenseVector<double> & x, unsigned long /*max_levels*/, double /*conv_rad*/,
                          MGInfo<float, SparseMatrixELL<float>> & info)
        {
            mg_kernel_detail::mixed_ell<methods::PROLMAT>(system, right_hand_side, x, info);
        }
    };
}
