// honei/math/solvers.hh -- Defect, transfers, smoother, CG variants and
// mixed-precision multigrid for tags::GPU::MI355X, with HONEI's interfaces:
//   Defect<Tag>                       honei/math/defect.hh:42-768 (ELL, CSR, Q1)
//   ApplyDirichletBoundaries<Tag>     honei/math/apply_dirichlet_boundaries.hh:28-85
//   Restriction/Prolongation<Tag, M>  honei/math/restriction.hh:40-200, prolongation.hh:41-170
//   Jacobi<Tag> (MG overloads)        honei/math/jacobi.hh:398-523
//   CGSolver<Tag, NONE / VAR>         honei/math/cg.hh:63-219
//   ConjugateGradients<Tag, NONE>     honei/math/conjugate_gradients.hh:490-526 (Q1, konv_rad)
//   MGInfo, Multigrid<..., MIXED>     honei/math/multigrid.hh:65-149, 1616-2393
// The solvers keep every scalar on the device (include/hsolve.h); the
// iteration counts and stopping tests are the reference's.
#pragma once

#include <honei/la/operations.hh>
#include <honei/math/methods.hh>

#include <limits>
#include <memory>
#include <vector>

namespace honei
{
    template <typename Tag_> struct Defect;
    template <typename Tag_> struct ApplyDirichletBoundaries;
    template <typename Tag_, typename Type_> struct Restriction;
    template <typename Tag_, typename Type_> struct Prolongation;
    template <typename Tag_> struct Jacobi;
    template <typename Tag_, typename Method_> struct CGSolver;
    template <typename Tag_, typename Method_> struct ConjugateGradients;

    template <> struct Defect<tags::GPU::MI355X>
    {
        /// r = rhs - A x (ELL)
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & r, const DenseVector<DT> & rhs, const SparseMatrixELL<DT> & a,
                                       const DenseVector<DT> & x)
        {
            using namespace mi355x;
            if (x.size() != a.columns())
                throw VectorSizeDoesNotMatch(x.size(), a.columns());
            if (rhs.size() != a.rows())
                throw VectorSizeDoesNotMatch(rhs.size(), a.rows());
            same_size(r.size(), a.rows());
            if (x.memid() == r.memid())
                throw InternalError("Defect: result must not share storage with x");
            const DT * pr = rd(rhs);
            check(hsolve_ell_spmv(a.device_handle(), rd(x), r.memid() == rhs.memid() ? rw(r) : wr(r), pr, stream()),
                  "Defect ELL");
            return r;
        }

        /// r = rhs - A x (blocked CSR) on rows [row_start, row_end) (row_end
        /// 0: all rows); defect.hh:386-418, defect-cuda.cc:309-390, :660-725
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & r, const DenseVector<DT> & rhs, const SparseMatrixCSR<DT> & a,
                                       const DenseVector<DT> & x, unsigned long row_start = 0,
                                       unsigned long row_end = 0)
        {
            using namespace mi355x;
            if (x.size() != a.columns())
                throw VectorSizeDoesNotMatch(x.size(), a.columns());
            if (rhs.size() != a.rows())
                throw VectorSizeDoesNotMatch(rhs.size(), a.rows());
            same_size(r.size(), a.rows());
            if (x.memid() == r.memid())
                throw InternalError("Defect: result must not share storage with x");
            if (row_end == 0)
                row_end = a.rows();
            const bool all = row_start == 0 && row_end == a.rows();
            const DT * pr = rd(rhs);
            check(hsolve_csr_spmv(a.device_handle(), rd(x), (r.memid() == rhs.memid() || ! all) ? rw(r) : wr(r), pr,
                                  row_start, row_end, stream()),
                  "Defect CSR");
            return r;
        }

        /// r = rhs - A x (Q1 bands)
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & r, const DenseVector<DT> & rhs,
                                       const BandedMatrixQx<Q1Type, DT> & a, const DenseVector<DT> & x)
        {
            using namespace mi355x;
            same_size(x.size(), a.columns());
            same_size(rhs.size(), a.rows());
            same_size(r.size(), a.rows());
            if (x.memid() == r.memid())
                throw InternalError("Defect: result must not share storage with x");
            const DT * bands[9];
            a.device_bands(bands);
            const DT * pr = rd(rhs);
            check(abi::defect_q1(pr, bands, rd(x), r.memid() == rhs.memid() ? rw(r) : wr(r), a.size(), a.root()),
                  "Defect Q1");
            return r;
        }

        template <typename MT, typename DT>
        static DenseVector<DT> value(const DenseVector<DT> & rhs, const MT & a, const DenseVector<DT> & x)
        {
            DenseVector<DT> r(rhs.size());
            return value(r, rhs, a, x);
        }
    };

    namespace mi355x
    {
        inline std::vector<uint64_t> mask_of(const DenseVector<unsigned long> & mask)
        {
            if (mask.size() != 8)
                throw VectorSizeDoesNotMatch(mask.size(), 8);
            std::vector<uint64_t> m(8);
            for (int i = 0; i < 8; ++i)
                m[i] = mask[i];
            return m;
        }
    }

    template <> struct ApplyDirichletBoundaries<tags::GPU::MI355X>
    {
        template <typename DT> static DenseVector<DT> & value(DenseVector<DT> & grid, const DenseVector<unsigned long> & mask)
        {
            using namespace mi355x;
            const std::vector<uint64_t> m = mask_of(mask);
            int st = std::is_same<DT, double>::value
                         ? hsolve_apply_dirichlet_double(rw(grid), grid.size(), m.data(), stream())
                         : hsolve_apply_dirichlet_float(rw(grid), grid.size(), m.data(), stream());
            check(st, "ApplyDirichletBoundaries");
            return grid;
        }
    };

    template <> struct Restriction<tags::GPU::MI355X, methods::NONE>
    {
        /// full weighting fine -> coarse with Dirichlet masking; the matrix argument is unused
        template <typename DT, typename MatrixType_>
        static DenseVector<DT> & value(DenseVector<DT> & coarse, const DenseVector<DT> & fine,
                                       const DenseVector<unsigned long> & mask, const MatrixType_ &)
        {
            using namespace mi355x;
            const std::vector<uint64_t> m = mask_of(mask);
            check(abi::restriction(wr(coarse), coarse.size(), rd(fine), fine.size(), m.data()), "Restriction");
            return coarse;
        }
    };

    template <> struct Restriction<tags::GPU::MI355X, methods::PROLMAT>
    {
        /// coarse = resmat * fine (restriction.hh:174-185)
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & coarse, const DenseVector<DT> & fine,
                                       const DenseVector<unsigned long> &, const SparseMatrixELL<DT> & resmat)
        {
            return Product<tags::GPU::MI355X>::value(coarse, resmat, fine);
        }
    };

    template <> struct Prolongation<tags::GPU::MI355X, methods::NONE>
    {
        template <typename DT, typename MatrixType_>
        static DenseVector<DT> & value(DenseVector<DT> & fine, const DenseVector<DT> & coarse,
                                       const DenseVector<unsigned long> & mask, const MatrixType_ &)
        {
            using namespace mi355x;
            const std::vector<uint64_t> m = mask_of(mask);
            check(abi::prolongation(wr(fine), fine.size(), rd(coarse), coarse.size(), m.data()), "Prolongation");
            return fine;
        }
    };

    template <> struct Prolongation<tags::GPU::MI355X, methods::PROLMAT>
    {
        /// fine = prolmat * coarse (prolongation.hh:141-155)
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & fine, const DenseVector<DT> & coarse,
                                       const DenseVector<unsigned long> &, const SparseMatrixELL<DT> & prolmat)
        {
            return Product<tags::GPU::MI355X>::value(fine, prolmat, coarse);
        }
    };

    template <> struct Jacobi<tags::GPU::MI355X>
    {
        /// First sweep from zero: x = dinv * rhs (jacobi.hh:398-436)
        template <typename DT, typename MatrixType_>
        static void value(const MatrixType_ &, const DenseVector<DT> & rhs, DenseVector<DT> & x, DT,
                          const DenseVector<DT> & dinv)
        {
            ElementProduct<tags::GPU::MI355X>::value(x, dinv, rhs);
        }

        /// iter_number sweeps x_new = x + dinv * (rhs - A x) on to_smooth; afterwards
        /// x shares to_smooth's storage (jacobi.hh:438-480, 117-139)
        template <typename DT>
        static void value(DenseVector<DT> & to_smooth, const BandedMatrixQx<Q1Type, DT> & a,
                          const DenseVector<DT> & rhs, DenseVector<DT> & x, unsigned long iter_number, DT,
                          const DenseVector<DT> & dinv)
        {
            using namespace mi355x;
            same_size(rhs.size(), a.rows());
            same_size(to_smooth.size(), a.rows());
            const DT * bands[9];
            a.device_bands(bands);
            DenseVector<DT> tmp(to_smooth.size());
            DT * cur = rw(to_smooth);
            DT * alt = wr(tmp);
            for (unsigned long i(0) ; i < iter_number ; ++i)
            {
                check(abi::jacobi_q1(rd(rhs), bands, rd(dinv), cur, alt, a.size()), "Jacobi");
                std::swap(cur, alt);
            }
            if (iter_number % 2)
                check(hsolve_copy(cur, alt, sizeof(DT) * a.size(), stream()), "Jacobi copy back");
            x = to_smooth;
        }

        template <typename DT>
        static void value(DenseVector<DT> & to_smooth, const SparseMatrixELL<DT> & a, const DenseVector<DT> & rhs,
                          DenseVector<DT> & x, unsigned long iter_number, DT, const DenseVector<DT> & dinv)
        {
            using namespace mi355x;
            same_size(rhs.size(), a.rows());
            same_size(to_smooth.size(), a.rows());
            DenseVector<DT> tmp(to_smooth.size());
            DT * cur = rw(to_smooth);
            DT * alt = wr(tmp);
            for (unsigned long i(0) ; i < iter_number ; ++i)
            {
                check(hsolve_ell_jacobi(a.device_handle(), rd(rhs), rd(dinv), cur, alt, stream()), "Jacobi ELL");
                std::swap(cur, alt);
            }
            if (iter_number % 2)
                check(hsolve_copy(cur, alt, sizeof(DT) * a.rows(), stream()), "Jacobi copy back");
            x = to_smooth;
        }
    };

    namespace mi355x
    {
        template <typename DT>
        inline void cg_call(const BandedMatrixQx<Q1Type, DT> & A, const DenseVector<DT> & b, DenseVector<DT> & x,
                            unsigned long max_iters, unsigned long & used, double eps)
        {
            const DT * bands[9];
            A.device_bands(bands);
            uint64_t u = 0;
            double def = 0;
            int st = hsolve_cg_q1(Suffix<DT>::dtype, (const void * const *)bands, A.size(), rd(b), rw(x), max_iters, eps,
                                  &u, &def, stream());
            if (st != HSOLVE_ERR_NOT_CONVERGED)
                check(st, "CGSolver");
            used = u;
        }

        template <typename DT>
        inline void cg_call(const SparseMatrixELL<DT> & A, const DenseVector<DT> & b, DenseVector<DT> & x,
                            unsigned long max_iters, unsigned long & used, double eps)
        {
            uint64_t u = 0;
            double def = 0;
            int st = hsolve_cg_ell(A.device_handle(), rd(b), rw(x), max_iters, eps, &u, &def, stream());
            if (st != HSOLVE_ERR_NOT_CONVERGED)
                check(st, "CGSolver");
            used = u;
        }
    }

    template <> struct CGSolver<tags::GPU::MI355X, methods::NONE>
    {
        /// Unpreconditioned CG; stops at ||r|| < eps*||r0||, ||r|| < eps or max_iters (cg.hh:63-133)
        template <typename DT, typename MatrixType_, typename PreconContType_>
        static DenseVector<DT> & value(const MatrixType_ & A, const PreconContType_ &, const DenseVector<DT> & b,
                                       DenseVector<DT> & x, unsigned long max_iters, unsigned long & used_iters,
                                       DT eps_relative = 1e-8)
        {
            mi355x::same_size(b.size(), A.rows());
            mi355x::same_size(x.size(), A.columns());
            mi355x::cg_call(A, b, x, max_iters, used_iters, (double)eps_relative);
            return x;
        }
    };

    template <> struct CGSolver<tags::GPU::MI355X, methods::VAR>
    {
        /// CG preconditioned by a diagonal given as DenseVector P (z = P * r) (cg.hh:141-219)
        template <typename DT>
        static DenseVector<DT> & value(const SparseMatrixELL<DT> & A, const DenseVector<DT> & P,
                                       const DenseVector<DT> & b, DenseVector<DT> & x, unsigned long max_iters,
                                       unsigned long & used_iters, DT eps_relative = 1e-8)
        {
            using namespace mi355x;
            same_size(b.size(), A.rows());
            same_size(P.size(), A.rows());
            uint64_t u = 0;
            double def = 0;
            int st = hsolve_pcg_ell(A.device_handle(), rd(P), rd(b), rw(x), max_iters, (double)eps_relative, &u, &def,
                                    stream());
            if (st != HSOLVE_ERR_NOT_CONVERGED)
                check(st, "CGSolver VAR");
            used_iters = u;
            return x;
        }

        /// CG preconditioned by a sparse matrix P (z = Product(P, r)) (cg.hh:141-219),
        /// e.g. the SPAI / SAINV matrices PreconFill reads for the operator-list multigrid
        template <typename DT>
        static DenseVector<DT> & value(const SparseMatrixELL<DT> & A, const SparseMatrixELL<DT> & P,
                                       const DenseVector<DT> & b, DenseVector<DT> & x, unsigned long max_iters,
                                       unsigned long & used_iters, DT eps_relative = 1e-8)
        {
            using namespace mi355x;
            same_size(b.size(), A.rows());
            same_size(P.rows(), A.rows());
            same_size(P.columns(), A.rows());
            uint64_t u = 0;
            double def = 0;
            int st = hsolve_pcg_ell_pmat(A.device_handle(), P.device_handle(), rd(b), rw(x), max_iters,
                                         (double)eps_relative, &u, &def, stream());
            if (st != HSOLVE_ERR_NOT_CONVERGED)
                check(st, "CGSolver VAR");
            used_iters = u;
            return x;
        }
    };

    template <> struct ConjugateGradients<tags::GPU::MI355X, methods::NONE>
    {
        /// Coarse-grid CG: x = 0, iterate while ||x||^2 grows by more than konv_rad
        /// (conjugate_gradients.hh:490-526), one workgroup on the device
        template <typename DT>
        static void value(const BandedMatrixQx<Q1Type, DT> & A, const DenseVector<DT> & rhs, DenseVector<DT> & x,
                          double konv_rad)
        {
            using namespace mi355x;
            same_size(rhs.size(), A.rows());
            same_size(x.size(), A.rows());
            const DT * bands[9];
            A.device_bands(bands);
            check(abi::cg_konv_q1(bands, rd(rhs), wr(x), A.size(), konv_rad, 0), "ConjugateGradients");
        }

        /// The same on an ELL matrix (conjugate_gradients.hh:588-620, kernel :153-200)
        template <typename DT>
        static void value(const SparseMatrixELL<DT> & A, const DenseVector<DT> & rhs, DenseVector<DT> & x,
                          double konv_rad)
        {
            using namespace mi355x;
            same_size(rhs.size(), A.rows());
            same_size(x.size(), A.rows());
            check(hsolve_cg_konv_ell(A.device_handle(), rd(rhs), wr(x), konv_rad, 0, nullptr, stream()),
                  "ConjugateGradients ELL");
        }

        /// At most `iters` steps from x = 0, stopping once the squared true
        /// defect is below 1e-8 of its initial value (conjugate_gradients.hh:526-569);
        /// one workgroup with the sequential sums, so bit-identical to tags::CPU.
        /// Like the reference, reaching `iters` is not an error.
        template <typename DT>
        static void value(const BandedMatrixQx<Q1Type, DT> & A, const DenseVector<DT> & rhs, DenseVector<DT> & x,
                          unsigned long iters)
        {
            using namespace mi355x;
            same_size(rhs.size(), A.rows());
            same_size(x.size(), A.rows());
            const DT * bands[9];
            A.device_bands(bands);
            check(abi::cg_iters_q1(bands, rd(rhs), wr(x), A.size(), iters), "ConjugateGradients");
        }

        /// Plain CG from the given x for at most max_iters steps, stopping at
        /// ||r|| < 1e-8 ||r0|| or ||r|| < 1e-8 (conjugate_gradients.hh:571-586,
        /// kernel :203-248); reaching max_iters is not an error.
        template <typename DT>
        static void value(const SparseMatrixELL<DT> & A, const DenseVector<DT> & rhs, DenseVector<DT> & x,
                          unsigned long max_iters)
        {
            using namespace mi355x;
            same_size(rhs.size(), A.rows());
            same_size(x.size(), A.rows());
            uint64_t u = 0;
            double def = 0;
            check(hsolve_cg_iters_ell(A.device_handle(), rd(rhs), rw(x), max_iters, &u, &def, stream()),
                  "ConjugateGradients ELL");
        }
    };

    template <> struct ConjugateGradients<tags::GPU::MI355X, methods::JAC>
    {
        /// PCG with the inverted diagonal dd_inverted on an ELL matrix
        /// (conjugate_gradients.hh:1110-1155, kernel :885-913): stops at
        /// ||r|| < eps*||r0|| or after max_iters; used_iters is set by either.
        template <typename DT>
        static void value(const SparseMatrixELL<DT> & A, const DenseVector<DT> & rhs, DenseVector<DT> & x,
                          const DenseVector<DT> & dd_inverted, unsigned long max_iters, unsigned long & used_iters,
                          DT eps_relative = 1e-8)
        {
            using namespace mi355x;
            same_size(rhs.size(), A.rows());
            same_size(x.size(), A.rows());
            same_size(dd_inverted.size(), A.rows());
            uint64_t u = 0;
            double def = 0;
            int st = hsolve_pcg_jac_ell(A.device_handle(), rd(dd_inverted), rd(rhs), rw(x), max_iters,
                                        (double)eps_relative, &u, &def, stream());
            if (st != HSOLVE_ERR_NOT_CONVERGED)
                check(st, "ConjugateGradients JAC");
            used_iters = u;
        }
    };

    /// Configuration and level data of the multigrid solvers (multigrid.hh:65-149).
    template <typename Prec_, typename DataLayout_> struct MGInfo
    {
        bool is_smoother;
        DenseVector<unsigned long> macro_border_mask;
        unsigned long min_level;
        unsigned long max_level;
        unsigned long n_max_iter;
        bool initial_zero;
        Prec_ tolerance;
        bool convergence_check;
        unsigned long inner_iterations;
        unsigned long n_pre_smooth;
        unsigned long n_post_smooth;
        unsigned long n_max_iter_coarse;
        Prec_ tolerance_coarse;
        Prec_ adapt_correction_factor;
        std::vector<DenseVector<Prec_>> c, d, rhs, x, diags_inverted, temp_0, temp_1;
        std::vector<DataLayout_> a, prolmats, resmats, spais;
        /// outer iterations of the last mixed-precision solve (printed by the reference)
        unsigned long outer_iterations;
        /// device hierarchy built from a and diags_inverted on first use
        std::shared_ptr<void> device_hierarchy;

        MGInfo() :
            is_smoother(false), macro_border_mask(8, 2ul), min_level(1), max_level(1), n_max_iter(2),
            initial_zero(true), tolerance(Prec_(1e-8)), convergence_check(false), inner_iterations(0),
            n_pre_smooth(2), n_post_smooth(2), n_max_iter_coarse(0), tolerance_coarse(Prec_(1e-2)),
            adapt_correction_factor(Prec_(1)), outer_iterations(0)
        {
        }
    };

    template <typename Tag_, typename OuterTag_, typename ProlType_, typename SmootherType_, typename CycleType_,
              typename Mode_>
    struct Multigrid;

    namespace mg_kernel_detail
    {
        /// The engine's parameters from an MGInfo (no cap on the outer loop,
        /// as in the reference).
        template <typename Layout_> hsolve_mg_params engine_params(const MGInfo<float, Layout_> & info)
        {
            hsolve_mg_params p;
            p.min_level = info.min_level;
            p.max_level = info.max_level;
            p.n_max_iter = info.n_max_iter;
            p.initial_zero = info.initial_zero ? 1 : 0;
            p.convergence_check = info.convergence_check ? 1 : 0;
            p.tolerance = (double)info.tolerance;
            p.n_pre_smooth = info.n_pre_smooth;
            p.adapt_correction_factor = (double)info.adapt_correction_factor;
            for (int i = 0; i < 8; ++i)
                p.macro_border_mask[i] = info.macro_border_mask[i];
            p.max_outer = 0;
            return p;
        }

        // MIXED solve on ELL matrices: the device engine, or host-driven (multigrid_kernel.hh)
        template <typename ProlType_>
        void mixed_ell(const SparseMatrixELL<double> & system, const DenseVector<double> & rhs, DenseVector<double> & x,
                       MGInfo<float, SparseMatrixELL<float>> & info);
    }

    /// FP32 V-cycles inside FP64 iterative refinement on Q1 bands (multigrid.hh:1616-2393)
    /// and on ELL matrices (multigrid.hh:1936-2543) on the device-resident engine (ELL
    /// hierarchies with more than one lane per row are driven from the host).
    template <> struct Multigrid<tags::GPU::MI355X, tags::GPU::MI355X, methods::NONE, methods::JAC, methods::CYCLE::V,
                                 methods::MIXED>
    {
        static void value(const SparseMatrixELL<double> & system, const DenseVector<double> & right_hand_side,
                          DenseVector<double> & x, unsigned long /*max_levels*/, double /*conv_rad*/,
                          MGInfo<float, SparseMatrixELL<float>> & info)
        {
            mg_kernel_detail::mixed_ell<methods::NONE>(system, right_hand_side, x, info);
        }

        static void value(const BandedMatrixQx<Q1Type, double> & system, const DenseVector<double> & right_hand_side,
                          DenseVector<double> & x, unsigned long /*max_levels*/, double /*conv_rad*/,
                          MGInfo<float, BandedMatrixQx<Q1Type, float>> & info)
        {
            using namespace mi355x;
            same_size(right_hand_side.size(), system.rows());
            same_size(x.size(), system.rows());
            if (! info.device_hierarchy)
            {
                hsolve_mg_params p = mg_kernel_detail::engine_params(info);
                if (info.a.size() <= info.max_level || info.diags_inverted.size() <= info.max_level)
                    throw InternalError("Multigrid: MGInfo holds fewer levels than max_level");
                if (system.size() != info.a[info.max_level].size())
                    throw VectorSizeDoesNotMatch(system.size(), info.a[info.max_level].size());
                const double * b64[9];
                for (int k = 0; k < 9; ++k)
                    b64[k] = system.band(k).read_elements();
                std::vector<const float *> b32(9 * (info.max_level + 1), nullptr);
                std::vector<const float *> d32(info.max_level + 1, nullptr);
                for (unsigned long L = info.min_level; L <= info.max_level; ++L)
                {
                    for (int k = 0; k < 9; ++k)
                        b32[9 * L + k] = info.a[L].band(k).read_elements();
                    d32[L] = info.diags_inverted[L].read_elements();
                }
                hsolve_mg h = nullptr;
                check(hsolve_mg_create(&p, b64, b32.data(), d32.data(), &h), "Multigrid setup");
                info.device_hierarchy = std::shared_ptr<void>(h, [](void * q) { hsolve_mg_destroy((hsolve_mg)q); });
            }
            uint64_t outer = 0, inner = 0;
            double fd = 0, idf = 0;
            check(hsolve_mg_solve((hsolve_mg)info.device_hierarchy.get(), rd(right_hand_side), rw(x), &outer, &inner,
                                  &fd, &idf, stream()),
                  "Multigrid");
            info.outer_iterations = outer;
            info.inner_iterations = inner;
        }
    };
}

// the host-driven FIXED and ELL MIXED solvers
#include <honei/math/multigrid_kernel.hh>
