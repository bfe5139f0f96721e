// honei/math/mg.hh -- operator-list multigrid (HONEI's honei/math/mg.hh):
// MGData (level matrices, transfer matrices, preconditioners and vectors),
// MGUtil (file loader and configuration), MGCycleCreation (builds the V, W
// or F cycle as an OperatorList), MGSolver and MGSmoother.
//
// The templates are generic over Tag_ like the reference's; with
// tags::GPU::MI355X every operator runs on the device through the backend
// (SparseMatrixELL handles upload once; DenseVector data stays resident),
// and the host only reads one norm per cycle.  Names, members and
// constructor arguments are kept alike so that HONEI callers compile
// unchanged.  Level indices: matrices and vectors at internal index
// level - 1, transfer matrices at level - 2 (MGDataIndex).
//
// Left out: the CYCLE::F::W shape (the reference's specialisation refers to
// members MGData does not have and cannot be instantiated) and the "spai"
// preconditioner built by SPAI2 (preconditioner matrices read from files,
// PreconFill's matrix form, are supported).
#pragma once

#include <honei/la/dense_vector.hh>
#include <honei/la/operations.hh>
#include <honei/la/sparse_matrix.hh>
#include <honei/la/sparse_matrix_ell.hh>
#include <honei/math/io.hh>
#include <honei/math/methods.hh>
#include <honei/math/operator.hh>
#include <honei/math/operator_list.hh>
#include <honei/math/ri.hh>
#include <honei/math/solvers.hh>
#include <honei/math/transposition.hh>
#include <honei/math/vectorpool.hh>
#include <honei/util/exception.hh>

#include <iostream>
#include <string>
#include <vector>

namespace honei
{
    struct MGDataIndex
    {
        static unsigned long internal_index_A(unsigned long base_index) { return base_index - 1; }
        static unsigned long base_index_A(unsigned long internal_index) { return internal_index + 1; }
        static unsigned long internal_index_Prol(unsigned long base_index) { return base_index - 2; }
        static unsigned long base_index_Prol(unsigned long internal_index) { return internal_index + 2; }
    };

    template <typename MatrixType_, typename VectorType_, typename TransferContType_, typename PreconContType_,
              typename DataType_>
    struct MGData
    {
        std::vector<MatrixType_> A;
        std::vector<TransferContType_> resmat;
        std::vector<TransferContType_> prolmat;
        std::vector<PreconContType_> P;
        std::vector<VectorType_> b, x, d, c, store;
        std::vector<std::vector<VectorType_>> smoother_temp;
        unsigned long max_iters;
        unsigned long max_iters_coarse;
        unsigned long used_iters_coarse;
        unsigned long used_iters;
        unsigned long n_pre_smooth;
        unsigned long n_post_smooth;
        unsigned long min_level;
        DataType_ eps_relative;

        MGData(std::vector<MatrixType_> & systems, std::vector<TransferContType_> & resmats,
               std::vector<TransferContType_> & prolmats, std::vector<PreconContType_> & precons,
               std::vector<VectorType_> & rhss, std::vector<VectorType_> & xs, std::vector<VectorType_> & ds,
               std::vector<VectorType_> & cs, std::vector<VectorType_> & stores,
               std::vector<std::vector<VectorType_>> & smoother_temps, unsigned long p_max_iters,
               unsigned long p_max_iters_coarse, unsigned long p_n_pre_smooth, unsigned long p_n_post_smooth,
               unsigned long p_min_level, DataType_ p_eps_rel) :
            A(systems), resmat(resmats), prolmat(prolmats), P(precons), b(rhss), x(xs), d(ds), c(cs), store(stores),
            smoother_temp(smoother_temps), max_iters(p_max_iters), max_iters_coarse(p_max_iters_coarse),
            used_iters_coarse(0), used_iters(0), n_pre_smooth(p_n_pre_smooth), n_post_smooth(p_n_post_smooth),
            min_level(p_min_level), eps_relative(p_eps_rel)
        {
            if (p_min_level > systems.size())
                throw InternalError("MGData: minimum level is larger than the number of levels");
        }

        /// deep copy of every matrix and vector
        MGData copy()
        {
            std::vector<MatrixType_> Ar;
            std::vector<TransferContType_> rr, pr;
            std::vector<PreconContType_> Pr;
            std::vector<VectorType_> br, xr, dr, cr, sr;
            std::vector<std::vector<VectorType_>> tr;
            for (auto & m : A) Ar.push_back(m.copy());
            for (auto & m : resmat) rr.push_back(m.copy());
            for (auto & m : prolmat) pr.push_back(m.copy());
            for (auto & m : P) Pr.push_back(m.copy());
            for (unsigned long i(0) ; i < b.size() ; ++i)
            {
                br.push_back(b[i].copy());
                xr.push_back(x[i].copy());
                dr.push_back(d[i].copy());
                cr.push_back(c[i].copy());
                sr.push_back(store[i].copy());
            }
            for (auto & v : smoother_temp)
            {
                std::vector<VectorType_> t;
                for (auto & e : v)
                    t.push_back(e.copy());
                tr.push_back(t);
            }
            MGData r(Ar, rr, pr, Pr, br, xr, dr, cr, sr, tr, max_iters, max_iters_coarse, n_pre_smooth,
                     n_post_smooth, min_level, eps_relative);
            r.used_iters = used_iters;
            return r;
        }
    };

    /// Builds level i's preconditioner (mg.hh:247-335).
    template <typename PreconContType_, typename MatrixType_, typename DataType_> struct PreconFill;

    template <typename MatrixType_, typename DataType_> struct PreconFill<DenseVector<DataType_>, MatrixType_, DataType_>
    {
        /// "jac": P = damping / diag(A), in DataType_
        static void value(unsigned long, std::vector<DenseVector<DataType_>> & target, std::string,
                          std::string precon_suffix, MatrixType_ & A, DataType_ damping_factor)
        {
            if (precon_suffix != "jac")
                throw InternalError("Preconditioner " + precon_suffix + " is not available on this backend");
            DenseVector<DataType_> P(A.rows(), damping_factor);
            DataType_ * p = P.elements();
            for (unsigned long j(0) ; j < P.size() ; ++j)
                p[j] /= A(j, j);
            target.push_back(P);
        }
    };

    template <typename MatrixType_, typename DataType_> struct PreconFill<MatrixType_, MatrixType_, DataType_>
    {
        /// the matrix in <filename><i+1>_<suffix>.ell, scaled by the damping
        /// factor through a SparseMatrix when that is not 1
        static void value(unsigned long i, std::vector<MatrixType_> & target, std::string filename,
                          std::string precon_suffix, MatrixType_ &, DataType_ damping_factor)
        {
            const std::string name = filename + std::to_string(i + 1) + "_" + precon_suffix + ".ell";
            MatrixType_ P(MatrixIO<io_formats::ELL>::read_matrix(name, DataType_(0)));
            if (damping_factor == DataType_(1))
            {
                target.push_back(P);
                return;
            }
            // SparseMatrix(P) (every slot, zeros skipped), Scale, SparseMatrixELL (mg.hh:336-339)
            SparseMatrix<DataType_> sm(P.rows(), P.columns());
            const unsigned long * aj = P.Aj().read_elements();
            const DataType_ * ax = P.Ax().read_elements();
            const unsigned long size = P.Aj().size();
            for (unsigned long row(0) ; row < P.rows() ; ++row)
                for (unsigned long pos(row * P.threads()) ; pos < size ; pos += P.stride())
                    for (unsigned long t(0) ; t < P.threads() ; ++t)
                        if (ax[pos + t] != DataType_(0))
                            sm(row, aj[pos + t], ax[pos + t] * damping_factor);
            target.push_back(MatrixType_(sm));
        }
    };

    /// Prints the level graph of a cycle (R restriction, P prolongation,
    /// S smoother, C coarse solver), as mg.hh:337-409.
    inline void print_cycle(OperatorList & ol, unsigned long max_level, unsigned long min_level)
    {
        std::vector<unsigned long> transfers;
        for (unsigned long i(0) ; i < ol.size() ; ++i)
            if (ol[i]->transfer_type() > 0)
                transfers.push_back(ol[i]->transfer_type());
        const unsigned long rows = (max_level - min_level + 1) * 2;
        std::vector<std::vector<unsigned long>> graph(rows, std::vector<unsigned long>(transfers.size(), 0));
        unsigned long at(0);
        for (unsigned long i(0) ; i < transfers.size() ; ++i)
        {
            if (transfers[i] == 1)
                graph[++at][i] = 1, ++at;
            else if (transfers[i] == 2)
                graph[--at][i] = 2, --at;
            else
                graph[at][i] = transfers[i];
        }
        static const char sym[] = {' ', 'R', 'P', 'S', 'C'};
        std::cout << "R: Restriction, P: Prolongation, S: Smoother, C: Solver" << std::endl << std::endl;
        for (auto & row : graph)
        {
            for (unsigned long v : row)
                std::cout << sym[v];
            std::cout << std::endl;
        }
    }

    template <typename Tag_, typename MatrixType_, typename VectorType_, typename TransferContType_,
              typename PreconContType_, typename MatIOType_, typename VecIOType_, typename DataType_>
    struct MGUtil
    {
        typedef MGData<MatrixType_, VectorType_, TransferContType_, PreconContType_, DataType_> Data;

        static void configure(Data & target, unsigned long max_iters, unsigned long max_iters_coarse,
                              unsigned long n_pre_smooth, unsigned long n_post_smooth, unsigned long min_level,
                              DataType_ eps_relative)
        {
            target.max_iters = max_iters;
            target.max_iters_coarse = max_iters_coarse;
            target.n_pre_smooth = n_pre_smooth;
            target.n_post_smooth = n_post_smooth;
            target.min_level = min_level;
            target.eps_relative = eps_relative;
        }

        /// Levels 1..max_level from <file_base>A_<l>.ell, prolongations from
        /// prol_<l>.ell (l >= 2) with restrictions as their transposes, the
        /// finest right-hand side and initial guess from rhs_<max> and
        /// init_<max>; every other vector starts at zero (mg.hh:449-545).
        static Data load_data(std::string file_base, unsigned long max_level, DataType_ damping_factor,
                              std::string precon_suffix)
        {
            std::vector<MatrixType_> A;
            std::vector<TransferContType_> Prol, Res;
            std::vector<PreconContType_> P;
            std::vector<VectorType_> b, x, d, c, store;
            std::vector<std::vector<VectorType_>> smoother_temp;
            for (unsigned long i(0) ; i < max_level ; ++i)
            {
                MatrixType_ Ai(MatIOType_::read_matrix(file_base + "A_" + std::to_string(i + 1) + ".ell",
                                                       DataType_(0)));
                A.push_back(Ai);
                PreconFill<PreconContType_, MatrixType_, DataType_>::value(i, P, file_base + "A_", precon_suffix, Ai,
                                                                           damping_factor);
                if (i + 1 < max_level)
                {
                    TransferContType_ pr(MatIOType_::read_matrix(file_base + "prol_" + std::to_string(i + 2) + ".ell",
                                                                 DataType_(0)));
                    Prol.push_back(pr);
                    Res.push_back(Transposition<Tag_>::value(pr.copy()));
                }
                const unsigned long n = Ai.rows();
                if (i + 1 == max_level)
                {
                    b.push_back(VecIOType_::read_vector(file_base + "rhs_" + std::to_string(max_level), DataType_(0)));
                    x.push_back(VecIOType_::read_vector(file_base + "init_" + std::to_string(max_level), DataType_(0)));
                }
                else
                {
                    b.push_back(VectorType_(n, DataType_(0)));
                    x.push_back(VectorType_(n, DataType_(0)));
                }
                d.push_back(VectorType_(n, DataType_(0)));
                c.push_back(VectorType_(n, DataType_(0)));
                store.push_back(VectorType_(n, DataType_(0)));
                smoother_temp.push_back(std::vector<VectorType_>());
            }
            return Data(A, Res, Prol, P, b, x, d, c, store, smoother_temp, 0, 0, 0, 0, 0, DataType_(0));
        }
    };

    /// Builds one cycle of the given shape into an OperatorList (mg.hh:548-1060).
    template <typename Tag_, typename CycleShape_, typename CoarseGridSolverType_, typename SmootherType_,
              typename ResType_, typename ProlType_, typename DataType_>
    struct MGCycleCreation;

    namespace mg_detail
    {
        enum Shape { V, W, F };

        template <typename Tag_, typename CoarseGridSolverType_, typename SmootherType_, typename ResType_,
                  typename ProlType_, typename DataType_>
        struct Builder
        {
            template <typename MatrixType_, typename VectorType_, typename TransferContType_, typename PreconContType_>
            static void build(Shape shape, std::vector<VectorType_> & x, std::vector<VectorType_> & b,
                              unsigned long level, OperatorList & cycle,
                              MGData<MatrixType_, VectorType_, TransferContType_, PreconContType_, DataType_> & data)
            {
                const unsigned long i = MGDataIndex::internal_index_A(level);
                if (level == data.min_level)
                {
                    cycle.push_back(new SolverOperator<CoarseGridSolverType_, MatrixType_, VectorType_,
                                    PreconContType_, DataType_>(data.A.at(i), data.P.at(i), b.at(i), x.at(i),
                                    data.max_iters_coarse, data.used_iters_coarse, data.eps_relative));
                    return;
                }
                const unsigned long t = MGDataIndex::internal_index_Prol(level);
                cycle.push_back(new SmootherOperator<SmootherType_, MatrixType_, VectorType_, PreconContType_>(
                                data.A.at(i), data.P.at(i), b.at(i), x.at(i), data.smoother_temp.at(i),
                                data.n_pre_smooth));
                cycle.push_back(new DefectOperator<Tag_, MatrixType_, VectorType_>(data.store.at(i), b.at(i),
                                data.A.at(i), x.at(i)));
                cycle.push_back(new TransferOperator<ResType_, TransferContType_, VectorType_>(data.d.at(i - 1),
                                data.store.at(i), data.resmat.at(t)));
                // the coarse correction starts from zero
                cycle.push_back(new FillOperator<Tag_, VectorType_>(data.c.at(i - 1)));
                if (shape == V)
                    build(V, data.c, data.d, level - 1, cycle, data);
                else if (shape == W)
                {
                    if (level - 1 != data.min_level)
                        build(W, data.c, data.d, level - 1, cycle, data);
                    build(W, data.c, data.d, level - 1, cycle, data);
                }
                else
                {
                    build(F, data.c, data.d, level - 1, cycle, data);
                    build(V, data.c, data.d, level - 1, cycle, data);
                }
                cycle.push_back(new TransferOperator<ProlType_, TransferContType_, VectorType_>(data.store.at(i),
                                data.c.at(i - 1), data.prolmat.at(t)));
                cycle.push_back(new SumOperator<Tag_, VectorType_>(x.at(i), data.store.at(i)));
                cycle.push_back(new SmootherOperator<SmootherType_, MatrixType_, VectorType_, PreconContType_>(
                                data.A.at(i), data.P.at(i), b.at(i), x.at(i), data.smoother_temp.at(i),
                                data.n_post_smooth));
            }
        };
    }

#define HONEI_MG_CYCLE_CREATION(CYCLE_TAG, SHAPE)                                                              \
    template <typename Tag_, typename CoarseGridSolverType_, typename SmootherType_, typename ResType_,       \
              typename ProlType_, typename DataType_>                                                        \
    struct MGCycleCreation<Tag_, CYCLE_TAG, CoarseGridSolverType_, SmootherType_, ResType_, ProlType_,       \
                           DataType_>                                                                        \
    {                                                                                                        \
        template <typename MatrixType_, typename VectorType_, typename TransferContType_,                    \
                  typename PreconContType_>                                                                  \
        static OperatorList value(MGData<MatrixType_, VectorType_, TransferContType_, PreconContType_,        \
                                         DataType_> & data)                                                  \
        {                                                                                                    \
            OperatorList cycle;                                                                              \
            mg_detail::Builder<Tag_, CoarseGridSolverType_, SmootherType_, ResType_, ProlType_,              \
                               DataType_>::build(mg_detail::SHAPE, data.x, data.b, data.A.size(), cycle,     \
                                                 data);                                                      \
            return cycle;                                                                                    \
        }                                                                                                    \
    };

    HONEI_MG_CYCLE_CREATION(methods::CYCLE::V::STATIC, V)
    HONEI_MG_CYCLE_CREATION(methods::CYCLE::W::STATIC, W)
    HONEI_MG_CYCLE_CREATION(methods::CYCLE::F::V::STATIC, F)
#undef HONEI_MG_CYCLE_CREATION

    /// Cycles until ||b - A x|| < eps_relative * ||b - A x0|| or max_iters
    /// (mg.hh:1064-1100); used_iters counts the cycles run.
    template <typename Tag_, typename NormType_> struct MGSolver
    {
        template <typename MatrixType_, typename VectorType_, typename TransferContType_, typename PreconContType_,
                  typename DataType_>
        static void value(MGData<MatrixType_, VectorType_, TransferContType_, PreconContType_, DataType_> & data,
                          OperatorList & cycle)
        {
            if (cycle.size() == 0)
                throw InternalError("MGSolver: the operator list is empty");
            VectorType_ & xf = data.x.back();
            VectorType_ & bf = data.b.back();
            MatrixType_ & Af = data.A.back();
            VectorType_ r(xf.size());
            Defect<Tag_>::value(r, bf, Af, xf);
            const DataType_ rnorm_initial = NormType_::value(r);
            std::cout << "Initial norm: " << rnorm_initial << std::endl;
            DataType_ rnorm_current(1e16);
            for (unsigned long i(0) ; i < data.max_iters ; ++i)
            {
                cycle.value();
                Defect<Tag_>::value(r, bf, Af, xf);
                rnorm_current = NormType_::value(r);
                data.used_iters = i + 1;
                if (rnorm_current < data.eps_relative * rnorm_initial)
                    break;
            }
            std::cout << "Final norm: " << rnorm_current << std::endl;
        }
    };

    /// max_iters cycles without a stopping test (mg.hh:1102-1118)
    struct MGSmoother
    {
        static void value(OperatorList & cycle, unsigned long max_iters)
        {
            if (cycle.size() == 0)
                throw InternalError("MGSmoother: the operator list is empty");
            for (unsigned long i(0) ; i < max_iters ; ++i)
                cycle.value();
        }
    };
}
