// honei/math/ri.hh -- preconditioned, damped Richardson iteration as solver
// and smoother (HONEI's honei/math/ri.hh:36-134):
//   repeat: t0 = b - A x; t1 = P t0; x += t1
// P is a DenseVector (elementwise product, e.g. damping / diag A) or a
// sparse matrix.  With a DenseVector P on tags::GPU::MI355X the three steps
// run as one fused pass per sweep (hsolve_ell_jacobi computes
// x + P * (b - A x) with the same rounding steps) into t1, which then
// becomes x.
#pragma once

#include <honei/la/dense_vector.hh>
#include <honei/la/operations.hh>
#include <honei/la/sparse_matrix_ell.hh>
#include <honei/math/solvers.hh>

#include <type_traits>
#include <vector>

namespace honei
{
    template <typename Tag_> struct RISolver
    {
        /// stops at ||t0|| < eps ||t0_initial||, ||t0|| < eps or after max_iters
        template <typename DT_, typename MatrixType_, typename VectorType_, typename PreconContType_>
        static VectorType_ & value(MatrixType_ & A, PreconContType_ & P, VectorType_ & b, VectorType_ & x,
                                   VectorType_ & temp_0, VectorType_ & temp_1, unsigned long max_iters,
                                   unsigned long & used_iters, DT_ eps_relative = 1e-8)
        {
            Defect<Tag_>::value(temp_0, b, A, x);
            const DT_ initial_defect = Norm<vnt_l_two, true, Tag_>::value(temp_0);
            used_iters = 0;
            while (true)
            {
                ++used_iters;
                Defect<Tag_>::value(temp_0, b, A, x);
                Product<Tag_>::value(temp_1, P, temp_0);
                Sum<Tag_>::value(x, temp_1);
                const DT_ current_defect = Norm<vnt_l_two, true, Tag_>::value(temp_0);
                if (current_defect < eps_relative * initial_defect || current_defect < eps_relative ||
                    used_iters >= max_iters)
                    break;
            }
            return x;
        }
    };

    template <typename Tag_> struct RISmoother
    {
        static const unsigned long NUM_TEMPVECS = 2;

        template <typename MatrixType_, typename VectorType_, typename PreconContType_>
        static void value(MatrixType_ & A, PreconContType_ & P, VectorType_ & b, VectorType_ & x,
                          std::vector<VectorType_> & temp_vecs, unsigned long max_iters)
        {
            VectorType_ temp_0(temp_vecs[0]);
            VectorType_ temp_1(temp_vecs[1]);
            for (unsigned long i(0) ; i < max_iters ; ++i)
            {
                Defect<Tag_>::value(temp_0, b, A, x);
                Product<Tag_>::value(temp_1, P, temp_0);
                Sum<Tag_>::value(x, temp_1);
            }
        }
    };

    template <> struct RISmoother<tags::GPU::MI355X>
    {
        static const unsigned long NUM_TEMPVECS = 2;

        template <typename DT, typename PreconContType_>
        static void value(SparseMatrixELL<DT> & A, PreconContType_ & P, DenseVector<DT> & b, DenseVector<DT> & x,
                          std::vector<DenseVector<DT>> & temp_vecs, unsigned long max_iters)
        {
            typedef tags::GPU::MI355X Tag;
            if constexpr (std::is_same<PreconContType_, DenseVector<DT>>::value)
            {
                using namespace mi355x;
                same_size(b.size(), A.rows());
                same_size(x.size(), A.rows());
                same_size(P.size(), A.rows());
                // sweeps alternate between x and t; an odd count ends with t -> x
                DenseVector<DT> t(temp_vecs[1]);
                for (unsigned long i(0) ; i < max_iters ; ++i)
                {
                    DenseVector<DT> & from = (i % 2 == 0) ? x : t;
                    DenseVector<DT> & to = (i % 2 == 0) ? t : x;
                    check(hsolve_ell_jacobi(A.device_handle(), rd(b), rd(P), rd(from), wr(to), stream()),
                          "RISmoother");
                }
                if (max_iters % 2 == 1)
                    copy<Tag>(t, x);
            }
            else
            {
                DenseVector<DT> temp_0(temp_vecs[0]);
                DenseVector<DT> temp_1(temp_vecs[1]);
                for (unsigned long i(0) ; i < max_iters ; ++i)
                {
                    Defect<Tag>::value(temp_0, b, A, x);
                    Product<Tag>::value(temp_1, P, temp_0);
                    Sum<Tag>::value(x, temp_1);
                }
            }
        }
    };
}
