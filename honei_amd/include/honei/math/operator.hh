// honei/math/operator.hh -- the operators an operator-list multigrid cycle
// is built from (HONEI's honei/math/operator.hh:33-441): each one holds
// shallow copies of its vectors and matrices (DenseVector and
// SparseMatrixELL share their storage) and runs one Tag_ operation when
// value() is called.  Class names, constructor arguments and
// transfer_type() codes (1 restriction, 2 prolongation, 3 smoother,
// 4 coarse solver) are kept alike so that callers and print_cycle work.
#pragma once

#include <honei/la/dense_vector.hh>
#include <honei/la/operations.hh>
#include <honei/math/solvers.hh>

#include <string>
#include <vector>

namespace honei
{
    class Operator
    {
        public:
            virtual void value() = 0;
            virtual ~Operator() {}
            virtual std::string to_string() = 0;
            virtual unsigned long transfer_type() { return 0; }
    };

    /// y = x
    template <typename Tag_, typename VectorType_> class CopyOperator : public Operator
    {
        private:
            VectorType_ _x, _y;

        public:
            CopyOperator(VectorType_ & x, VectorType_ & y) : _x(x), _y(y) {}
            virtual void value() { copy<Tag_>(_x, _y); }
            virtual std::string to_string() { return "CopyOperator"; }
    };

    /// x = 0
    template <typename Tag_, typename VectorType_> class FillOperator : public Operator
    {
        private:
            VectorType_ _x;

        public:
            explicit FillOperator(VectorType_ & x) : _x(x) {}
            virtual void value() { fill<Tag_>(_x, 0); }
            virtual std::string to_string() { return "FillOperator"; }
    };

    /// y = b - A x
    template <typename Tag_, typename MatType_, typename VectorType_> class DefectOperator : public Operator
    {
        private:
            VectorType_ _y, _b;
            MatType_ _A;
            VectorType_ _x;

        public:
            DefectOperator(VectorType_ & y, VectorType_ & b, MatType_ & A, VectorType_ & x) :
                _y(y), _b(b), _A(A), _x(x)
            {
            }
            virtual void value() { Defect<Tag_>::value(_y, _b, _A, _x); }
            virtual std::string to_string() { return "DefectOperator"; }
    };

    /// y = ||x|| (or its square)
    template <typename Tag_, VectorNormType normtype, typename DT_, bool root> class NormOperator : public Operator
    {
        private:
            DenseVector<DT_> _x;
            DT_ & _y;

        public:
            NormOperator(DT_ & y, DenseVector<DT_> & x) : _x(x), _y(y) {}
            virtual void value() { _y = Norm<normtype, root, Tag_>::value(_x); }
            virtual std::string to_string() { return "NormOperator"; }
    };

    /// y += x
    template <typename Tag_, typename VectorType_> class SumOperator : public Operator
    {
        private:
            VectorType_ _x, _y;

        public:
            SumOperator(VectorType_ & y, VectorType_ & x) : _x(x), _y(y) {}
            virtual void value() { Sum<Tag_>::value(_y, _x); }
            virtual std::string to_string()
            {
                return "SumOperator\n  left.size()= " + std::to_string(_x.size()) + "\n  right.size()= " +
                    std::to_string(_y.size()) + "\n";
            }
    };

    /// y += alpha x
    template <typename Tag_, typename DT_> class ScaledSumOperator : public Operator
    {
        private:
            DenseVector<DT_> _x, _y;
            DT_ _alpha;

        public:
            ScaledSumOperator(DenseVector<DT_> & y, DenseVector<DT_> & x, DT_ alpha) : _x(x), _y(y), _alpha(alpha) {}
            virtual void value() { ScaledSum<Tag_>::value(_y, _x, _alpha); }
            virtual std::string to_string() { return "ScaledSumOperator"; }
    };

    /// coarse-grid solve: SolverType_::value(A, P, b, x, max_iters, used_iters, eps)
    template <typename SolverType_, typename MatrixType_, typename VectorType_, typename PreconContType_,
              typename DataType_>
    class SolverOperator : public Operator
    {
        private:
            MatrixType_ _A;
            PreconContType_ _P;
            VectorType_ _b, _x;
            unsigned long _max_iters;
            unsigned long & _used_iters;
            DataType_ _eps_relative;

        public:
            SolverOperator(MatrixType_ & A, PreconContType_ & P, VectorType_ & b, VectorType_ & x,
                           unsigned long max_iters, unsigned long & used_iters, DataType_ eps_relative = 1e-8) :
                _A(A), _P(P), _b(b), _x(x), _max_iters(max_iters), _used_iters(used_iters), _eps_relative(eps_relative)
            {
            }
            virtual void value() { SolverType_::value(_A, _P, _b, _x, _max_iters, _used_iters, _eps_relative); }
            virtual std::string to_string() { return "SolverOperator"; }
            virtual unsigned long transfer_type() { return 4; }
    };

    /// smoothing: SmootherType_::value(A, P, b, x, temp_vecs, max_iters); the
    /// constructor tops temp_vecs up to SmootherType_::NUM_TEMPVECS vectors
    template <typename SmootherType_, typename MatrixType_, typename VectorType_, typename PreconContType_>
    class SmootherOperator : public Operator
    {
        private:
            MatrixType_ _A;
            PreconContType_ _P;
            VectorType_ _b, _x;
            std::vector<VectorType_> & _temp_vecs;
            unsigned long _max_iters;

        public:
            SmootherOperator(MatrixType_ & A, PreconContType_ & P, VectorType_ & b, VectorType_ & x,
                             std::vector<VectorType_> & temp_vecs, unsigned long max_iters) :
                _A(A), _P(P), _b(b), _x(x), _temp_vecs(temp_vecs), _max_iters(max_iters)
            {
                while (_temp_vecs.size() < SmootherType_::NUM_TEMPVECS)
                    _temp_vecs.push_back(VectorType_(_b.size()));
            }
            virtual void value() { SmootherType_::value(_A, _P, _b, _x, _temp_vecs, _max_iters); }
            virtual std::string to_string() { return "SmootherOperator"; }
            virtual unsigned long transfer_type() { return 3; }
    };

    /// grid transfer: TransferType_::value(left, right, mask, mat)
    template <typename TransferType_, typename MatrixType_, typename VectorType_> class TransferOperator : public Operator
    {
        private:
            VectorType_ _left, _right;
            MatrixType_ _mat;

        public:
            TransferOperator(VectorType_ & left, VectorType_ & right, MatrixType_ & mat) :
                _left(left), _right(right), _mat(mat)
            {
            }
            virtual void value()
            {
                DenseVector<unsigned long> mask(8, 2ul);
                TransferType_::value(_left, _right, mask, _mat);
            }
            virtual std::string to_string()
            {
                return "TransferOperator\n  left.size()= " + std::to_string(_left.size()) + "\n  right.size()= " +
                    std::to_string(_right.size()) + "\n  mat_dims= " + std::to_string(_mat.rows()) + " x " +
                    std::to_string(_mat.columns()) + "\n";
            }
            virtual unsigned long transfer_type() { return _left.size() > _right.size() ? 2 : 1; }
    };
}
