// honei/la/operations.hh -- the linear-algebra operation templates and their
// tags::GPU::MI355X specialisations.
//
// Each op keeps HONEI's class-template-per-tag interface and argument
// meaning (declarations: honei/la/product.hh:1728-1794,
// dot_product.hh:343-367, norm.hh:538-580, scaled_sum.hh:444-488, sum.hh,
// scale.hh, difference.hh, element_product.hh, element_inverse.hh,
// algorithm.hh:50-200): sizes are checked on the host and mismatches throw
// VectorSizeDoesNotMatch; the work is one call of the C-ABI on device-
// resident data (DenseVector::lock(..., tv_gpu_hip)).  Reductions return
// the value on the host, as in HONEI.
#pragma once

#include <honei/backends/mi355x/abi.hh>
#include <honei/la/banded_matrix_qx.hh>
#include <honei/la/dense_vector.hh>
#include <honei/la/sparse_matrix_csr.hh>
#include <honei/la/sparse_matrix_ell.hh>

namespace honei
{
    enum VectorNormType
    {
        vnt_max = 0,
        vnt_l_one,
        vnt_l_two
    };

    template <typename Tag_> struct Product;
    template <typename Tag_> struct DotProduct;
    template <VectorNormType norm_type_, bool root_, typename Tag_> struct Norm;
    template <typename Tag_> struct ScaledSum;
    template <typename Tag_> struct Sum;
    template <typename Tag_> struct Scale;
    template <typename Tag_> struct Difference;
    template <typename Tag_> struct ElementProduct;
    template <typename Tag_> struct ElementInverse;

    namespace mi355x
    {
        template <typename DT> inline DT * wr(DenseVector<DT> & v) { return static_cast<DT *>(v.lock(lm_write_only, tv_gpu_hip)); }
        template <typename DT> inline DT * rw(DenseVector<DT> & v) { return static_cast<DT *>(v.lock(lm_read_and_write, tv_gpu_hip)); }
        template <typename DT> inline const DT * rd(const DenseVector<DT> & v)
        {
            return static_cast<const DT *>(v.lock(lm_read_only, tv_gpu_hip));
        }
        inline void same_size(unsigned long a, unsigned long b)
        {
            if (a != b)
                throw VectorSizeDoesNotMatch(a, b);
        }
    }

    template <> struct Product<tags::GPU::MI355X>
    {
        /// y = A x on rows [row_start, row_end) (row_end 0: all rows); product.hh:148-206
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & result, const SparseMatrixELL<DT> & a, const DenseVector<DT> & b,
                                       unsigned long row_start = 0, unsigned long row_end = 0)
        {
            using namespace mi355x;
            if (b.size() != a.columns())
                throw VectorSizeDoesNotMatch(b.size(), a.columns());
            if (a.rows() != result.size())
                throw VectorSizeDoesNotMatch(a.rows(), result.size());
            if (b.memid() == result.memid())
                throw InternalError("Product: result must not share storage with the operand");
            if (row_end == 0)
                row_end = a.rows();
            if (row_start > row_end || row_end > a.rows())
                throw VectorSizeDoesNotMatch(row_end, a.rows());
            if (row_start == 0 && row_end == a.rows())
                check(hsolve_ell_spmv(a.device_handle(), rd(b), wr(result), nullptr, stream()), "Product ELL");
            else if (row_start < row_end)
                // rows outside the range keep their values: the result is read-write here
                check(hsolve_ell_spmv_rows(a.device_handle(), rd(b), rw(result), nullptr, row_start, row_end, stream()),
                      "Product ELL rows");
            return result;
        }

        template <typename DT>
        static DenseVector<DT> value(const SparseMatrixELL<DT> & a, const DenseVector<DT> & b)
        {
            DenseVector<DT> result(a.rows());
            return value(result, a, b);
        }

        /// y = A x (blocked CSR) on rows [row_start, row_end) (row_end 0: all
        /// rows; the others keep their values); product.hh:208-245,
        /// product-cuda.cc:643-700
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & result, const SparseMatrixCSR<DT> & a, const DenseVector<DT> & b,
                                       unsigned long row_start = 0, unsigned long row_end = 0)
        {
            using namespace mi355x;
            if (b.size() != a.columns())
                throw VectorSizeDoesNotMatch(b.size(), a.columns());
            if (a.rows() != result.size())
                throw VectorSizeDoesNotMatch(a.rows(), result.size());
            if (b.memid() == result.memid())
                throw InternalError("Product: result must not share storage with the operand");
            if (row_end == 0)
                row_end = a.rows();
            const bool all = row_start == 0 && row_end == a.rows();
            check(hsolve_csr_spmv(a.device_handle(), rd(b), all ? wr(result) : rw(result), nullptr, row_start, row_end,
                                  stream()),
                  "Product CSR");
            return result;
        }

        template <typename DT>
        static DenseVector<DT> value(const SparseMatrixCSR<DT> & a, const DenseVector<DT> & b)
        {
            DenseVector<DT> result(a.rows());
            return value(result, a, b);
        }

        /// y = A x for Q1 bands; product.hh:370-403 summation order
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & result, const BandedMatrixQx<Q1Type, DT> & a,
                                       const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(b.size(), a.columns());
            same_size(result.size(), a.rows());
            if (b.memid() == result.memid())
                throw InternalError("Product: result must not share storage with the operand");
            const DT * bands[9];
            a.device_bands(bands);
            check(abi::product_q1(bands, rd(b), wr(result), a.size(), a.root()), "Product Q1");
            return result;
        }

        template <typename DT>
        static DenseVector<DT> value(const BandedMatrixQx<Q1Type, DT> & a, const DenseVector<DT> & b)
        {
            DenseVector<DT> result(a.rows());
            return value(result, a, b);
        }

        /// A DenseVector as a diagonal matrix: y = a * b elementwise (product.hh:1198-1203)
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & y, const DenseVector<DT> & a, const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(a.size(), b.size());
            same_size(y.size(), a.size());
            check(abi::element_product3(rw(y), rd(a), rd(b), y.size()), "Product diagonal");
            return y;
        }
    };

    template <> struct DotProduct<tags::GPU::MI355X>
    {
        template <typename DT> static DT value(const DenseVector<DT> & x, const DenseVector<DT> & y)
        {
            using namespace mi355x;
            if (x.size() != y.size())
                throw VectorSizeDoesNotMatch(y.size(), x.size());
            DT r(0);
            check(abi::dot(rd(x), rd(y), x.size(), &r), "DotProduct");
            return r;
        }
    };

    template <bool root_> struct Norm<vnt_l_two, root_, tags::GPU::MI355X>
    {
        /// squared L2 norm (root_ false) or L2 norm (root_ true); norm.hh:290-420
        template <typename DT> static DT value(const DenseVector<DT> & x)
        {
            using namespace mi355x;
            DT r(0);
            check(abi::norm(rd(x), x.size(), root_ ? 1 : 0, &r), "Norm");
            return r;
        }
    };

    template <> struct ScaledSum<tags::GPU::MI355X>
    {
        /// x += b * y
        template <typename DT> static DenseVector<DT> & value(DenseVector<DT> & x, const DenseVector<DT> & y, DT b)
        {
            using namespace mi355x;
            same_size(y.size(), x.size());
            DT * px = rw(x);
            check(abi::scaled_sum_s(px, px, rd(y), b, x.size()), "ScaledSum");
            return x;
        }

        /// r = x + b * y
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & r, const DenseVector<DT> & x, const DenseVector<DT> & y, DT b)
        {
            using namespace mi355x;
            same_size(y.size(), x.size());
            same_size(r.size(), x.size());
            const DT * px = rd(x);
            const DT * py = rd(y);
            check(abi::scaled_sum_s(r.memid() == x.memid() ? rw(r) : wr(r), px, py, b, x.size()), "ScaledSum");
            return r;
        }

        /// a += b * c (elementwise)
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & a, const DenseVector<DT> & b, const DenseVector<DT> & c)
        {
            using namespace mi355x;
            same_size(b.size(), a.size());
            same_size(c.size(), a.size());
            check(abi::scaled_sum_v(rw(a), rd(b), rd(c), a.size()), "ScaledSum");
            return a;
        }
    };

    template <> struct Sum<tags::GPU::MI355X>
    {
        template <typename DT> static DenseVector<DT> & value(DenseVector<DT> & a, const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(b.size(), a.size());
            check(abi::sum(rw(a), rd(b), a.size()), "Sum");
            return a;
        }
    };

    template <> struct Scale<tags::GPU::MI355X>
    {
        /// x *= a, with a converted to the vector's type (as scale.cu:48-75)
        template <typename DT, typename ST> static DenseVector<DT> & value(DenseVector<DT> & x, const ST a)
        {
            using namespace mi355x;
            check(abi::scale(rw(x), DT(a), x.size()), "Scale");
            return x;
        }
    };

    template <> struct Difference<tags::GPU::MI355X>
    {
        /// a -= b
        template <typename DT> static DenseVector<DT> & value(DenseVector<DT> & a, const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(b.size(), a.size());
            check(abi::difference(rw(a), rd(b), a.size()), "Difference");
            return a;
        }

        /// r = a - b
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & r, const DenseVector<DT> & a, const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(b.size(), a.size());
            same_size(r.size(), a.size());
            const DT * pa = rd(a);
            const DT * pb = rd(b);
            const bool alias = r.memid() == a.memid() || r.memid() == b.memid();
            check(abi::difference3(alias ? rw(r) : wr(r), pa, pb, a.size()), "Difference");
            return r;
        }
    };

    template <> struct ElementProduct<tags::GPU::MI355X>
    {
        /// a *= b
        template <typename DT> static DenseVector<DT> & value(DenseVector<DT> & a, const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(b.size(), a.size());
            DT * pa = rw(a);
            check(abi::element_product3(pa, pa, rd(b), a.size()), "ElementProduct");
            return a;
        }

        /// r = a * b
        template <typename DT>
        static DenseVector<DT> & value(DenseVector<DT> & r, const DenseVector<DT> & a, const DenseVector<DT> & b)
        {
            using namespace mi355x;
            same_size(b.size(), a.size());
            same_size(r.size(), a.size());
            const DT * pa = rd(a);
            const DT * pb = rd(b);
            const bool alias = r.memid() == a.memid() || r.memid() == b.memid();
            check(abi::element_product3(alias ? rw(r) : wr(r), pa, pb, a.size()), "ElementProduct");
            return r;
        }
    };

    template <> struct ElementInverse<tags::GPU::MI355X>
    {
        /// x = 1/x where x != 0 (element_inverse.hh:66-80)
        template <typename DT> static DenseVector<DT> & value(DenseVector<DT> & x)
        {
            using namespace mi355x;
            check(abi::element_inverse(rw(x), x.size()), "ElementInverse");
            return x;
        }
    };

    /// copy<Tag>(source, dest): device copy (algorithm.hh:50-86)
    template <typename Tag_, typename DT> void copy(const DenseVector<DT> & source, DenseVector<DT> & dest)
    {
        using namespace mi355x;
        same_size(source.size(), dest.size());
        check(hsolve_copy(rd(source), wr(dest), sizeof(DT) * source.size(), stream()), "copy");
    }

    /// fill<Tag>(dest, value) (algorithm.hh:132-160)
    template <typename Tag_, typename DT, typename VT> void fill(DenseVector<DT> & dest, const VT value)
    {
        using namespace mi355x;
        check(abi::fill(wr(dest), DT(value), dest.size()), "fill");
    }

    /// convert<Tag>(dest, source): FP64 <-> FP32, round to nearest (algorithm.hh:163-200)
    template <typename Tag_> void convert(DenseVector<float> & dest, const DenseVector<double> & source)
    {
        using namespace mi355x;
        same_size(source.size(), dest.size());
        check(hsolve_convert_double_float(rd(source), wr(dest), source.size(), stream()), "convert");
    }

    template <typename Tag_> void convert(DenseVector<double> & dest, const DenseVector<float> & source)
    {
        using namespace mi355x;
        same_size(source.size(), dest.size());
        check(hsolve_convert_float_double(rd(source), wr(dest), source.size(), stream()), "convert");
    }
}
