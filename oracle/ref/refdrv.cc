/*
 * refdrv.cc -- C entry points over HONEI's own tags::CPU templates, compiled
 * from the reference sources in place (see oracle/ref/Makefile).
 *
 * TEST INFRASTRUCTURE ONLY.  Used in the development container to check the
 * C restatement (oracle.c) against the reference itself and to write golden
 * fixtures under tests/golden/.  The built library sits in oracle/_ref/,
 * outside version control; it is copied along with the working tree like
 * every built library.  Nothing in the product library, the GPU tests or
 * smoke loads it; the benchmark times ref_sse_product_smell_dv_f64 from it
 * as the CPU baseline (HONEI's own SSE kernel on the GPU host's cores).
 */
#include <honei/la/dense_vector.hh>
#include <honei/la/banded_matrix_qx.hh>
#include <honei/la/sparse_matrix.hh>
#include <honei/la/sparse_matrix_ell.hh>
#include <honei/la/sparse_matrix_csr.hh>
#include <honei/la/product.hh>
#include <honei/la/dot_product.hh>
#include <honei/la/norm.hh>
#include <honei/la/scaled_sum.hh>
#include <honei/la/element_inverse.hh>
#include <honei/la/scale.hh>
#include <honei/math/defect.hh>
#include <honei/math/cg.hh>
#include <honei/math/conjugate_gradients.hh>
#include <honei/math/jacobi.hh>
#include <honei/math/multigrid.hh>
#include <honei/math/restriction.hh>
#include <honei/math/prolongation.hh>
#include <honei/math/fill_matrix.hh>
#include <honei/math/mg.hh>
#include <honei/math/ri.hh>
#include <honei/math/ir.hh>
#include <honei/math/matrix_io.hh>
#include <honei/math/vector_io.hh>
#include <honei/util/configuration.hh>
#include <honei/backends/sse/operations.hh>

#include <cstring>
#include <iostream>
#include <sstream>
#include <stdint.h>
#include <string>

using namespace honei;

namespace
{
    template <typename DT_> DenseVector<DT_> dv(const DT_ * src, unsigned long n)
    {
        DenseVector<DT_> v(n);
        for (unsigned long i(0) ; i < n ; ++i)
            v[i] = src[i];
        return v;
    }

    template <typename DT_> void out(DT_ * dst, const DenseVector<DT_> & v)
    {
        for (unsigned long i(0) ; i < v.size() ; ++i)
            dst[i] = v[i];
    }

    template <typename DT_> BandedMatrixQx<Q1Type, DT_> q1(const DT_ * const * b, unsigned long n)
    {
        return BandedMatrixQx<Q1Type, DT_>(n, dv(b[0], n), dv(b[1], n), dv(b[2], n), dv(b[3], n),
                dv(b[4], n), dv(b[5], n), dv(b[6], n), dv(b[7], n), dv(b[8], n));
    }

    template <typename DT_> SparseMatrixELL<DT_> ell(unsigned long rows, unsigned long cols,
            unsigned long stride, unsigned long ncpr, unsigned long threads,
            const uint64_t * Aj, const DT_ * Ax)
    {
        DenseVector<unsigned long> aj(ncpr * stride);
        DenseVector<DT_> ax(ncpr * stride);
        for (unsigned long i(0) ; i < ncpr * stride ; ++i)
        {
            aj[i] = Aj[i];
            ax[i] = Ax[i];
        }
        return SparseMatrixELL<DT_>(rows, cols, stride, ncpr, aj, ax, threads);
    }

    // swallows what a solver prints to std::cout while it lives
    struct quiet
    {
        std::stringstream sink;
        std::streambuf * old;
        quiet() : old(std::cout.rdbuf(sink.rdbuf())) {}
        ~quiet() { std::cout.rdbuf(old); }
    };

    template <typename DT_> void ell_from_csr(unsigned long rows, unsigned long cols, const uint64_t * rowptr,
            const uint64_t * col, const DT_ * val, unsigned long threads,
            uint64_t * stride, uint64_t * ncpr, uint64_t * Aj, DT_ * Ax, uint64_t * Arl)
    {
        Configuration::instance()->set_value("ell::threads", (int)threads);
        SparseMatrix<DT_> sm(rows, cols);
        for (unsigned long r(0) ; r < rows ; ++r)
            for (uint64_t i(rowptr[r]) ; i < rowptr[r + 1] ; ++i)
                sm(r, col[i], val[i]);
        SparseMatrixELL<DT_> e(sm);
        *stride = e.stride();
        *ncpr = e.num_cols_per_row();
        if (Aj != 0)
        {
            for (unsigned long i(0) ; i < e.Aj().size() ; ++i)
            {
                Aj[i] = e.Aj()[i];
                Ax[i] = e.Ax()[i];
            }
            for (unsigned long i(0) ; i < rows ; ++i)
                Arl[i] = e.Arl()[i];
        }
        Configuration::instance()->set_value("ell::threads", 1);
    }
}

extern "C"
{

#define REF_TYPED(DT, S)                                                                                   \
void ref_ell_from_csr_##S(uint64_t rows, uint64_t cols, const uint64_t * rowptr, const uint64_t * col,     \
        const DT * val, uint64_t threads, uint64_t * stride, uint64_t * ncpr, uint64_t * Aj, DT * Ax,       \
        uint64_t * Arl)                                                                                    \
{                                                                                                          \
    ell_from_csr<DT>(rows, cols, rowptr, col, val, threads, stride, ncpr, Aj, Ax, Arl);                     \
}                                                                                                          \
                                                                                                           \
void ref_ell_product_##S(DT * y, uint64_t rows, uint64_t cols, uint64_t stride, uint64_t ncpr,            \
        uint64_t threads, const uint64_t * Aj, const DT * Ax, const DT * x, uint64_t * Arl_out)            \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, cols, stride, ncpr, threads, Aj, Ax));                            \
    DenseVector<DT> xv(dv(x, cols)), yv(rows);                                                             \
    Product<tags::CPU>::value(yv, A, xv);                                                                  \
    out(y, yv);                                                                                            \
    if (Arl_out)                                                                                           \
        for (unsigned long i(0) ; i < rows ; ++i) Arl_out[i] = A.Arl()[i];                                 \
}                                                                                                          \
                                                                                                           \
void ref_ell_defect_##S(DT * r, const DT * rhs, uint64_t rows, uint64_t stride, uint64_t ncpr,            \
        uint64_t threads, const uint64_t * Aj, const DT * Ax, const DT * x)                                \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, rows, stride, ncpr, threads, Aj, Ax));                            \
    DenseVector<DT> xv(dv(x, rows)), rv(rows), bv(dv(rhs, rows));                                          \
    Defect<tags::CPU>::value(rv, bv, A, xv);                                                               \
    out(r, rv);                                                                                            \
}                                                                                                          \
                                                                                                           \
void ref_q1_product_##S(DT * y, const DT * const * bands, const DT * x, uint64_t n)                       \
{                                                                                                          \
    BandedMatrixQx<Q1Type, DT> A(q1(bands, n));                                                            \
    DenseVector<DT> xv(dv(x, n)), yv(n);                                                                   \
    Product<tags::CPU>::value(yv, A, xv);                                                                  \
    out(y, yv);                                                                                            \
}                                                                                                          \
                                                                                                           \
void ref_q1_defect_##S(DT * r, const DT * rhs, const DT * const * bands, const DT * x, uint64_t n)        \
{                                                                                                          \
    BandedMatrixQx<Q1Type, DT> A(q1(bands, n));                                                            \
    DenseVector<DT> xv(dv(x, n)), rv(n), bv(dv(rhs, n));                                                   \
    Defect<tags::CPU>::value(rv, bv, A, xv);                                                               \
    out(r, rv);                                                                                            \
}                                                                                                          \
                                                                                                           \
DT ref_dot_##S(const DT * x, const DT * y, uint64_t n)                                                    \
{                                                                                                          \
    return DotProduct<tags::CPU>::value(dv(x, n), dv(y, n));                                               \
}                                                                                                          \
                                                                                                           \
DT ref_norm_##S(const DT * x, uint64_t n, int root)                                                       \
{                                                                                                          \
    DenseVector<DT> v(dv(x, n));                                                                           \
    return root ? Norm<vnt_l_two, true, tags::CPU>::value(v) : Norm<vnt_l_two, false, tags::CPU>::value(v);\
}                                                                                                          \
                                                                                                           \
void ref_restriction_##S(DT * coarse, uint64_t n_coarse, const DT * fine, uint64_t n_fine,                \
        const uint64_t * mask)                                                                             \
{                                                                                                          \
    DenseVector<DT> c(n_coarse, DT(0)), f(dv(fine, n_fine));                                               \
    DenseVector<unsigned long> m(8);                                                                       \
    for (int i(0) ; i < 8 ; ++i) m[i] = mask[i];                                                           \
    BandedMatrixQx<Q1Type, DT> dummy(9, DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)),              \
            DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)),               \
            DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)),               \
            DenseVector<DT>(9, DT(0)));                                                                    \
    Restriction<tags::CPU, methods::NONE>::value(c, f, m, dummy);                                          \
    out(coarse, c);                                                                                        \
}                                                                                                          \
                                                                                                           \
void ref_prolongation_##S(DT * fine, uint64_t n_fine, const DT * coarse, uint64_t n_coarse,               \
        const uint64_t * mask)                                                                             \
{                                                                                                          \
    DenseVector<DT> f(n_fine, DT(0)), c(dv(coarse, n_coarse));                                             \
    DenseVector<unsigned long> m(8);                                                                       \
    for (int i(0) ; i < 8 ; ++i) m[i] = mask[i];                                                           \
    BandedMatrixQx<Q1Type, DT> dummy(9, DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)),              \
            DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)),               \
            DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)), DenseVector<DT>(9, DT(0)),               \
            DenseVector<DT>(9, DT(0)));                                                                    \
    Prolongation<tags::CPU, methods::NONE>::value(f, c, m, dummy);                                         \
    out(fine, f);                                                                                          \
}                                                                                                          \
                                                                                                           \
void ref_jacobi_q1_##S(DT * x, const DT * const * bands, const DT * rhs, const DT * dinv, uint64_t n,     \
        uint64_t iters)                                                                                    \
{                                                                                                          \
    BandedMatrixQx<Q1Type, DT> A(q1(bands, n));                                                            \
    DenseVector<DT> xv(dv(x, n)), bv(dv(rhs, n)), di(dv(dinv, n));                                        \
    Jacobi<tags::CPU>::value(xv, A, bv, xv, iters, DT(0.7), di);                                          \
    out(x, xv);                                                                                            \
}                                                                                                          \
                                                                                                           \
uint64_t ref_cg_none_q1_##S(const DT * const * bands, const DT * b, DT * x, uint64_t n, uint64_t max_iters,\
        DT eps)                                                                                            \
{                                                                                                          \
    BandedMatrixQx<Q1Type, DT> A(q1(bands, n));                                                            \
    DenseVector<DT> bv(dv(b, n)), xv(dv(x, n));                                                            \
    unsigned long used(0);                                                                                 \
    CGSolver<tags::CPU, methods::NONE>::value(A, A, bv, xv, (unsigned long)max_iters, used, eps);         \
    out(x, xv);                                                                                            \
    return used;                                                                                           \
}                                                                                                          \
                                                                                                           \
uint64_t ref_cg_none_ell_##S(uint64_t rows, uint64_t stride, uint64_t ncpr, uint64_t threads,             \
        const uint64_t * Aj, const DT * Ax, const DT * b, DT * x, uint64_t max_iters, DT eps)              \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, rows, stride, ncpr, threads, Aj, Ax));                            \
    DenseVector<DT> bv(dv(b, rows)), xv(dv(x, rows));                                                      \
    unsigned long used(0);                                                                                 \
    CGSolver<tags::CPU, methods::NONE>::value(A, A, bv, xv, (unsigned long)max_iters, used, eps);         \
    out(x, xv);                                                                                            \
    return used;                                                                                           \
}                                                                                                          \
                                                                                                           \
uint64_t ref_cg_var_ell_##S(uint64_t rows, uint64_t stride, uint64_t ncpr, uint64_t threads,              \
        const uint64_t * Aj, const DT * Ax, const DT * P, const DT * b, DT * x, uint64_t max_iters,        \
        DT eps)                                                                                            \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, rows, stride, ncpr, threads, Aj, Ax));                            \
    DenseVector<DT> bv(dv(b, rows)), xv(dv(x, rows)), pv(dv(P, rows));                                     \
    unsigned long used(0);                                                                                 \
    CGSolver<tags::CPU, methods::VAR>::value(A, pv, bv, xv, (unsigned long)max_iters, used, eps);         \
    out(x, xv);                                                                                            \
    return used;                                                                                           \
}                                                                                                          \
                                                                                                           \
uint64_t ref_pcg_jac_ell_##S(uint64_t rows, uint64_t stride, uint64_t ncpr, uint64_t threads,            \
        const uint64_t * Aj, const DT * Ax, const DT * dd, const DT * b, DT * x, uint64_t max_iters,       \
        DT eps)                                                                                            \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, rows, stride, ncpr, threads, Aj, Ax));                            \
    DenseVector<DT> bv(dv(b, rows)), xv(dv(x, rows)), ddv(dv(dd, rows));                                    \
    unsigned long used(0);                                                                                 \
    ConjugateGradients<tags::CPU, methods::JAC>::value(A, bv, xv, ddv, (unsigned long)max_iters, used,    \
                                                       eps);                                               \
    out(x, xv);                                                                                            \
    return used;                                                                                           \
}                                                                                                          \
                                                                                                           \
void ref_cg_konv_q1_##S(const DT * const * bands, const DT * rhs, DT * x, uint64_t n, double konv_rad)    \
{                                                                                                          \
    BandedMatrixQx<Q1Type, DT> A(q1(bands, n));                                                            \
    DenseVector<DT> bv(dv(rhs, n)), xv(dv(x, n));                                                          \
    ConjugateGradients<tags::CPU, methods::NONE>::value(A, bv, xv, konv_rad);                             \
    out(x, xv);                                                                                            \
}                                                                                                          \
                                                                                                           \
void ref_cg_iters_q1_##S(const DT * const * bands, const DT * rhs, DT * x, uint64_t n, uint64_t iters)    \
{                                                                                                          \
    BandedMatrixQx<Q1Type, DT> A(q1(bands, n));                                                            \
    DenseVector<DT> bv(dv(rhs, n)), xv(dv(x, n));                                                          \
    quiet q;                                                                                               \
    ConjugateGradients<tags::CPU, methods::NONE>::value(A, bv, xv, (unsigned long)iters);                 \
    out(x, xv);                                                                                            \
}                                                                                                          \
                                                                                                           \
void ref_cg_iters_ell_##S(uint64_t rows, uint64_t stride, uint64_t ncpr, uint64_t threads,                 \
        const uint64_t * Aj, const DT * Ax, const DT * rhs, DT * x, uint64_t max_iters)                    \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, rows, stride, ncpr, threads, Aj, Ax));                            \
    DenseVector<DT> bv(dv(rhs, rows)), xv(dv(x, rows));                                                    \
    quiet q;                                                                                               \
    ConjugateGradients<tags::CPU, methods::NONE>::value(A, bv, xv, (unsigned long)max_iters);             \
    out(x, xv);                                                                                            \
}                                                                                                          \
                                                                                                           \
uint64_t ref_cg_var_ell_pm_##S(uint64_t rows, uint64_t stride, uint64_t ncpr, uint64_t threads,           \
        const uint64_t * Aj, const DT * Ax, uint64_t p_stride, uint64_t p_ncpr, uint64_t p_threads,        \
        const uint64_t * PAj, const DT * PAx, const DT * b, DT * x, uint64_t max_iters, DT eps)            \
{                                                                                                          \
    SparseMatrixELL<DT> A(ell<DT>(rows, rows, stride, ncpr, threads, Aj, Ax));                            \
    SparseMatrixELL<DT> P(ell<DT>(rows, rows, p_stride, p_ncpr, p_threads, PAj, PAx));                    \
    DenseVector<DT> bv(dv(b, rows)), xv(dv(x, rows));                                                      \
    unsigned long used(0);                                                                                 \
    CGSolver<tags::CPU, methods::VAR>::value(A, P, bv, xv, (unsigned long)max_iters, used, eps);          \
    out(x, xv);                                                                                            \
    return used;                                                                                           \
}

REF_TYPED(double, f64)
REF_TYPED(float, f32)

/* Multigrid<CPU, CPU, NONE, JAC, V, MIXED> set up as in
 * honei/math/poisson_mg_mixed_banded_TEST.cc:176-423; bands a32[level*9+b]
 * for levels min..max, dinv computed there as ElementInverse + Scale(0.7).
 * Returns the outer iteration count printed by the solver; *inner gets the
 * inner total. */
int ref_mg_mixed_q1(const double * const * A64, const double * b, double * x,
        const float * const * a32, uint64_t min_level, uint64_t max_level, uint64_t n_max_iter,
        double tolerance, uint64_t n_pre_smooth, double adapt, const uint64_t * mask,
        uint64_t * outer, uint64_t * inner)
{
    unsigned long n(((1ul << max_level) + 1) * ((1ul << max_level) + 1));
    BandedMatrixQx<Q1Type, double> A(q1(A64, n));
    MGInfo<float, BandedMatrixQx<Q1Type, float> > info;
    info.is_smoother = false;
    for (int i(0) ; i < 8 ; ++i)
        info.macro_border_mask[i] = mask[i];
    info.min_level = min_level;
    info.max_level = max_level;
    info.n_max_iter = n_max_iter;
    info.initial_zero = true;
    info.tolerance = tolerance;
    info.convergence_check = false;
    info.n_pre_smooth = n_pre_smooth;
    info.n_post_smooth = n_pre_smooth;
    info.n_max_iter_coarse = 0;
    info.tolerance_coarse = 1e-2;
    info.adapt_correction_factor = adapt;
    for (unsigned long i(0) ; i < min_level ; ++i)
    {
        unsigned long size((((1ul << i) + 1) * ((1ul << i) + 1)));
        if (i == 0)
            size = 9;
        DenseVector<float> dummy_band(size, float(0));
        BandedMatrixQx<Q1Type, float> ac_a(size, dummy_band, dummy_band, dummy_band, dummy_band, dummy_band,
                dummy_band, dummy_band, dummy_band, dummy_band);
        info.a.push_back(ac_a);
        info.c.push_back(DenseVector<float>(size, float(0)));
        info.d.push_back(DenseVector<float>(size, float(0)));
        info.rhs.push_back(DenseVector<float>(size, float(0)));
        info.x.push_back(DenseVector<float>(size, float(0)));
        info.diags_inverted.push_back(dummy_band.copy());
    }
    for (unsigned long i(min_level) ; i <= max_level ; ++i)
    {
        unsigned long size((((1ul << i) + 1) * ((1ul << i) + 1)));
        info.c.push_back(DenseVector<float>(size, float(0)));
        info.d.push_back(DenseVector<float>(size, float(0)));
        info.x.push_back(DenseVector<float>(size, float(0)));
        info.diags_inverted.push_back(DenseVector<float>(size, float(0)));
        info.rhs.push_back(DenseVector<float>(size, float(0)));
        info.a.push_back(q1(a32 + 9 * i, size));
    }
    for (unsigned long i(0) ; i <= max_level ; ++i)
    {
        DenseVector<float> scaled_diag_inverted(info.a[i].band(DD).copy());
        ElementInverse<tags::CPU>::value(scaled_diag_inverted);
        Scale<tags::CPU>::value(scaled_diag_inverted, 0.7);
        info.diags_inverted[i] = scaled_diag_inverted.copy();
    }
    DenseVector<double> bv(dv(b, n)), xv(dv(x, n));
    std::stringstream captured;
    std::streambuf * old(std::cout.rdbuf(captured.rdbuf()));
    Multigrid<tags::CPU, tags::CPU, methods::NONE, methods::JAC, methods::CYCLE::V, methods::MIXED>::value(
            A, bv, xv, (unsigned long)11, std::numeric_limits<double>::epsilon(), info);
    std::cout.rdbuf(old);
    out(x, xv);
    std::string line;
    *outer = 0;
    *inner = 0;
    while (std::getline(captured, line))
    {
        if (line.find("TN of outer iters: ") == 0)
            *outer = std::strtoul(line.c_str() + 19, 0, 10);
        if (line.find("TN of inner iters: ") == 0)
            *inner = std::strtoul(line.c_str() + 19, 0, 10);
    }
    return (int)*outer;
}

}  // extern "C"

/* ---- SparseMatrixCSR: construction (from a row-wise SparseMatrix or from an
 * ELL matrix, csr::blocksize = bs), Product and Defect on row ranges. */
namespace
{
    template <typename DT_> void csr_out(const SparseMatrixCSR<DT_> & c, uint64_t * Ar, uint64_t * Aj, DT_ * Ax,
            uint64_t * nblocks, uint64_t * used, uint64_t * axsize)
    {
        const unsigned long nb(c.Ar()[c.rows()]);
        *nblocks = nb;
        *used = c.used_elements();
        *axsize = c.Ax().size();
        if (Ar != 0)
        {
            for (unsigned long i(0) ; i <= c.rows() ; ++i)
                Ar[i] = c.Ar()[i];
            for (unsigned long i(0) ; i < nb ; ++i)
                Aj[i] = c.Aj()[i];
            for (unsigned long i(0) ; i < c.Ax().size() ; ++i)
                Ax[i] = c.Ax()[i];
        }
    }

    template <typename DT_> void csr_from_rows(uint64_t rows, uint64_t cols, const uint64_t * rowptr,
            const uint64_t * col, const DT_ * val, uint64_t bs, uint64_t * Ar, uint64_t * Aj, DT_ * Ax,
            uint64_t * nblocks, uint64_t * used, uint64_t * axsize)
    {
        Configuration::instance()->set_value("csr::blocksize", (int)bs);
        SparseMatrix<DT_> sm(rows, cols);
        for (unsigned long r(0) ; r < rows ; ++r)
            for (uint64_t i(rowptr[r]) ; i < rowptr[r + 1] ; ++i)
                sm(r, col[i], val[i]);
        SparseMatrixCSR<DT_> c(sm);
        csr_out(c, Ar, Aj, Ax, nblocks, used, axsize);
        Configuration::instance()->set_value("csr::blocksize", 1);
    }

    template <typename DT_> void csr_from_ell(uint64_t rows, uint64_t cols, uint64_t stride, uint64_t ncpr,
            uint64_t threads, const uint64_t * eAj, const DT_ * eAx, uint64_t bs, uint64_t * Ar, uint64_t * Aj,
            DT_ * Ax, uint64_t * nblocks, uint64_t * used, uint64_t * axsize)
    {
        Configuration::instance()->set_value("csr::blocksize", (int)bs);
        SparseMatrixELL<DT_> e(ell<DT_>(rows, cols, stride, ncpr, threads, eAj, eAx));
        SparseMatrixCSR<DT_> c(e);
        csr_out(c, Ar, Aj, Ax, nblocks, used, axsize);
        Configuration::instance()->set_value("csr::blocksize", 1);
    }

    template <typename DT_> SparseMatrixCSR<DT_> csr(uint64_t rows, uint64_t cols, uint64_t bs,
            uint64_t nblocks, uint64_t used, const uint64_t * Ar, const uint64_t * Aj, const DT_ * Ax)
    {
        DenseVector<unsigned long> ar(rows + 1), aj(nblocks ? nblocks : 1, 0ul);
        DenseVector<DT_> ax(nblocks ? nblocks * bs : 1, DT_(0));
        for (unsigned long i(0) ; i <= rows ; ++i)
            ar[i] = Ar[i];
        for (unsigned long i(0) ; i < nblocks ; ++i)
            aj[i] = Aj[i];
        for (unsigned long i(0) ; i < nblocks * bs ; ++i)
            ax[i] = Ax[i];
        return SparseMatrixCSR<DT_>(bs, rows, cols, aj, ax, ar, used);
    }

    template <typename DT_> void csr_product(DT_ * y, uint64_t rows, uint64_t cols, uint64_t bs,
            uint64_t nblocks, const uint64_t * Ar, const uint64_t * Aj, const DT_ * Ax, const DT_ * x,
            uint64_t row_start, uint64_t row_end)
    {
        SparseMatrixCSR<DT_> A(csr<DT_>(rows, cols, bs, nblocks, 0, Ar, Aj, Ax));
        DenseVector<DT_> xv(dv(x, cols)), yv(dv(y, rows));
        Product<tags::CPU>::value(yv, A, xv, row_start, row_end);
        out(y, yv);
    }

    template <typename DT_> void csr_defect(DT_ * r, const DT_ * rhs, uint64_t rows, uint64_t cols,
            uint64_t bs, uint64_t nblocks, const uint64_t * Ar, const uint64_t * Aj, const DT_ * Ax,
            const DT_ * x, uint64_t row_start, uint64_t row_end)
    {
        SparseMatrixCSR<DT_> A(csr<DT_>(rows, cols, bs, nblocks, 0, Ar, Aj, Ax));
        DenseVector<DT_> xv(dv(x, cols)), rv(dv(r, rows)), bv(dv(rhs, rows));
        Defect<tags::CPU>::value(rv, bv, A, xv, row_start, row_end);
        out(r, rv);
    }
}

extern "C"
{
#define REF_CSR(DT, S)                                                                                     \
void ref_csr_from_rows_##S(uint64_t rows, uint64_t cols, const uint64_t * rowptr, const uint64_t * col,    \
        const DT * val, uint64_t bs, uint64_t * Ar, uint64_t * Aj, DT * Ax, uint64_t * nblocks,             \
        uint64_t * used, uint64_t * axsize)                                                                \
{                                                                                                          \
    csr_from_rows<DT>(rows, cols, rowptr, col, val, bs, Ar, Aj, Ax, nblocks, used, axsize);                \
}                                                                                                          \
void ref_csr_from_ell_##S(uint64_t rows, uint64_t cols, uint64_t stride, uint64_t ncpr, uint64_t threads,  \
        const uint64_t * eAj, const DT * eAx, uint64_t bs, uint64_t * Ar, uint64_t * Aj, DT * Ax,          \
        uint64_t * nblocks, uint64_t * used, uint64_t * axsize)                                            \
{                                                                                                          \
    csr_from_ell<DT>(rows, cols, stride, ncpr, threads, eAj, eAx, bs, Ar, Aj, Ax, nblocks, used, axsize);  \
}                                                                                                          \
void ref_csr_product_##S(DT * y, uint64_t rows, uint64_t cols, uint64_t bs, uint64_t nblocks,              \
        const uint64_t * Ar, const uint64_t * Aj, const DT * Ax, const DT * x, uint64_t row_start,          \
        uint64_t row_end)                                                                                  \
{                                                                                                          \
    csr_product<DT>(y, rows, cols, bs, nblocks, Ar, Aj, Ax, x, row_start, row_end);                        \
}                                                                                                          \
void ref_csr_defect_##S(DT * r, const DT * rhs, uint64_t rows, uint64_t cols, uint64_t bs,                 \
        uint64_t nblocks, const uint64_t * Ar, const uint64_t * Aj, const DT * Ax, const DT * x,            \
        uint64_t row_start, uint64_t row_end)                                                              \
{                                                                                                          \
    csr_defect<DT>(r, rhs, rows, cols, bs, nblocks, Ar, Aj, Ax, x, row_start, row_end);                    \
}
REF_CSR(double, f64)
REF_CSR(float, f32)
#undef REF_CSR
}  // extern "C"

/* Operator-list MG from the reference's own loader and templates:
 * MGUtil::load_data(base, levels, damping, "jac") + configure, the cycle of
 * `cycle` (0 V, 1 W, 2 F::V) with CGSolver<NONE> (coarse_var 0) or <VAR>
 * (1), RISmoother and PROLMAT transfers, then MGSolver with the l2 norm.
 * x receives the finest iterate. */
template <typename DT_, typename Cycle_, typename Coarse_, typename Precon_>
static void mgop_run(const char * base, unsigned long levels, unsigned long min_level, unsigned long max_iters,
        unsigned long max_iters_coarse, unsigned long npre, unsigned long npost, double eps, double damping,
        DT_ * x, unsigned long * used, unsigned long * used_coarse, const char * suffix)
{
    typedef MGUtil<tags::CPU, SparseMatrixELL<DT_>, DenseVector<DT_>, SparseMatrixELL<DT_>, Precon_,
            MatrixIO<io_formats::ELL>, VectorIO<io_formats::EXP>, DT_> Util;
    MGData<SparseMatrixELL<DT_>, DenseVector<DT_>, SparseMatrixELL<DT_>, Precon_, DT_> data(
            Util::load_data(base, levels, DT_(damping), suffix));
    Util::configure(data, max_iters, max_iters_coarse, npre, npost, min_level, DT_(eps));
    OperatorList ol(MGCycleCreation<tags::CPU, Cycle_, Coarse_, RISmoother<tags::CPU>,
            Restriction<tags::CPU, methods::PROLMAT>, Prolongation<tags::CPU, methods::PROLMAT>, DT_>::value(data));
    std::stringstream captured;
    std::streambuf * old(std::cout.rdbuf(captured.rdbuf()));
    MGSolver<tags::CPU, Norm<vnt_l_two, true, tags::CPU> >::value(data, ol);
    std::cout.rdbuf(old);
    const DenseVector<DT_> & xr(data.x.at(data.x.size() - 1));
    for (unsigned long i(0) ; i < xr.size() ; ++i)
        x[i] = xr[i];
    *used = data.used_iters;
    *used_coarse = data.used_iters_coarse;
}

template <typename DT_, typename Coarse_, typename Precon_>
static int mgop_by_cycle(int cycle, const char * base, unsigned long levels, unsigned long min_level,
        unsigned long max_iters, unsigned long max_iters_coarse, unsigned long npre, unsigned long npost, double eps,
        double damping, DT_ * x, unsigned long * used, unsigned long * used_coarse, const char * suffix)
{
    if (cycle == 0)
        mgop_run<DT_, methods::CYCLE::V::STATIC, Coarse_, Precon_>(base, levels, min_level, max_iters,
                max_iters_coarse, npre, npost, eps, damping, x, used, used_coarse, suffix);
    else if (cycle == 1)
        mgop_run<DT_, methods::CYCLE::W::STATIC, Coarse_, Precon_>(base, levels, min_level, max_iters,
                max_iters_coarse, npre, npost, eps, damping, x, used, used_coarse, suffix);
    else if (cycle == 2)
        mgop_run<DT_, methods::CYCLE::F::V::STATIC, Coarse_, Precon_>(base, levels, min_level, max_iters,
                max_iters_coarse, npre, npost, eps, damping, x, used, used_coarse, suffix);
    else
        return -1;
    return 0;
}

extern "C" {

int ref_mgop_solve_f64(const char * base, int cycle, int coarse_var, uint64_t levels, uint64_t min_level,
        uint64_t max_iters, uint64_t max_iters_coarse, uint64_t npre, uint64_t npost, double eps, double damping,
        double * x, uint64_t * used, uint64_t * used_coarse)
{
    unsigned long u(0), uc(0);
    int st = coarse_var ?
        mgop_by_cycle<double, CGSolver<tags::CPU, methods::VAR>, DenseVector<double> >(cycle, base, levels,
                min_level, max_iters, max_iters_coarse, npre, npost, eps, damping, x, &u, &uc, "jac") :
        mgop_by_cycle<double, CGSolver<tags::CPU, methods::NONE>, DenseVector<double> >(cycle, base, levels,
                min_level, max_iters, max_iters_coarse, npre, npost, eps, damping, x, &u, &uc, "jac");
    *used = u;
    *used_coarse = uc;
    return st;
}

/* The same with SparseMatrixELL preconditioners read by PreconFill from
 * <base>A_<l>_<suffix>.ell (e.g. "spai", "sainv"), scaled by damping when
 * that is not 1. */
int ref_mgop_solve_pm_f64(const char * base, int cycle, int coarse_var, uint64_t levels, uint64_t min_level,
        uint64_t max_iters, uint64_t max_iters_coarse, uint64_t npre, uint64_t npost, double eps, double damping,
        const char * suffix, double * x, uint64_t * used, uint64_t * used_coarse)
{
    unsigned long u(0), uc(0);
    int st = coarse_var ?
        mgop_by_cycle<double, CGSolver<tags::CPU, methods::VAR>, SparseMatrixELL<double> >(cycle, base, levels,
                min_level, max_iters, max_iters_coarse, npre, npost, eps, damping, x, &u, &uc, suffix) :
        mgop_by_cycle<double, CGSolver<tags::CPU, methods::NONE>, SparseMatrixELL<double> >(cycle, base, levels,
                min_level, max_iters, max_iters_coarse, npre, npost, eps, damping, x, &u, &uc, suffix);
    *used = u;
    *used_coarse = uc;
    return st;
}

/* The IRSolver test's set-up (ir_TEST.cc:30-80): FP32 operator-list data
 * from `base` (V cycle, CGSolver<VAR> coarse, RISmoother, jac damping),
 * FP64 A_levels, rhs_levels, init_levels, then IRSolver. */
int ref_ir_mgop(const char * base, uint64_t levels, uint64_t min_level, uint64_t max_iters_inner,
        uint64_t max_iters_coarse, uint64_t npre, uint64_t npost, double eps_inner, double damping,
        double eps, uint64_t max_iters, double * x_out, uint64_t * used, uint64_t * inner_used,
        uint64_t * coarse_used)
{
    typedef MGUtil<tags::CPU, SparseMatrixELL<float>, DenseVector<float>, SparseMatrixELL<float>, DenseVector<float>,
            MatrixIO<io_formats::ELL>, VectorIO<io_formats::EXP>, float> Util;
    MGData<SparseMatrixELL<float>, DenseVector<float>, SparseMatrixELL<float>, DenseVector<float>, float> data(
            Util::load_data(base, levels, float(damping), "jac"));
    Util::configure(data, max_iters_inner, max_iters_coarse, npre, npost, min_level, float(eps_inner));
    OperatorList ol(MGCycleCreation<tags::CPU, methods::CYCLE::V::STATIC, CGSolver<tags::CPU, methods::VAR>,
            RISmoother<tags::CPU>, Restriction<tags::CPU, methods::PROLMAT>, Prolongation<tags::CPU, methods::PROLMAT>,
            float>::value(data));
    std::string b(base);
    SparseMatrixELL<double> A(MatrixIO<io_formats::ELL>::read_matrix(b + "A_" + stringify(levels) + ".ell", double(0)));
    DenseVector<double> rhs(VectorIO<io_formats::EXP>::read_vector(b + "rhs_" + stringify(levels), double(0)));
    DenseVector<double> x(VectorIO<io_formats::EXP>::read_vector(b + "init_" + stringify(levels), double(0)));
    unsigned long u(0);
    std::stringstream captured;
    std::streambuf * old(std::cout.rdbuf(captured.rdbuf()));
    IRSolver<tags::CPU, tags::CPU, Norm<vnt_l_two, true, tags::CPU> >::value(A, rhs, x, data, ol, eps,
            (unsigned long)max_iters, u);
    std::cout.rdbuf(old);
    for (unsigned long i(0) ; i < x.size() ; ++i)
        x_out[i] = x[i];
    *used = u;
    *inner_used = data.used_iters;
    *coarse_used = data.used_iters_coarse;
    return 0;
}

}  // extern "C"

/* ---- Multigrid<CPU, CPU, NONE|PROLMAT, JAC, CYCLE::V, FIXED|MIXED> on Q1 bands
 * or ELL matrices, with the MGInfo set-up of the reference's tests
 * (poisson_mg_*_TEST.cc): levels below min_level hold dummies, dinv is given. */
struct ref_ell
{
    uint64_t rows, cols, stride, ncpr;
    const uint64_t * Aj;
    const void * Ax;
};

namespace
{
    template <typename DT_> BandedMatrixQx<Q1Type, DT_> q1_dummy(unsigned long size)
    {
        DenseVector<DT_> z(size, DT_(0));
        return BandedMatrixQx<Q1Type, DT_>(size, z, z, z, z, z, z, z, z, z);
    }

    template <typename DT_> SparseMatrixELL<DT_> ell_of(const ref_ell & e)
    {
        return ell<DT_>(e.rows, e.cols, e.stride, e.ncpr, 1, e.Aj, (const DT_ *)e.Ax);
    }

    template <typename DT_> SparseMatrixELL<DT_> ell_dummy()
    {
        SparseMatrix<DT_> s(1, 1);
        s(0, 0, DT_(1));
        return SparseMatrixELL<DT_>(s);
    }

    template <typename Layout_> struct MGMats;
    template <typename DT_> struct MGMats<BandedMatrixQx<Q1Type, DT_> >
    {
        static BandedMatrixQx<Q1Type, DT_> level(const void * const * a_q1, const ref_ell *, unsigned long L,
                                                 unsigned long n)
        {
            return q1((const DT_ * const *)a_q1 + 9 * L, n);
        }
        static BandedMatrixQx<Q1Type, DT_> dummy(unsigned long n) { return q1_dummy<DT_>(n); }
        static BandedMatrixQx<Q1Type, DT_> transfer(const ref_ell *, unsigned long, unsigned long n)
        {
            return q1_dummy<DT_>(n);
        }
    };
    template <typename DT_> struct MGMats<SparseMatrixELL<DT_> >
    {
        static SparseMatrixELL<DT_> level(const void * const *, const ref_ell * a_ell, unsigned long L, unsigned long)
        {
            return ell_of<DT_>(a_ell[L]);
        }
        static SparseMatrixELL<DT_> dummy(unsigned long) { return ell_dummy<DT_>(); }
        static SparseMatrixELL<DT_> transfer(const ref_ell * t, unsigned long L, unsigned long)
        {
            return t ? ell_of<DT_>(t[L]) : ell_dummy<DT_>();
        }
    };

    template <typename DT_, typename Layout_>
    void mg_info(MGInfo<DT_, Layout_> & info, const void * const * a_q1, const ref_ell * a_ell,
                 const ref_ell * prol, const ref_ell * res, const void * const * dinv, uint64_t min_level,
                 uint64_t max_level, uint64_t n_max_iter, double tolerance, uint64_t n_pre, double adapt,
                 const uint64_t * mask, int initial_zero, int convergence_check)
    {
        info.is_smoother = false;
        info.macro_border_mask = DenseVector<unsigned long>(8);
        for (int i(0) ; i < 8 ; ++i)
            info.macro_border_mask[i] = mask[i];
        info.min_level = min_level;
        info.max_level = max_level;
        info.n_max_iter = n_max_iter;
        info.initial_zero = initial_zero != 0;
        info.tolerance = tolerance;
        info.convergence_check = convergence_check != 0;
        info.n_pre_smooth = n_pre;
        info.n_post_smooth = n_pre;
        info.n_max_iter_coarse = 0;
        info.tolerance_coarse = 1e-2;
        info.adapt_correction_factor = adapt;
        for (unsigned long L(0) ; L <= max_level ; ++L)
        {
            const bool real = L >= min_level;
            unsigned long size = ((1ul << L) + 1) * ((1ul << L) + 1);
            if (real && a_ell)
                size = a_ell[L].rows;
            info.a.push_back(real ? MGMats<Layout_>::level(a_q1, a_ell, L, size) : MGMats<Layout_>::dummy(size));
            info.prolmats.push_back(real && L > min_level ? MGMats<Layout_>::transfer(prol, L, size)
                                                          : MGMats<Layout_>::dummy(size));
            info.resmats.push_back(real && L > min_level ? MGMats<Layout_>::transfer(res, L, size)
                                                         : MGMats<Layout_>::dummy(size));
            info.c.push_back(DenseVector<DT_>(size, DT_(0)));
            info.d.push_back(DenseVector<DT_>(size, DT_(0)));
            info.rhs.push_back(DenseVector<DT_>(size, DT_(0)));
            info.x.push_back(DenseVector<DT_>(size, DT_(0)));
            info.diags_inverted.push_back(real ? dv((const DT_ *)dinv[L], size) : DenseVector<DT_>(size, DT_(0)));
        }
    }

    template <typename Prol_, typename Layout_, typename Outer_>
    void mg_mixed_run(const Outer_ & A, const double * b, double * x, MGInfo<float, Layout_> & info,
                      uint64_t n, uint64_t * outer, uint64_t * inner)
    {
        DenseVector<double> bv(dv(b, n)), xv(dv(x, n));
        Outer_ Ac(A);
        std::stringstream captured;
        std::streambuf * old(std::cout.rdbuf(captured.rdbuf()));
        Multigrid<tags::CPU, tags::CPU, Prol_, methods::JAC, methods::CYCLE::V, methods::MIXED>::value(
                Ac, bv, xv, (unsigned long)11, std::numeric_limits<double>::epsilon(), info);
        std::cout.rdbuf(old);
        out(x, xv);
        std::string line;
        *outer = 0;
        *inner = 0;
        while (std::getline(captured, line))
        {
            if (line.find("TN of outer iters: ") == 0)
                *outer = std::strtoul(line.c_str() + 19, 0, 10);
            if (line.find("TN of inner iters: ") == 0)
                *inner = std::strtoul(line.c_str() + 19, 0, 10);
        }
    }

    template <typename Prol_, typename DT_, typename Layout_>
    void mg_fixed_run(Layout_ & A, const DT_ * b, DT_ * x, MGInfo<DT_, Layout_> & info, uint64_t n)
    {
        DenseVector<DT_> bv(dv(b, n)), xv(dv(x, n));
        info.rhs[info.max_level] = bv;
        std::stringstream captured;
        std::streambuf * old(std::cout.rdbuf(captured.rdbuf()));
        Multigrid<tags::CPU, tags::CPU, Prol_, methods::JAC, methods::CYCLE::V, methods::FIXED>::value(
                A, bv, xv, (unsigned long)11, std::numeric_limits<DT_>::epsilon(), info);
        std::cout.rdbuf(old);
        out(x, xv);
    }

    template <typename DT_, typename Layout_>
    int mg_fixed_any(int prolmat, const void * const * a_q1, const ref_ell * a_ell, const ref_ell * prol,
                     const ref_ell * res, const void * const * dinv, const DT_ * b, DT_ * x, uint64_t min_level,
                     uint64_t max_level, uint64_t n_max_iter, double tolerance, uint64_t n_pre, double adapt,
                     const uint64_t * mask, int initial_zero, int convergence_check)
    {
        MGInfo<DT_, Layout_> info;
        mg_info(info, a_q1, a_ell, prol, res, dinv, min_level, max_level, n_max_iter, tolerance, n_pre, adapt, mask,
                initial_zero, convergence_check);
        const uint64_t n = info.a[max_level].rows();
        Layout_ A(info.a[max_level]);
        if (prolmat)
            mg_fixed_run<methods::PROLMAT>(A, b, x, info, n);
        else
            mg_fixed_run<methods::NONE>(A, b, x, info, n);
        return 0;
    }
}

extern "C" {

/* kind 0: Q1 (a32_q1 bands [9 L + b], A64_q1 finest FP64 bands); kind 1: ELL
 * (a32_ell[L], A64_ell).  prol/res: PROLMAT matrices per level, or NULL. */
int ref_mg_mixed(int kind, int prolmat, const void * const * A64_q1, const ref_ell * A64_ell, const double * b,
                 double * x, const void * const * a32_q1, const ref_ell * a32_ell, const ref_ell * prol32,
                 const ref_ell * res32, const void * const * dinv32, uint64_t min_level, uint64_t max_level,
                 uint64_t n_max_iter, double tolerance, uint64_t n_pre, double adapt, const uint64_t * mask,
                 int initial_zero, int convergence_check, uint64_t * outer, uint64_t * inner)
{
    if (kind == 0)
    {
        MGInfo<float, BandedMatrixQx<Q1Type, float> > info;
        mg_info(info, a32_q1, (const ref_ell *)0, prol32, res32, dinv32, min_level, max_level, n_max_iter, tolerance,
                n_pre, adapt, mask, initial_zero, convergence_check);
        const unsigned long n = ((1ul << max_level) + 1) * ((1ul << max_level) + 1);
        BandedMatrixQx<Q1Type, double> A(q1((const double * const *)A64_q1, n));
        if (prolmat)
            mg_mixed_run<methods::PROLMAT>(A, b, x, info, n, outer, inner);
        else
            mg_mixed_run<methods::NONE>(A, b, x, info, n, outer, inner);
    }
    else
    {
        MGInfo<float, SparseMatrixELL<float> > info;
        mg_info(info, (const void * const *)0, a32_ell, prol32, res32, dinv32, min_level, max_level, n_max_iter,
                tolerance, n_pre, adapt, mask, initial_zero, convergence_check);
        SparseMatrixELL<double> A(ell_of<double>(*A64_ell));
        if (prolmat)
            mg_mixed_run<methods::PROLMAT>(A, b, x, info, A64_ell->rows, outer, inner);
        else
            mg_mixed_run<methods::NONE>(A, b, x, info, A64_ell->rows, outer, inner);
    }
    return 0;
}

int ref_mg_fixed_f32(int kind, int prolmat, const void * const * a_q1, const ref_ell * a_ell, const ref_ell * prol,
                     const ref_ell * res, const void * const * dinv, const float * b, float * x, uint64_t min_level,
                     uint64_t max_level, uint64_t n_max_iter, double tolerance, uint64_t n_pre, double adapt,
                     const uint64_t * mask, int initial_zero, int convergence_check)
{
    if (kind == 0)
        return mg_fixed_any<float, BandedMatrixQx<Q1Type, float> >(prolmat, a_q1, a_ell, prol, res, dinv, b, x,
                min_level, max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check);
    return mg_fixed_any<float, SparseMatrixELL<float> >(prolmat, a_q1, a_ell, prol, res, dinv, b, x, min_level,
            max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check);
}

int ref_mg_fixed_f64(int kind, int prolmat, const void * const * a_q1, const ref_ell * a_ell, const ref_ell * prol,
                     const ref_ell * res, const void * const * dinv, const double * b, double * x, uint64_t min_level,
                     uint64_t max_level, uint64_t n_max_iter, double tolerance, uint64_t n_pre, double adapt,
                     const uint64_t * mask, int initial_zero, int convergence_check)
{
    if (kind == 0)
        return mg_fixed_any<double, BandedMatrixQx<Q1Type, double> >(prolmat, a_q1, a_ell, prol, res, dinv, b, x,
                min_level, max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check);
    return mg_fixed_any<double, SparseMatrixELL<double> >(prolmat, a_q1, a_ell, prol, res, dinv, b, x, min_level,
            max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check);
}

/* The ell::threads value the reference's ELL constructor reads. */
void ref_set_ell_threads(int threads)
{
    Configuration::instance()->set_value("ell::threads", threads);
}

/* The reference's own SSE ELL product kernel (honei/backends/sse/product.cc,
 * product_smell_dv) on rows [row_start, row_end): what Product<tags::CPU::SSE>
 * runs, and what each thread of Product<tags::CPU::MultiCore::SSE> runs on
 * its equal share of the rows (honei/la/product.hh:2087-2106).  Timed as the
 * CPU baseline of the ELL SpMV. */
void ref_sse_product_smell_dv_f64(double * y, const uint64_t * Aj, const double * Ax, const uint64_t * Arl,
                                  const double * x, uint64_t stride, uint64_t rows, uint64_t ncpr,
                                  uint64_t row_start, uint64_t row_end, uint64_t threads)
{
    honei::sse::product_smell_dv(y, (const unsigned long *)Aj, Ax, (const unsigned long *)Arl, x, stride, rows,
                                 ncpr, row_start, row_end, threads);
}

}
