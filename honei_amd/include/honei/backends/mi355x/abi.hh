// honei/backends/mi355x/abi.hh -- float/double overloads over the C-ABI,
// so the op templates can be written once for both value types.
#pragma once

#include <honei/backends/mi355x/hsolve.hh>

#include <cstdint>

namespace honei
{
    namespace mi355x
    {
        namespace abi
        {
#define HS_PAIR(NAME, ...)                                                                     \
            inline int NAME(float, __VA_ARGS__);                                               \
            inline int NAME(double, __VA_ARGS__);

            // ELL in HONEI's layout (uint64 Aj / Arl)
            inline int product_smell(const float * x, float * y, const void * Aj, const float * Ax, const void * Arl,
                                     uint64_t rs, uint64_t re, uint64_t ncpr, uint64_t stride, uint64_t threads)
            { return hsolve_product_smell_dv_float(x, y, Aj, Ax, Arl, rs, re, ncpr, stride, threads, stream()); }
            inline int product_smell(const double * x, double * y, const void * Aj, const double * Ax, const void * Arl,
                                     uint64_t rs, uint64_t re, uint64_t ncpr, uint64_t stride, uint64_t threads)
            { return hsolve_product_smell_dv_double(x, y, Aj, Ax, Arl, rs, re, ncpr, stride, threads, stream()); }
            inline int defect_smell(const float * rhs, float * r, const void * Aj, const float * Ax, const void * Arl,
                                    const float * b, uint64_t rs, uint64_t re, uint64_t ncpr, uint64_t stride,
                                    uint64_t threads)
            { return hsolve_defect_smell_dv_float(rhs, r, Aj, Ax, Arl, b, rs, re, ncpr, stride, threads, stream()); }
            inline int defect_smell(const double * rhs, double * r, const void * Aj, const double * Ax,
                                    const void * Arl, const double * b, uint64_t rs, uint64_t re, uint64_t ncpr,
                                    uint64_t stride, uint64_t threads)
            { return hsolve_defect_smell_dv_double(rhs, r, Aj, Ax, Arl, b, rs, re, ncpr, stride, threads, stream()); }

            // Q1 bands (9 device pointers LL..UU)
            inline int product_q1(const float * const * b, const float * x, float * y, uint64_t n, uint64_t m)
            { return hsolve_product_bmdv_q1_float(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], x, y, n, m, stream()); }
            inline int product_q1(const double * const * b, const double * x, double * y, uint64_t n, uint64_t m)
            { return hsolve_product_bmdv_q1_double(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], x, y, n, m, stream()); }
            inline int defect_q1(const float * rhs, const float * const * b, const float * x, float * y, uint64_t n, uint64_t m)
            { return hsolve_defect_q1_float(rhs, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], x, y, n, m, stream()); }
            inline int defect_q1(const double * rhs, const double * const * b, const double * x, double * y, uint64_t n,
                                 uint64_t m)
            { return hsolve_defect_q1_double(rhs, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], x, y, n, m, stream()); }

            // reductions returning host values
            inline int dot(const float * a, const float * b, uint64_t n, float * r)
            { return hsolve_dot_product_two_float_host(a, b, n, r, stream()); }
            inline int dot(const double * a, const double * b, uint64_t n, double * r)
            { return hsolve_dot_product_two_double_host(a, b, n, r, stream()); }
            inline int norm(const float * a, uint64_t n, int root, float * r)
            { return hsolve_norm_l2_one_float_host(a, n, root, r, stream()); }
            inline int norm(const double * a, uint64_t n, int root, double * r)
            { return hsolve_norm_l2_one_double_host(a, n, root, r, stream()); }

            // BLAS-1
            inline int scaled_sum_s(float * r, const float * x, const float * y, float b, uint64_t n)
            { return hsolve_scaled_sum_three_float_s(r, x, y, b, n, stream()); }
            inline int scaled_sum_s(double * r, const double * x, const double * y, double b, uint64_t n)
            { return hsolve_scaled_sum_three_double_s(r, x, y, b, n, stream()); }
            inline int scaled_sum_v(float * x, const float * y, const float * z, uint64_t n)
            { return hsolve_scaled_sum_three_float(x, y, z, n, stream()); }
            inline int scaled_sum_v(double * x, const double * y, const double * z, uint64_t n)
            { return hsolve_scaled_sum_three_double(x, y, z, n, stream()); }
            inline int sum(float * a, const float * b, uint64_t n) { return hsolve_sum_two_float(a, b, n, stream()); }
            inline int sum(double * a, const double * b, uint64_t n) { return hsolve_sum_two_double(a, b, n, stream()); }
            inline int scale(float * x, float a, uint64_t n) { return hsolve_scale_one_float(x, a, n, stream()); }
            inline int scale(double * x, double a, uint64_t n) { return hsolve_scale_one_double(x, a, n, stream()); }
            inline int difference(float * a, const float * b, uint64_t n)
            { return hsolve_difference_two_float(a, b, n, stream()); }
            inline int difference(double * a, const double * b, uint64_t n)
            { return hsolve_difference_two_double(a, b, n, stream()); }
            inline int difference3(float * r, const float * a, const float * b, uint64_t n)
            { return hsolve_difference_three_float(r, a, b, n, stream()); }
            inline int difference3(double * r, const double * a, const double * b, uint64_t n)
            { return hsolve_difference_three_double(r, a, b, n, stream()); }
            inline int element_product3(float * r, const float * a, const float * b, uint64_t n)
            { return hsolve_element_product_three_float(r, a, b, n, stream()); }
            inline int element_product3(double * r, const double * a, const double * b, uint64_t n)
            { return hsolve_element_product_three_double(r, a, b, n, stream()); }
            inline int element_inverse(float * x, uint64_t n) { return hsolve_element_inverse_one_float(x, n, stream()); }
            inline int element_inverse(double * x, uint64_t n) { return hsolve_element_inverse_one_double(x, n, stream()); }
            inline int fill(float * x, float v, uint64_t n) { return hsolve_fill_float(x, v, n, stream()); }
            inline int fill(double * x, double v, uint64_t n) { return hsolve_fill_double(x, v, n, stream()); }

            // transfers
            inline int prolongation(float * f, uint64_t nf, const float * c, uint64_t nc, const uint64_t * mask)
            { return hsolve_prolongation_float(f, nf, c, nc, mask, stream()); }
            inline int prolongation(double * f, uint64_t nf, const double * c, uint64_t nc, const uint64_t * mask)
            { return hsolve_prolongation_double(f, nf, c, nc, mask, stream()); }
            inline int restriction(float * c, uint64_t nc, const float * f, uint64_t nf, const uint64_t * mask)
            { return hsolve_restriction_float(c, nc, f, nf, mask, stream()); }
            inline int restriction(double * c, uint64_t nc, const double * f, uint64_t nf, const uint64_t * mask)
            { return hsolve_restriction_double(c, nc, f, nf, mask, stream()); }

            // smoother and coarse solver
            inline int jacobi_q1(const float * rhs, const float * const * b, const float * dinv, const float * xo,
                                 float * xn, uint64_t n)
            { return hsolve_jacobi_q1_float(rhs, (const void * const *)b, dinv, xo, xn, n, stream()); }
            inline int jacobi_q1(const double * rhs, const double * const * b, const double * dinv, const double * xo,
                                 double * xn, uint64_t n)
            { return hsolve_jacobi_q1_double(rhs, (const void * const *)b, dinv, xo, xn, n, stream()); }
            inline int cg_konv_q1(const float * const * b, const float * rhs, float * x, uint64_t n, double konv,
                                  uint64_t max_iters)
            { return hsolve_cg_konv_q1_float((const void * const *)b, rhs, x, n, konv, max_iters, nullptr, stream()); }
            inline int cg_konv_q1(const double * const * b, const double * rhs, double * x, uint64_t n, double konv,
                                  uint64_t max_iters)
            { return hsolve_cg_konv_q1_double((const void * const *)b, rhs, x, n, konv, max_iters, nullptr, stream()); }
            inline int cg_iters_q1(const float * const * b, const float * rhs, float * x, uint64_t n, uint64_t iters)
            { return hsolve_cg_iters_q1_float((const void * const *)b, rhs, x, n, iters, nullptr, stream()); }
            inline int cg_iters_q1(const double * const * b, const double * rhs, double * x, uint64_t n,
                                   uint64_t iters)
            { return hsolve_cg_iters_q1_double((const void * const *)b, rhs, x, n, iters, nullptr, stream()); }
#undef HS_PAIR
        }
    }
}
