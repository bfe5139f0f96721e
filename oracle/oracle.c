/*
 * oracle.c -- CPU restatement of HONEI's tags::CPU sparse-solver path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP backend.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
 * the product library.  Built with -ffp-contract=off (see Makefile).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_T double
#define OR_S f64
#define OR_IS_FLOAT 0
#include "oracle_impl.h"
#undef OR_T
#undef OR_S
#undef OR_IS_FLOAT

#define OR_T float
#define OR_S f32
#define OR_IS_FLOAT 1
#include "oracle_impl.h"
#undef OR_T
#undef OR_S
#undef OR_IS_FLOAT

/* honei/la/sparse_matrix_ell-impl.hh:85-101 */
void or_ell_shape(uint64_t rows, const uint64_t *rowptr, uint64_t threads,
                  uint64_t *stride, uint64_t *num_cols_per_row)
{
    uint64_t ncpr = 1;
    for (uint64_t i = 0; i < rows; ++i) {
        uint64_t used = rowptr[i + 1] - rowptr[i];
        if (used > ncpr)
            ncpr = used;
    }
    *num_cols_per_row = (uint64_t)ceil((double)ncpr / (double)threads);
    const uint64_t alignment = 32;
    *stride = alignment * (((rows * threads) + alignment - 1) / alignment);
}

/* SplitMix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom
 * number generators", OOPSLA 2014). */
static uint64_t or_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t or_rng_bits(uint64_t seed, uint64_t stream, uint64_t index)
{
    const uint64_t gamma = 0x9E3779B97F4A7C15ull;
    uint64_t key = or_mix64(seed + gamma * (stream + 1));
    return or_mix64(key + gamma * (index + 1));
}

double or_rng_uniform(uint64_t seed, uint64_t stream, uint64_t index, double lo, double hi)
{
    double u = (double)(or_rng_bits(seed, stream, index) >> 11) * 0x1.0p-53;
    return lo + (hi - lo) * u;
}

/* TypeTraits<double>::convert / TypeTraits<float>::convert
 * (honei/util/type_traits.cc:107-123): plain C casts, round to nearest. */
void or_convert_d2f(float *dst, const double *src, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        dst[i] = (float)src[i];
}

void or_convert_f2d(double *dst, const float *src, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        dst[i] = (double)src[i];
}

/* Multigrid<...MIXED>::value: honei/math/multigrid.hh:2253-2393 (Q1) and
 * 2400-2543 (ELL), the same outer loop: x0 with b injected on the first row,
 * the last row and the first column; result = 0; d = b - A x0 (product then
 * difference); loop { rhs32 = (float)d; inner FP32 solve; result += scale *
 * x32; d = b - A result; stop at |d| < eps |d0|; scale = |d|; d /= scale }.
 * The inner FP32 iterate x32 carries over from one outer step to the next. */
int or_mg_mixed(int kind, const double *const *A64_q1, const or_ell_f64 *A64_ell, const double *b,
                double *x, const float *const *a32_q1, const or_ell_f32 *a32_ell,
                const float *const *dinv32, const or_ell_f32 *prol32, const or_ell_f32 *res32,
                const or_mg_params *p, uint64_t *outer_iters, uint64_t *inner_iters,
                double *final_defect, double *initial_defect_out)
{
    if (p->max_level < p->min_level || p->max_level > 14 || p->n_pre_smooth < 1)
        return -1;
    const uint64_t maxl = p->max_level;
    /* MGInfo<float, ...> stores tolerance and the correction factor as float */
    const float tol = (float)p->tolerance;
    const float adapt = (float)p->adapt_correction_factor;
    or_mgl_f32 lv[15];
    or_mg_levels_f32(lv, kind, a32_q1, a32_ell, dinv32, prol32, res32, p);
    or_mgl_f32 *top = &lv[maxl];
    const uint64_t n = top->n;
    or_matrix_f64 A64;
    if (kind == 0) {
        or_matrix_f64 m = {0, A64_q1, NULL, NULL, NULL, 0, 0};
        A64 = m;
    } else {
        or_matrix_f64 m = {1, NULL, A64_ell->Aj, A64_ell->Ax, A64_ell->Arl, A64_ell->stride, A64_ell->threads};
        A64 = m;
    }

    double *result = (double *)calloc(n, sizeof(double));
    double *ig = (double *)malloc(sizeof(double) * n);
    double *od = (double *)malloc(sizeof(double) * n);
    double *prod = (double *)malloc(sizeof(double) * n);
    double *x_outer = (double *)malloc(sizeof(double) * n);
    memcpy(ig, x, sizeof(double) * n);

    uint64_t sqrt_N = (uint64_t)sqrt((double)n);
    for (uint64_t i = 0; i < sqrt_N; ++i) ig[i] = b[i];
    for (uint64_t i = n - sqrt_N; i < n; ++i) ig[i] = b[i];
    for (uint64_t i = 0; i < n; i += sqrt_N) ig[i] = b[i];
    or_convert_d2f(top->x, ig, n);

    uint64_t inner_total = 0, outer = 0;
    double scale_factor = 1.0;
    or_apply_f64(&A64, prod, ig, n);
    memcpy(od, b, sizeof(double) * n);
    or_difference_f64(od, prod, n);
    double initial_defect = or_norm_l2_f64(od, n, 1);
    double def_norm = initial_defect;
    const double outer_eps = (double)tol;
    int status = 0;

    while (1) {
        if (p->max_outer && outer >= p->max_outer) {
            status = 1;
            break;
        }
        or_convert_d2f(top->rhs, od, n);
        uint64_t inner = or_mg_inner_f32(lv, p, tol, adapt, NULL, NULL);
        ++outer;
        inner_total += inner;
        or_convert_f2d(x_outer, top->x, n);
        or_scaled_sum_f64(result, x_outer, scale_factor, n);
        or_defect_any_f64(&A64, od, b, result, n);
        def_norm = or_norm_l2_f64(od, n, 1);
        if (def_norm < outer_eps * initial_defect)
            break;
        scale_factor = def_norm;
        double inv = 1.0 / scale_factor;
        or_scale_f64(od, inv, n);
    }
    memcpy(x, result, sizeof(double) * n);
    *outer_iters = outer;
    *inner_iters = inner_total;
    if (final_defect) *final_defect = def_norm;
    if (initial_defect_out) *initial_defect_out = initial_defect;
    or_mg_free_f32(lv, p);
    free(result); free(ig); free(od); free(prod); free(x_outer);
    return status;
}

int or_mg_mixed_q1(const double *const *A64, const double *b, double *x,
                   const float *const *a32, const float *const *dinv32,
                   const or_mg_params *p, uint64_t *outer_iters, uint64_t *inner_iters,
                   double *final_defect, double *initial_defect_out)
{
    return or_mg_mixed(0, A64, NULL, b, x, a32, NULL, dinv32, NULL, NULL, p, outer_iters, inner_iters,
                       final_defect, initial_defect_out);
}

/* IRSolver<OuterTag, InnerTag, Norm<vnt_l_two, true, OuterTag>>::value
 * (honei/math/ir.hh:30-79). */
int or_ir_mgop(const or_mgop_params *p, const or_ell_f64 *A64, const double *b, double *x,
               const or_ell_f32 *A, const float *const *P, const or_ell_f32 *prol,
               const or_ell_f32 *res, float *x_inner, double eps, uint64_t max_iters,
               uint64_t *used_iters, uint64_t *inner_used, uint64_t *coarse_used)
{
    const uint64_t n = A64->rows;
    if (p->levels < 1 || A[p->levels - 1].rows != n)
        return -1;
    double *r = (double *)malloc(sizeof(double) * n), *c = (double *)malloc(sizeof(double) * n);
    float *b_inner = (float *)malloc(sizeof(float) * n);
    double rnorm_initial = 1e16, rnorm_current = 1e16;
    int status = 0;
    for (uint64_t i = 0; i < max_iters; ++i) {
        or_ell_defect_f64(r, b, A64->Aj, A64->Ax, A64->Arl, x, n, A64->stride, A64->threads);
        if (i == 0)
            rnorm_initial = or_norm_l2_f64(r, n, 1);
        else {
            rnorm_current = or_norm_l2_f64(r, n, 1);
            *used_iters = i;
            if (rnorm_current < eps * rnorm_initial)
                break;
        }
        or_convert_d2f(b_inner, r, n);
        status = or_mgop_solve_f32(p, A, P, prol, res, b_inner, x_inner, inner_used, coarse_used, NULL);
        if (status != 0)
            break;
        or_convert_f2d(c, x_inner, n);
        or_sum_f64(x, c, n);
    }
    free(r);
    free(c);
    free(b_inner);
    return status;
}
