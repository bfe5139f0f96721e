/*
 * oracle.h -- CPU restatement of HONEI's tags::CPU sparse-solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * backend: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product library never links or calls it.
 *
 * Every function restates the loop order of the cited reference function
 * (paths relative to the HONEI source tree) so that summation order, and
 * therefore rounding, is the reference's own.  The build uses
 * -ffp-contract=off so that no multiply-add is fused, as on the reference's
 * x86-64 SSE2 build.
 *
 * Band order of Q1 matrices everywhere: LL, LD, LU, DL, DD, DU, UL, UD, UU
 * (honei/la/band_type.hh:30-38), offsets -m-1, -m, -m+1, -1, 0, 1, m-1, m, m+1.
 *
 * Functions whose results no fixture, closed form or compiled-reference run
 * checks are marked "parity unpinned" below; at present there are none.
 */
#ifndef HSOLVE_ORACLE_H
#define HSOLVE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of Multigrid<..., JAC, CYCLE::V, MIXED> (MGInfo fields,
 * honei/math/multigrid.hh:65-149). */
typedef struct or_mg_params {
    uint64_t min_level;
    uint64_t max_level;
    uint64_t n_max_iter;          /* V-cycles per inner solve */
    int32_t initial_zero;
    int32_t convergence_check;
    double tolerance;             /* inner relative tolerance and outer eps */
    uint64_t n_pre_smooth;        /* also the post-smoothing count (multigrid.hh:1851) */
    double adapt_correction_factor;
    uint64_t macro_border_mask[8];
    uint64_t max_outer;           /* safety cap: 0 = none (reference has none) */
} or_mg_params;

/* Operator-list multigrid (honei/math/mg.hh): MGCycleCreation shape, coarse
 * solver and MGData configuration (MGUtil::configure, mg.hh:434-447). */
typedef struct or_mgop_params {
    int32_t cycle;                /* 0: CYCLE::V, 1: CYCLE::W, 2: CYCLE::F::V (all STATIC) */
    int32_t coarse_var;           /* 0: CGSolver<NONE>, 1: CGSolver<VAR> with the level's P */
    uint64_t levels;              /* max level; levels 1..levels (internal index 0..levels-1) */
    uint64_t min_level;
    uint64_t max_iters, max_iters_coarse;
    uint64_t n_pre_smooth, n_post_smooth;
    double eps_relative;          /* stored as the data type, like MGData::eps_relative */
} or_mgop_params;

/* ---- shape helpers (type independent) ---------------------------------- */
/* SparseMatrixELL(SparseMatrix) shape: honei/la/sparse_matrix_ell-impl.hh:78-101 */
void or_ell_shape(uint64_t rows, const uint64_t *rowptr, uint64_t threads,
                  uint64_t *stride, uint64_t *num_cols_per_row);

/* ---- SplitMix64 counter RNG shared by all synthetic inputs ------------- */
uint64_t or_rng_bits(uint64_t seed, uint64_t stream, uint64_t index);
double or_rng_uniform(uint64_t seed, uint64_t stream, uint64_t index, double lo, double hi);

#define OR_ELL_TYPE(T, S)                                                                 \
typedef struct or_ell_##S {                                                              \
    uint64_t rows, cols, stride, threads;                                                \
    const uint64_t *Aj;                                                                  \
    const T *Ax;                                                                         \
    const uint64_t *Arl;                                                                 \
} or_ell_##S;
OR_ELL_TYPE(double, f64)
OR_ELL_TYPE(float, f32)

#define OR_DECLARE(T, S)                                                                  \
void or_ell_from_csr_##S(uint64_t rows, const uint64_t *rowptr, const uint64_t *col,     \
                         const T *val, uint64_t threads, uint64_t stride,                \
                         uint64_t num_cols_per_row, uint64_t *Aj, T *Ax, uint64_t *Arl); \
void or_ell_row_length_##S(uint64_t rows, uint64_t stride, uint64_t threads,             \
                           uint64_t size, const T *Ax, uint64_t *Arl);                   \
void or_ell_product_##S(T *y, const uint64_t *Aj, const T *Ax, const uint64_t *Arl,      \
                        const T *x, uint64_t row_start, uint64_t row_end,                \
                        uint64_t stride, uint64_t threads);                              \
void or_ell_defect_##S(T *r, const T *rhs, const uint64_t *Aj, const T *Ax,              \
                       const uint64_t *Arl, const T *x, uint64_t rows,                   \
                       uint64_t stride, uint64_t threads);                               \
void or_csr_from_rows_##S(uint64_t rows, const uint64_t *rowptr, const uint64_t *col,    \
                          const T *val, uint64_t bs, uint64_t *Ar, uint64_t *Aj, T *Ax, \
                          uint64_t *nblocks, uint64_t *used);                           \
void or_csr_product_##S(T *y, const uint64_t *Ar, const uint64_t *Aj, const T *Ax,      \
                        const T *x, uint64_t columns, uint64_t bs,                      \
                        uint64_t row_start, uint64_t row_end);                          \
void or_csr_defect_##S(T *r, const T *rhs, const uint64_t *Ar, const uint64_t *Aj,      \
                       const T *Ax, const T *x, uint64_t columns, uint64_t bs,          \
                       uint64_t row_start, uint64_t row_end);                           \
void or_q1_product_##S(T *y, const T *const *bands, const T *x, uint64_t n);             \
void or_q1_defect_##S(T *r, const T *rhs, const T *const *bands, const T *x, uint64_t n);\
T or_dot_##S(const T *x, const T *y, uint64_t n);                                        \
T or_norm_l2_##S(const T *x, uint64_t n, int32_t root);                                  \
void or_scaled_sum_##S(T *x, const T *y, T b, uint64_t n);                               \
void or_scaled_sum3_##S(T *r, const T *x, const T *y, T b, uint64_t n);                  \
void or_scaled_sum_vv_##S(T *a, const T *b, const T *c, uint64_t n);                     \
void or_sum_##S(T *x, const T *y, uint64_t n);                                           \
void or_scale_##S(T *x, T a, uint64_t n);                                                \
void or_scale_by_double_##S(T *x, double a, uint64_t n);                                 \
void or_difference_##S(T *x, const T *y, uint64_t n);                                    \
void or_difference3_##S(T *r, const T *x, const T *y, uint64_t n);                       \
void or_element_product_##S(T *a, const T *b, uint64_t n);                               \
void or_element_product3_##S(T *r, const T *a, const T *b, uint64_t n);                  \
void or_element_inverse_##S(T *x, uint64_t n);                                           \
void or_apply_dirichlet_##S(T *grid, uint64_t n, const uint64_t *mask);                  \
void or_restriction_##S(T *coarse, const T *fine, uint64_t n_fine, const uint64_t *mask);\
void or_prolongation_##S(T *fine, const T *coarse, uint64_t n_fine, const uint64_t *mask);\
void or_jacobi_q1_##S(T *x, const T *const *bands, const T *rhs, const T *dinv,          \
                      uint64_t n, uint64_t iters, T *temp);                              \
void or_jacobi_ell_##S(T *x, const uint64_t *Aj, const T *Ax, const uint64_t *Arl,       \
                       uint64_t stride, uint64_t threads, const T *rhs, const T *dinv,   \
                       uint64_t n, uint64_t iters, T *temp);                             \
void or_fill_poisson_dirichlet0_##S(T *const *bands, uint64_t n);                        \
void or_fill_poisson_dirichlet_neumann_##S(T *const *bands, uint64_t n);                 \
void or_fill_vector_dirichlet_neumann_##S(T *v, uint64_t n);                             \
int or_cg_none_q1_##S(const T *const *bands, const T *b, T *x, uint64_t n,              \
                      uint64_t max_iters, T eps, uint64_t *used_iters);                  \
int or_cg_none_ell_##S(const uint64_t *Aj, const T *Ax, const uint64_t *Arl,             \
                       uint64_t stride, uint64_t threads, const T *b, T *x, uint64_t n,  \
                       uint64_t max_iters, T eps, uint64_t *used_iters);                 \
int or_cg_var_ell_##S(const uint64_t *Aj, const T *Ax, const uint64_t *Arl,              \
                      uint64_t stride, uint64_t threads, const T *P, const T *b, T *x,   \
                      uint64_t n, uint64_t max_iters, T eps, uint64_t *used_iters);      \
int or_pcg_jac_ell_##S(const uint64_t *Aj, const T *Ax, const uint64_t *Arl,             \
                       uint64_t stride, uint64_t threads, const T *dd_inv, const T *rhs, \
                       T *x, uint64_t n, uint64_t max_iters, T eps, uint64_t *used_iters);\
int or_cg_konv_q1_##S(const T *const *bands, const T *rhs, T *x, uint64_t n,            \
                      double konv_rad, uint64_t max_iters, uint64_t *iters);             \
int or_cg_konv_ell_##S(const or_ell_##S *M, const T *rhs, T *x, double konv_rad,          \
                       uint64_t max_iters, uint64_t *iters);                             \
int or_mg_fixed_##S(int kind, const T *const *a_q1, const or_ell_##S *a_ell,             \
                    const T *const *dinv, const or_ell_##S *prol, const or_ell_##S *res, \
                    const T *b, T *x, const or_mg_params *p, uint64_t *iters,            \
                    T *final_defect, T *initial_defect);                                 \
int or_mgop_solve_##S(const or_mgop_params *p, const or_ell_##S *A, const T *const *P,  \
                      const or_ell_##S *prol, const or_ell_##S *res, const T *b, T *x,   \
                      uint64_t *used_iters, uint64_t *used_iters_coarse, T *final_norm); \
int or_mgop_solve_pm_##S(const or_mgop_params *p, const or_ell_##S *A, const or_ell_##S *Pm,\
                         const or_ell_##S *prol, const or_ell_##S *res, const T *b, T *x,\
                         uint64_t *used_iters, uint64_t *used_iters_coarse, T *final_norm);\
int or_cg_var_ell_pm_##S(const or_ell_##S *A, const or_ell_##S *P, const T *b, T *x,     \
                         uint64_t max_iters, T eps, uint64_t *used_iters);               \
int or_cg_iters_q1_##S(const T *const *bands, const T *rhs, T *x, uint64_t n,           \
                       uint64_t iters, uint64_t *used);                                  \
int or_cg_iters_ell_##S(const or_ell_##S *M, const T *rhs, T *x, uint64_t max_iters,     \
                        uint64_t *used);

OR_DECLARE(double, f64)
OR_DECLARE(float, f32)

void or_convert_d2f(float *dst, const double *src, uint64_t n);
void or_convert_f2d(double *dst, const float *src, uint64_t n);

/* Multigrid<Tag,OuterTag,NONE,JAC,CYCLE::V,MIXED>::value on Q1 bands
 * (honei/math/multigrid.hh:1616-1933 inner kernel, 2253-2393 outer loop).
 *   A64      : 9 finest-level FP64 bands, size n = (2^max_level+1)^2
 *   b, x     : FP64 right-hand side and initial guess / result
 *   a32      : FP32 bands per level, a32[level*9 + band], levels 0..max_level
 *   dinv32   : FP32 inverted, omega-scaled diagonal per level
 * Returns 0 on convergence, 1 if max_outer was reached, <0 on bad input.   */
int or_mg_mixed_q1(const double *const *A64, const double *b, double *x,
                   const float *const *a32, const float *const *dinv32,
                   const or_mg_params *p, uint64_t *outer_iters, uint64_t *inner_iters,
                   double *final_defect, double *initial_defect);

/* IRSolver<OuterTag, InnerTag, Norm<vnt_l_two, true>>::value (honei/math/ir.hh:
 * 30-79): FP64 defect and update around MGSolver on FP32 operator-list data.
 * x_inner is the inner finest iterate (MGData::x of the top level), which the
 * reference keeps from one outer step to the next: in = its initial value,
 * out = its last value.  used_iters as the reference sets it. */
int or_ir_mgop(const or_mgop_params *p, const or_ell_f64 *A64, const double *b, double *x,
               const or_ell_f32 *A, const float *const *P, const or_ell_f32 *prol,
               const or_ell_f32 *res, float *x_inner, double eps, uint64_t max_iters,
               uint64_t *used_iters, uint64_t *inner_used, uint64_t *coarse_used);

/* Multigrid<..., ProlType, JAC, CYCLE::V, MIXED> on Q1 bands (kind 0:
 * A64_q1 finest FP64 bands, a32_q1[9 L + b]) or ELL matrices (kind 1:
 * A64_ell finest, a32_ell[L]); dinv32[L]; prol32[L] / res32[L] for
 * Prolongation/Restriction<PROLMAT> (NULL: structured transfers with the
 * border mask).  Outer loop: multigrid.hh:2253-2393 (Q1), 2400-2543 (ELL). */
int or_mg_mixed(int kind, const double *const *A64_q1, const or_ell_f64 *A64_ell, const double *b,
                double *x, const float *const *a32_q1, const or_ell_f32 *a32_ell,
                const float *const *dinv32, const or_ell_f32 *prol32, const or_ell_f32 *res32,
                const or_mg_params *p, uint64_t *outer_iters, uint64_t *inner_iters,
                double *final_defect, double *initial_defect);

#ifdef __cplusplus
}
#endif
#endif
