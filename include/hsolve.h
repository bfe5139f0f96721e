/*
 * hsolve.h -- C-ABI of the MI355X (gfx950) backend for HONEI's sparse
 * solver path.
 *
 * Conventions
 *   - Every entry point returns an int status (HSOLVE_OK == 0) and never
 *     throws; hsolve_last_error() gives the message of the last failure on
 *     the calling thread.  Every HIP call inside is checked.
 *   - Vector and matrix arguments are device pointers unless named *_host.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Kernels are enqueued asynchronously; functions that return a host
 *     value synchronise that stream.
 *   - Sizes are uint64_t (the reference's `unsigned long` on LP64).
 *
 * The first block mirrors the reference's CUDA C-ABI
 * (honei/backends/cuda/operations.hh:28-130, transfer.hh, multi_gpu.hh):
 * same names with the `hsolve_` prefix instead of `cuda_`, same argument
 * order and meaning, minus the per-call `blocksize`/`gridsize` tuning
 * arguments (launch shapes are chosen for gfx950 inside) and plus an
 * explicit stream.  Reductions write their result to a device scalar or,
 * in the *_host forms, return it.
 *
 * The second block is new: device-resident matrix handles (32-bit column
 * indices, compact row lengths), fused smoother/transfer kernels and the
 * solver entry points (CG, MG-IR) that the C++ template layer in
 * honei_amd/include implements on top of the kernels.
 */
#ifndef HSOLVE_H
#define HSOLVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum hsolve_status_code {
    HSOLVE_OK = 0,
    HSOLVE_ERR_INVALID = 1,       /* bad argument (null pointer, bad layout) */
    HSOLVE_ERR_SIZE = 2,          /* sizes do not match (VectorSizeDoesNotMatch) */
    HSOLVE_ERR_HIP = 3,           /* a HIP runtime call failed */
    HSOLVE_ERR_ALLOC = 4,         /* device allocation failed (InternalError) */
    HSOLVE_ERR_NOT_CONVERGED = 5, /* solver hit its iteration cap */
    HSOLVE_ERR_BUDGET = 6         /* request exceeds the device-memory budget */
};

const char *hsolve_last_error(void);
const char *hsolve_version(void);

/* ---------------- runtime: transfer.hh:26-60, multi_gpu.hh:24-40 ---------- */
int hsolve_device_count(int *count);                     /* cuda_device_count */
int hsolve_get_device(int *device);                      /* cuda_get_device */
int hsolve_set_device(int device);                       /* cuda_set_device */
int hsolve_malloc(void **ptr, uint64_t bytes);           /* cuda_malloc */
int hsolve_malloc_host(void **ptr, uint64_t bytes);      /* cuda_malloc_host (pinned) */
int hsolve_free(void *ptr);                              /* cuda_free */
int hsolve_free_host(void *ptr);                         /* cuda_free_host */
int hsolve_upload(const void *src_host, void *dst, uint64_t bytes, void *stream);    /* cuda_upload(_async) */
int hsolve_download(const void *src, void *dst_host, uint64_t bytes, void *stream);  /* cuda_download(_async) */
int hsolve_copy(const void *src, void *dst, uint64_t bytes, void *stream);           /* cuda_copy */
int hsolve_fill_zero(void *dst, uint64_t bytes, void *stream);                       /* cuda_fill_zero (kernel, no memset) */
int hsolve_fill_float(void *dst, float value, uint64_t size, void *stream);
int hsolve_fill_double(void *dst, double value, uint64_t size, void *stream);
int hsolve_synchronize(void *stream);                    /* cuda_thread_synchronize */
int hsolve_stream_create(void **stream);
int hsolve_stream_destroy(void *stream);
/* HIP events for timing work on a given stream (no reference counterpart:
 * the reference times with gettimeofday around synchronous calls,
 * benchmark/benchmark.cc:194-400). */
int hsolve_event_create(void **event);
int hsolve_event_destroy(void *event);
int hsolve_event_record(void *event, void *stream);
int hsolve_event_elapsed_ms(void *start, void *stop, float *ms);   /* synchronises stop */
/* Plain device copy kernel (16-byte vector loads/stores), used to calibrate
 * hardware byte counters against a kernel of known traffic. */
int hsolve_copy_kernel(const void *src, void *dst, uint64_t bytes, void *stream);

/* ---------------- operations.hh mirror ------------------------------------- */
/* ELL SpMV y = A x on rows [row_start, row_end); pass row_end = rows for
 * the whole matrix.  An empty range (row_end <= row_start) is a no-op: y is
 * left unchanged and HSOLVE_OK is returned.  HONEI's SMELL layout: element
 * k of lane t of row r at r*threads + t + k*stride; Aj uint64 column
 * indices, Arl uint64 row step counts.
 * Replaces cuda_product_smell_dv_{float,double} (product.cu:356-402). */
int hsolve_product_smell_dv_float(const void *x, void *y, const void *Aj, const void *Ax,
                                  const void *Arl, uint64_t row_start, uint64_t row_end,
                                  uint64_t num_cols_per_row, uint64_t stride, uint64_t threads,
                                  void *stream);
int hsolve_product_smell_dv_double(const void *x, void *y, const void *Aj, const void *Ax,
                                   const void *Arl, uint64_t row_start, uint64_t row_end,
                                   uint64_t num_cols_per_row, uint64_t stride, uint64_t threads,
                                   void *stream);
/* ELL defect result = rhs - A b.  Replaces cuda_defect_smell_dv_* (defect.cu:367-417). */
int hsolve_defect_smell_dv_float(const void *rhs, void *result, const void *Aj, const void *Ax,
                                 const void *Arl, const void *b, uint64_t row_start,
                                 uint64_t row_end, uint64_t num_cols_per_row, uint64_t stride,
                                 uint64_t threads, void *stream);
int hsolve_defect_smell_dv_double(const void *rhs, void *result, const void *Aj, const void *Ax,
                                  const void *Arl, const void *b, uint64_t row_start,
                                  uint64_t row_end, uint64_t num_cols_per_row, uint64_t stride,
                                  uint64_t threads, void *stream);
/* Blocked-CSR SpMV y = A x on rows [row_start, row_end) (an empty range,
 * row_end <= row_start, is a no-op returning HSOLVE_OK) in HONEI's
 * SparseMatrixCSR layout: Ar uint64 block offsets (rows + 1),
 * Aj uint64 block start columns, Ax atomicsize values per block; positions at
 * or past `columns` (the length of x) add nothing.  Row sums in the CPU
 * order, bit-identical to Product<tags::CPU>.  Replaces
 * cuda_product_csr_dv_{float,double} (operations.hh:95-99, product.cu:257-296,
 * :404-447); the CUDA launcher's blocksize is dropped and `columns` is added
 * (the CUDA kernel leans on texture bounds for that check). */
int hsolve_product_csr_dv_float(const void *x, void *y, const void *Aj, const void *Ax,
                                const void *Ar, uint64_t row_start, uint64_t row_end,
                                uint64_t atomicsize, uint64_t columns, void *stream);
int hsolve_product_csr_dv_double(const void *x, void *y, const void *Aj, const void *Ax,
                                 const void *Ar, uint64_t row_start, uint64_t row_end,
                                 uint64_t atomicsize, uint64_t columns, void *stream);
/* Blocked-CSR defect result = rhs - A b on rows [0, rows).  Replaces
 * cuda_defect_csr_dv_{float,double} (operations.hh:49-53, defect.cu:260-301, :419-466). */
int hsolve_defect_csr_dv_float(const void *rhs, void *result, const void *Aj, const void *Ax,
                               const void *Ar, const void *b, uint64_t rows, uint64_t atomicsize,
                               uint64_t columns, void *stream);
int hsolve_defect_csr_dv_double(const void *rhs, void *result, const void *Aj, const void *Ax,
                                const void *Ar, const void *b, uint64_t rows, uint64_t atomicsize,
                                uint64_t columns, void *stream);
/* Q1 banded SpMV y = A x, m = sqrt(size).  Replaces cuda_product_bmdv_q1_*
 * (product.cu:300-354); sums in tags::CPU order DD,LL,LD,LU,DL,DU,UL,UD,UU. */
int hsolve_product_bmdv_q1_float(const void *ll, const void *ld, const void *lu, const void *dl,
                                 const void *dd, const void *du, const void *ul, const void *ud,
                                 const void *uu, const void *x, void *y, uint64_t size, uint64_t m,
                                 void *stream);
int hsolve_product_bmdv_q1_double(const void *ll, const void *ld, const void *lu, const void *dl,
                                  const void *dd, const void *du, const void *ul, const void *ud,
                                  const void *uu, const void *x, void *y, uint64_t size,
                                  uint64_t m, void *stream);
/* Q1 defect y = rhs - A x.  Replaces cuda_defect_q1_* (defect.cu:305-365). */
int hsolve_defect_q1_float(const void *rhs, const void *ll, const void *ld, const void *lu,
                           const void *dl, const void *dd, const void *du, const void *ul,
                           const void *ud, const void *uu, const void *x, void *y, uint64_t size,
                           uint64_t m, void *stream);
int hsolve_defect_q1_double(const void *rhs, const void *ll, const void *ld, const void *lu,
                            const void *dl, const void *dd, const void *du, const void *ul,
                            const void *ud, const void *uu, const void *x, void *y,
                            uint64_t size, uint64_t m, void *stream);
/* Dot product and squared L2 norm (root = 0) or L2 norm (root = 1), one
 * pass over the data plus a one-block finish, deterministic order; result_dev is a device scalar of the vector's
 * type.  Replace cuda_dot_product_two_* (dot_product.cu:77-155) and
 * cuda_norm_l2_one_* (norm.cu:77-155). */
int hsolve_dot_product_two_float(const void *a, const void *b, uint64_t size, void *result_dev,
                                 void *stream);
int hsolve_dot_product_two_double(const void *a, const void *b, uint64_t size, void *result_dev,
                                  void *stream);
int hsolve_norm_l2_one_float(const void *a, uint64_t size, int root, void *result_dev,
                             void *stream);
int hsolve_norm_l2_one_double(const void *a, uint64_t size, int root, void *result_dev,
                              void *stream);
int hsolve_dot_product_two_float_host(const void *a, const void *b, uint64_t size, float *result,
                                      void *stream);
int hsolve_dot_product_two_double_host(const void *a, const void *b, uint64_t size,
                                       double *result, void *stream);
int hsolve_norm_l2_one_float_host(const void *a, uint64_t size, int root, float *result,
                                  void *stream);
int hsolve_norm_l2_one_double_host(const void *a, uint64_t size, int root, double *result,
                                   void *stream);
/* r = x + b*y (r may equal x).  cuda_scaled_sum_three_*_s (scaled_sum.cu:68-99). */
int hsolve_scaled_sum_three_float_s(void *r, const void *x, const void *y, float b,
                                    uint64_t size, void *stream);
int hsolve_scaled_sum_three_double_s(void *r, const void *x, const void *y, double b,
                                     uint64_t size, void *stream);
/* x += y * z.  cuda_scaled_sum_three_* (scaled_sum.cu:102-133). */
int hsolve_scaled_sum_three_float(void *x, const void *y, const void *z, uint64_t size,
                                  void *stream);
int hsolve_scaled_sum_three_double(void *x, const void *y, const void *z, uint64_t size,
                                   void *stream);
/* a += b.  cuda_sum_two_* (sum.cu:48-77). */
int hsolve_sum_two_float(void *a, const void *b, uint64_t size, void *stream);
int hsolve_sum_two_double(void *a, const void *b, uint64_t size, void *stream);
/* x *= a.  cuda_scale_one_* (scale.cu:48-75). */
int hsolve_scale_one_float(void *x, float a, uint64_t size, void *stream);
int hsolve_scale_one_double(void *x, double a, uint64_t size, void *stream);
/* a -= b; r = a - b.  cuda_difference_{two,three}_* (difference.cu:68-131). */
int hsolve_difference_two_float(void *a, const void *b, uint64_t size, void *stream);
int hsolve_difference_two_double(void *a, const void *b, uint64_t size, void *stream);
int hsolve_difference_three_float(void *r, const void *a, const void *b, uint64_t size,
                                  void *stream);
int hsolve_difference_three_double(void *r, const void *a, const void *b, uint64_t size,
                                   void *stream);
/* r = a * b elementwise.  cuda_element_product_three_* (element_product.cu:48-79). */
int hsolve_element_product_three_float(void *r, const void *a, const void *b, uint64_t size,
                                       void *stream);
int hsolve_element_product_three_double(void *r, const void *a, const void *b, uint64_t size,
                                        void *stream);
/* x = 1/x where x != 0.  cuda_element_inverse_one_* (element_inverse.cu:50-75). */
int hsolve_element_inverse_one_float(void *x, uint64_t size, void *stream);
int hsolve_element_inverse_one_double(void *x, uint64_t size, void *stream);
/* Q1 bilinear prolongation / full-weighting restriction on (2^L+1)^2 grids,
 * then zeroing of Dirichlet nodes per the 8-entry host mask (2 = Dirichlet).
 * cuda_prolongation_* (prolongation.cu:191-234), cuda_restriction_*
 * (restriction.cu:240-285); arithmetic follows the tags::CPU loops bit for bit. */
int hsolve_prolongation_float(void *fine, uint64_t size_fine, const void *coarse,
                              uint64_t size_coarse, const uint64_t *macro_border_mask_host,
                              void *stream);
int hsolve_prolongation_double(void *fine, uint64_t size_fine, const void *coarse,
                               uint64_t size_coarse, const uint64_t *macro_border_mask_host,
                               void *stream);
int hsolve_restriction_float(void *coarse, uint64_t size_coarse, const void *fine,
                             uint64_t size_fine, const uint64_t *macro_border_mask_host,
                             void *stream);
int hsolve_restriction_double(void *coarse, uint64_t size_coarse, const void *fine,
                              uint64_t size_fine, const uint64_t *macro_border_mask_host,
                              void *stream);
/* Zero the Dirichlet nodes of a (2^L+1)^2 grid vector per the 8-entry host
 * mask: ApplyDirichletBoundaries<Tag>::value
 * (honei/math/apply_dirichlet_boundaries.hh:31-80; CUDA applies it inside
 * the transfer kernels only). */
int hsolve_apply_dirichlet_float(void *grid, uint64_t size, const uint64_t *macro_border_mask_host,
                                 void *stream);
int hsolve_apply_dirichlet_double(void *grid, uint64_t size, const uint64_t *macro_border_mask_host,
                                  void *stream);
/* FP32 <-> FP64 conversion.  cuda_convert_float_double / _double_float
 * (transfer.cu:110-138). */
int hsolve_convert_float_double(const void *src_float, void *dst_double, uint64_t size,
                                void *stream);
int hsolve_convert_double_float(const void *src_double, void *dst_float, uint64_t size,
                                void *stream);
/* FillMatrix<Tag, POISSON, DIRICHLET_0>::value on a BandedMatrixQx<Q1>
 * (honei/math/fill_matrix.hh:37-76), on the device: bands9 holds the nine
 * device band pointers (LL..UU) of a size = M * M matrix.  Boundary rows
 * become identity rows, interior rows the stencil DD 8/3, others -1/3.
 * The multigrid engine assembles every level with the same kernel. */
int hsolve_fill_poisson_q1_float(void *const *bands9, uint64_t size, void *stream);
int hsolve_fill_poisson_q1_double(void *const *bands9, uint64_t size, void *stream);

/* ---------------- new fused kernels ---------------------------------------- */
/* One damped-Jacobi sweep in one pass: x_new = x_old + dinv * (rhs - A x_old)
 * (Jacobi<Tag>::value MG overload, honei/math/jacobi.hh:438-480); x_new must
 * not alias x_old. */
int hsolve_jacobi_q1_float(const void *rhs, const void *const *bands9, const void *dinv,
                           const void *x_old, void *x_new, uint64_t size, void *stream);
int hsolve_jacobi_q1_double(const void *rhs, const void *const *bands9, const void *dinv,
                            const void *x_old, void *x_new, uint64_t size, void *stream);
/* fine_x += alpha * (P coarse) without materialising P coarse
 * (Prolongation then ScaledSum, honei/math/multigrid.hh:1818-1843). */
int hsolve_prolongation_axpy_float(void *fine_x, uint64_t size_fine, const void *coarse,
                                   uint64_t size_coarse, float alpha,
                                   const uint64_t *macro_border_mask_host, void *stream);
int hsolve_prolongation_axpy_double(void *fine_x, uint64_t size_fine, const void *coarse,
                                    uint64_t size_coarse, double alpha,
                                    const uint64_t *macro_border_mask_host, void *stream);
/* ConjugateGradients<Tag, NONE>::value(Q1, rhs, x, konv_rad) in one
 * workgroup (honei/math/conjugate_gradients.hh:490-526); size <= 16641.
 * iters_dev (uint64 device scalar, may be NULL) receives the iteration
 * count; max_iters is a termination guarantee the reference lacks. */
int hsolve_cg_konv_q1_float(const void *const *bands9, const void *rhs, void *x, uint64_t size,
                            double konv_rad, uint64_t max_iters, void *iters_dev, void *stream);
int hsolve_cg_konv_q1_double(const void *const *bands9, const void *rhs, void *x,
                             uint64_t size, double konv_rad, uint64_t max_iters, void *iters_dev,
                             void *stream);
/* ConjugateGradients<Tag, NONE>::value(Q1, rhs, x, unsigned long iters)
 * (honei/math/conjugate_gradients.hh:526-569) in one workgroup with the
 * sequential sums of tags::CPU, so the iterate is bit-identical to it:
 * x = 0, then up to `iters` steps, stopping once the squared norm of the
 * true defect falls below 1e-8 times its initial value.  size <= 2^20.
 * used_dev (uint64 device scalar, may be NULL) receives the steps taken;
 * like the reference, reaching `iters` is not an error. */
int hsolve_cg_iters_q1_float(const void *const *bands9, const void *rhs, void *x, uint64_t size,
                             uint64_t iters, void *used_dev, void *stream);
int hsolve_cg_iters_q1_double(const void *const *bands9, const void *rhs, void *x, uint64_t size,
                              uint64_t iters, void *used_dev, void *stream);

/* ---------------- device-resident CSR handle -------------------------------- */
typedef struct hsolve_csr_s *hsolve_csr;
/* Uploads a SparseMatrixCSR given in HONEI's host layout (uint64 Ar/Aj,
 * blocksize values per block in Ax); offsets and columns are kept as uint32
 * when the element count and columns + blocksize fit.  dtype: 0 = float,
 * 1 = double.  used_elements is stored for hsolve_csr_info only.  Backs
 * SparseMatrixCSR of the template layer (sparse_matrix_csr.hh:55-160). */
int hsolve_csr_create(int dtype, uint64_t rows, uint64_t columns, uint64_t blocksize,
                      const uint64_t *Ar_host, const uint64_t *Aj_host, const void *Ax_host,
                      uint64_t used_elements, hsolve_csr *out);
int hsolve_csr_destroy(hsolve_csr m);
int hsolve_csr_info(hsolve_csr m, uint64_t *rows, uint64_t *columns, uint64_t *blocksize,
                    uint64_t *nblocks, uint64_t *used_elements, uint64_t *device_bytes);
/* y = A x, or y = rhs - A x when rhs != NULL, on rows [row_start, row_end)
 * (row_end 0 = all rows); other rows of y are left as they are.  With
 * blocksize 1, rows of at most 16 entries and at most 50% padding, the
 * upload also builds an ELL handle of the same matrix (each row's entries
 * as its slots, in order) and full-range products run on it; results are
 * identical either way. */
int hsolve_csr_spmv(hsolve_csr m, const void *x, void *y, const void *rhs, uint64_t row_start,
                    uint64_t row_end, void *stream);
/* 0 = automatic, 1 = always the CSR kernel (tuning knob, no reference
 * counterpart); hsolve_csr_form reports 1 when the handle holds the ELL form. */
int hsolve_csr_set_variant(hsolve_csr m, int variant);
int hsolve_csr_form(hsolve_csr m, int *form);

/* ---------------- device-resident ELL handle -------------------------------- */
typedef struct hsolve_ell_s *hsolve_ell;
/* Uploads a SparseMatrixELL given in HONEI's host layout (uint64 Aj/Arl),
 * converting columns to uint32 and row lengths to the narrowest unsigned
 * type once.  dtype: 0 = float, 1 = double.  Arl_host may be NULL: row
 * lengths are then derived by scanning for the first zero value
 * (sparse_matrix_ell-impl.hh:186-207). */
int hsolve_ell_create(int dtype, uint64_t rows, uint64_t columns, uint64_t stride,
                      uint64_t num_cols_per_row, uint64_t threads, const uint64_t *Aj_host,
                      const void *Ax_host, const uint64_t *Arl_host, hsolve_ell *out);
int hsolve_ell_destroy(hsolve_ell m);
int hsolve_ell_info(hsolve_ell m, uint64_t *rows, uint64_t *columns, uint64_t *stride,
                    uint64_t *num_cols_per_row, uint64_t *threads, uint64_t *nnz,
                    uint64_t *device_bytes);
/* Kernel variant of the handle (tuning knob, no reference counterpart):
 * 0 = automatic; else bits 1 = one row per lane (default two), 2 = load all
 * num_cols_per_row slots without waiting for row lengths, 4 = XCD-contiguous
 * block order, 16 = stage the x window of each block of 1024 (FP64) or
 * 512 (FP32) rows in LDS (used when every stored column lies within a
 * window of at most 48 KiB around its row; automatic for such matrices
 * unless they fill 90% of their slots), 32 = with 16: read a copy whose rows are sorted by length inside
 * each such block, so slots past a row's length are not fetched (the
 * copy is built at upload when the LDS form is the automatic choice, and
 * the automatic choice then uses it), 64 = with 16: a fixed grid of
 * workgroups walks the LDS blocks and loads the next block's row lengths
 * while it works on the current one.  Results are identical for every
 * variant. */
int hsolve_ell_set_variant(hsolve_ell m, int variant);
/* Grid side M when the handle holds a 9-point structured matrix (rows =
 * columns = M^2, threads == 1, every stored column of row (y, x) one of
 * (y + dy, x + dx), |dy|, |dx| <= 1, ascending), else 0.  The many-RHS
 * kernels then keep a sliding 3 x 3 window of rows (same results); an
 * explicit hsolve_ell_set_variant choice turns that off.  No reference
 * counterpart. */
int hsolve_ell_q1_side(hsolve_ell m, uint64_t *side);
/* y = A x, or y = rhs - A x when rhs != NULL */
int hsolve_ell_spmv(hsolve_ell m, const void *x, void *y, const void *rhs, void *stream);
/* The same on rows [row_start, row_end) only (other entries of y untouched;
 * an empty range is a no-op): HONEI's ELL Product / Defect on a row range
 * (product.hh:148-206, defect.hh:357-384 loop over [row_start, row_end)),
 * as its MultiCore row split (product.hh:2087-2106) and the CUDA launcher
 * (product.cu:356-402) use it; one rank's share of a row-partitioned
 * product.  Same bits as those rows of hsolve_ell_spmv. */
int hsolve_ell_spmv_rows(hsolve_ell m, const void *x, void *y, const void *rhs, uint64_t row_start,
                         uint64_t row_end, void *stream);
/* ConjugateGradients<Tag, NONE>::value(ELL, rhs, x, konv_rad) in one
 * workgroup (honei/math/conjugate_gradients.hh:588-620, kernel :153-200):
 * x = 0, then CG steps while ||x||^2 grows by more than konv_rad; rows <=
 * 2^20.  iters_dev (uint64 device scalar, may be NULL) receives the count;
 * max_iters (0: 100 n + 1000) is a termination guarantee the reference lacks. */
int hsolve_cg_konv_ell(hsolve_ell m, const void *rhs, void *x, double konv_rad, uint64_t max_iters,
                       void *iters_dev, void *stream);
/* x_new = x_old + dinv * (rhs - A x_old) */
int hsolve_ell_jacobi(hsolve_ell m, const void *rhs, const void *dinv, const void *x_old,
                      void *x_new, void *stream);

/* ---------------- solvers ---------------------------------------------------- */
/* CGSolver<Tag, methods::NONE>::value (honei/math/cg.hh:63-133) on a Q1
 * matrix (bands9: 9 device pointers, LL..UU) or an ELL handle; FP64 or
 * FP32 by dtype (0 float, 1 double); b, x device vectors; x is the initial
 * guess and the result.  used_iters and final relative defect on output.
 * Stops when ||r|| < eps*||r0|| or ||r|| < eps or iters == max_iters. */
int hsolve_cg_q1(int dtype, const void *const *bands9, uint64_t size, const void *b, void *x,
                 uint64_t max_iters, double eps, uint64_t *used_iters, double *defect_out,
                 void *stream);
int hsolve_cg_ell(hsolve_ell m, const void *b, void *x, uint64_t max_iters, double eps,
                  uint64_t *used_iters, double *defect_out, void *stream);
/* hsolve_cg_ell that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels, same results, bit for
 * bit; the CG workspace (four vectors, reduction partials, scalars) comes
 * from a fixed 64 MiB device array that each device holds until the process
 * ends, instead of a per-call allocation.  Rows up to 2,092,000 (FP64) /
 * 4,184,000 (FP32); larger matrices return HSOLVE_ERR_SIZE (use
 * hsolve_cg_ell).  Every call on one device shares that array: calls that
 * overlap on one device (other host threads, other streams) race. */
int hsolve_cg_ell_noalloc(hsolve_ell m, const void *b, void *x, uint64_t max_iters, double eps,
                          uint64_t *used_iters, double *defect_out, void *stream);
/* CGSolver<Tag, methods::VAR> with a diagonal preconditioner P (device
 * vector, z = P * r elementwise), honei/math/cg.hh:141-219. */
int hsolve_pcg_ell(hsolve_ell m, const void *P, const void *b, void *x, uint64_t max_iters,
                   double eps, uint64_t *used_iters, double *defect_out, void *stream);
/* CGSolver<Tag, methods::VAR> with a SparseMatrixELL preconditioner P (an
 * ELL handle of the same value type, rows = columns = rows of A; z =
 * Product(P, r)), honei/math/cg.hh:141-219: the preconditioner form of the
 * operator-list multigrid's file-read "spai"/"sainv" matrices. */
int hsolve_pcg_ell_pmat(hsolve_ell m, hsolve_ell P, const void *b, void *x, uint64_t max_iters,
                        double eps, uint64_t *used_iters, double *defect_out, void *stream);
/* ConjugateGradients<Tag, NONE>::value(ELL, rhs, x, unsigned long
 * max_iters) (honei/math/conjugate_gradients.hh:571-586, cg_kernel
 * :203-248): CGSolver<NONE>'s steps from the given x without its epsilon
 * guards (lambda = alpha / (v.p), p = p * (alpha / alpha_old) + r),
 * stopping when ||r|| < 1e-8 ||r0|| or ||r|| < 1e-8 (both formed in
 * double) or after max_iters.  Like the reference, reaching max_iters is
 * not an error; used_iters and the final relative defect on output. */
int hsolve_cg_iters_ell(hsolve_ell m, const void *rhs, void *x, uint64_t max_iters, uint64_t *used_iters,
                        double *defect_out, void *stream);

/* ConjugateGradients<Tag, methods::JAC>::value(ELL A, rhs, x, dd_inverted,
 * max_iters, used_iters, eps) (honei/math/conjugate_gradients.hh:1110-1155,
 * cg_kernel :885-913): PCG with z = dd_inverted * r; stops when ||r|| <
 * eps*||r0|| (no absolute test) or after max_iters. */
int hsolve_pcg_jac_ell(hsolve_ell m, const void *dd_inverted, const void *b, void *x, uint64_t max_iters,
                       double eps, uint64_t *used_iters, double *defect_out, void *stream);

/* ---------------- many right-hand sides on one ELL matrix -------------------
 * A column-blocked multivector of nrhs <= 64 columns stores row i's entries
 * at [i*nrhs, i*nrhs + nrhs) (device memory, the handle's value type).  The
 * handle must have threads == 1.  The reference applies Product / Defect /
 * CGSolver to one DenseVector at a time; these run every column in the same
 * launches with each column's arithmetic unchanged. */
/* Y = A X, or Y = RHS - A X when RHS != NULL: Product<Tag>::value(y, A, x)
 * (honei/la/product.hh:148-206) and Defect<Tag> (honei/math/defect.hh:357-384)
 * per column, bit-identical to hsolve_ell_spmv on that column. */
int hsolve_ell_spmm(hsolve_ell m, uint64_t nrhs, const void *X, void *Y, const void *RHS,
                    void *stream);
/* CGSolver<Tag, methods::VAR> (P != NULL: diagonal preconditioner, z = P * r)
 * or CGSolver<Tag, methods::NONE> (P == NULL), honei/math/cg.hh:63-219, on
 * every column of B at once; X holds the initial guesses and the results.
 * Each column keeps its own scalars, stopping test and iteration count;
 * used_iters[nrhs] and defect_out[nrhs] (final relative defect) are host
 * arrays or NULL.  HSOLVE_ERR_NOT_CONVERGED when a column hits max_iters. */
int hsolve_cg_ell_multi(hsolve_ell m, uint64_t nrhs, const void *P, const void *B, void *X,
                        uint64_t max_iters, double eps, uint64_t *used_iters, double *defect_out,
                        void *stream);
/* Seeded right-hand sides on a side x side grid (FP64): entry i of column j
 * at b[i*row_stride + j*col_stride] is -1 + 2u, u = (SplitMix64 output i of
 * stream (seed + j, stream_id) >> 11) * 2^-53, and 0 on the nodes whose
 * macro_border_mask part (corners 0..3, bottom, right, top, left edges) is
 * 2 (Dirichlet); mask NULL = every border node.  No reference counterpart
 * (synthetic input generator of the benchmarks). */
int hsolve_rhs_uniform_multi(void *b, uint64_t side, uint64_t count, uint64_t row_stride,
                             uint64_t col_stride, const uint64_t *macro_border_mask, uint64_t seed,
                             uint64_t stream_id, void *stream);

/* Mixed-precision MG-IR: Multigrid<Tag, OuterTag, NONE, JAC, CYCLE::V, MIXED>
 * on Q1 bands (honei/math/multigrid.hh:1616-2393).  A hierarchy holds the
 * FP64 finest matrix, FP32 matrices and inverted diagonals on every level
 * and all iterates, device resident. */
typedef struct hsolve_mg_params_s {
    uint64_t min_level, max_level;
    uint64_t n_max_iter;          /* V-cycles per inner solve */
    int32_t initial_zero, convergence_check;
    double tolerance;             /* stored as FP32 like MGInfo<float> */
    uint64_t n_pre_smooth;        /* pre- and (reference quirk) post-smoothing sweeps */
    double adapt_correction_factor;
    uint64_t macro_border_mask[8];
    uint64_t max_outer;           /* 0 = no cap (reference behaviour) */
} hsolve_mg_params;
typedef struct hsolve_mg_s *hsolve_mg;
/* bands64_host: 9 host pointers (finest, FP64).  bands32_host: 9*(max_level+1)
 * host pointers, entry [L*9+b] for L in [min_level, max_level] (others
 * ignored); dinv32_host[L] likewise, or NULL to use the inverted diagonal
 * scaled by 0.7 in double as the reference tests build it. */
int hsolve_mg_create(const hsolve_mg_params *params, const double *const *bands64_host,
                     const float *const *bands32_host, const float *const *dinv32_host,
                     hsolve_mg *out);
/* Variant that assembles every level with FillMatrix<POISSON, DIRICHLET_0>
 * on the device (no host bands). */
int hsolve_mg_create_poisson(const hsolve_mg_params *params, hsolve_mg *out);
int hsolve_mg_destroy(hsolve_mg mg);
int hsolve_mg_device_bytes(hsolve_mg mg, uint64_t *bytes);
/* No counterpart in the reference (it concerns this backend's storage of the
 * band values): a Q1 hierarchy whose levels' bands and inverted diagonals hold
 * at most 16 distinct values each also keeps them as value codes (4-bit table
 * indices, 8 bytes per node) and its V-cycle and outer-defect kernels read
 * those.  enable = 0 makes the handle's solves read the plain bands, 1 (the
 * default) the codes where they exist; both give the same bits. */
int hsolve_mg_set_value_codes(hsolve_mg mg, int enable);
/* b, x: FP64 device vectors of the finest size; x in: initial guess, out: result. */
int hsolve_mg_solve(hsolve_mg mg, const void *b, void *x, uint64_t *outer_iters,
                    uint64_t *inner_iters, double *final_defect, double *initial_defect,
                    void *stream);
/* An ELL matrix in HONEI's host layout with one lane per row (slot k of row
 * r at r + k*stride; uint64 columns and row lengths; values float or double
 * as the context says). */
typedef struct hsolve_ell_host_s {
    uint64_t rows, cols, stride, num_cols_per_row;
    const uint64_t *Aj;
    const void *Ax;
    const uint64_t *Arl;
} hsolve_ell_host;
/* Multigrid<..., NONE | PROLMAT, JAC, CYCLE::V, MIXED> on ELL matrices
 * (multigrid.hh:1936-2543): a32[L] FP32 level matrices for L in [min_level,
 * max_level] (entries below min_level ignored), A64 the FP64 finest matrix,
 * dinv32[L] or NULL (then (float)((double)(1/diag) * 0.7)), prol32[L] /
 * res32[L] (L in (min_level, max_level]) for matrix transfers or both NULL
 * for the structured Q1 transfers (then level L has (2^L+1)^2 rows).
 * `slots` right-hand sides can be solved together (hsolve_mg_solve_batch). */
int hsolve_mg_create_ell(const hsolve_mg_params *params, const hsolve_ell_host *A64,
                         const hsolve_ell_host *a32, const float *const *dinv32,
                         const hsolve_ell_host *prol32, const hsolve_ell_host *res32,
                         uint64_t slots, hsolve_mg *out);
int hsolve_mg_slots(hsolve_mg mg, uint64_t *slots);
/* count <= slots solves in the same launches: b[i] into x[i] (x[i] in:
 * initial guess); per-solve counts and defects into the arrays (may be NULL). */
int hsolve_mg_solve_batch(hsolve_mg mg, uint64_t count, const void *const *b, void *const *x,
                          uint64_t *outer_iters, uint64_t *inner_iters, double *final_defect,
                          double *initial_defect, void *stream);

/* Request server: Poisson Q1 solves (FillMatrix<POISSON, DIRICHLET_0>, side
 * 2^L+1, L in [min_served_level, max_served_level]) by the MG-IR above with
 * x0 = 0.  Requests are admitted into per-level slot tables as slots and the
 * device-memory budget allow (first come, first served on memory); every
 * level's admitted solves advance together, one V-cycle per tick, in the same
 * kernel launches, and levels tick independently on their own streams, so a
 * request that arrives while others are on the GPU joins at the next tick of
 * its level.  A solve gives the same bits alone and in any mix.  No reference
 * counterpart: HONEI solves one system per call (multigrid.hh:2245); this is
 * the serving layer around that call. */
typedef struct hsolve_serve_params_s {
    uint64_t min_level;                 /* coarsest level of every hierarchy */
    uint64_t min_served_level, max_served_level;
    uint64_t n_max_iter;
    int32_t initial_zero, convergence_check;
    double tolerance;
    uint64_t n_pre_smooth;
    double adapt_correction_factor;
    uint64_t macro_border_mask[8];
    uint64_t max_outer;
    uint64_t memory_budget;             /* bytes of request state held at once */
    uint64_t max_slots;                 /* concurrent solves per level */
    uint64_t rows_per_tick;             /* also caps a level's slots at rows_per_tick / (2^L+1)^2
                                           (at least 1), which bounds one tick's work; 0: no cap */
    int32_t use_graphs;                 /* replay each level's tick as a captured graph */
} hsolve_serve_params;
typedef struct hsolve_serve_s *hsolve_serve;
typedef struct hsolve_serve_result_s {
    uint64_t id, tag, level;
    int32_t status;                     /* HSOLVE_OK or HSOLVE_ERR_NOT_CONVERGED */
    uint64_t outer_iters, inner_iters, ticks;
    double initial_defect, final_defect;
    double t_submit, t_admit, t_done;   /* seconds on the server clock */
} hsolve_serve_result;
typedef struct hsolve_serve_stats_s {
    uint64_t bytes_in_use, bytes_held, bytes_shared, budget;
    uint64_t submitted, admitted, completed, queued, running, peak_running;
    uint64_t ticks, graph_captures;
} hsolve_serve_stats;
int hsolve_serve_create(const hsolve_serve_params *params, hsolve_serve *out);
int hsolve_serve_destroy(hsolve_serve srv);
/* Bytes one request of level L holds while admitted. */
int hsolve_serve_request_bytes(hsolve_serve srv, uint64_t level, uint64_t *bytes);
/* Queues a request.  b: FP64 device vector of (2^L+1)^2 entries, copied at
 * admission (keep it valid until the request completes), or NULL to generate
 * b = U[-1,1] from SplitMix64(rhs_seed, stream 1) with zero Dirichlet nodes.
 * x: FP64 device output, or NULL to discard the solution.  t_submit: the
 * arrival time on the server clock (< 0: now).  HSOLVE_ERR_BUDGET when one
 * request alone exceeds the budget. */
int hsolve_serve_submit(hsolve_serve srv, uint64_t level, const double *b, uint64_t rhs_seed,
                        double *x, uint64_t tag, double t_submit, uint64_t *id);
/* One scheduling pass: collects every level whose tick finished, admits,
 * and launches the next tick of every level with work.  Never blocks on the
 * GPU.  *busy = levels with a tick in flight. */
int hsolve_serve_step(hsolve_serve srv, uint64_t *busy);
/* Moves up to max finished requests into out. */
int hsolve_serve_poll(hsolve_serve srv, hsolve_serve_result *out, uint64_t max, uint64_t *count);
int hsolve_serve_now(hsolve_serve srv, double *seconds);
int hsolve_serve_get_stats(hsolve_serve srv, hsolve_serve_stats *out);
/* Replays an arrival trace in real time: request i (level levels[i], seed
 * seeds[i], output x[i] or NULL) is submitted at arrival_s[i] seconds after
 * the call starts; returns when every request is done, with results[i] for
 * request i. */
int hsolve_serve_run(hsolve_serve srv, uint64_t count, const double *arrival_s,
                     const uint64_t *levels, const uint64_t *seeds, double *const *x,
                     hsolve_serve_result *results);

#ifdef __cplusplus
}
#endif
#endif
