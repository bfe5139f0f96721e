/*
 * oracle_impl.h -- type-generic body of the CPU oracle.  Included twice by
 * oracle.c with OR_T = double / float and OR_S = f64 / f32.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#define OR_CAT2(a, b) a##_##b
#define OR_CAT(a, b) OR_CAT2(a, b)
#define FN(name) OR_CAT(name, OR_S)
#define T OR_T

/* honei/la/sparse_matrix_ell-impl.hh:78-117: rows of a SparseMatrix into the
 * column-major "threads" layout; explicit zeros are skipped, Arl counts the
 * stored entries. */
void FN(or_ell_from_csr)(uint64_t rows, const uint64_t *rowptr, const uint64_t *col,
                         const T *val, uint64_t threads, uint64_t stride,
                         uint64_t num_cols_per_row, uint64_t *Aj, T *Ax, uint64_t *Arl)
{
    uint64_t size = num_cols_per_row * stride;
    for (uint64_t i = 0; i < size; ++i) {
        Aj[i] = 0;
        Ax[i] = (T)0;
    }
    for (uint64_t r = 0; r < rows; ++r) {
        uint64_t used = rowptr[r + 1] - rowptr[r];
        Arl[r] = (uint64_t)ceil((double)used / (double)threads);
        uint64_t target = 0;
        for (uint64_t i = rowptr[r]; i < rowptr[r + 1]; ++i) {
            if (val[i] != (T)0) {
                uint64_t pos = (target % threads) + r * threads + (target / threads) * stride;
                Aj[pos] = col[i];
                Ax[pos] = val[i];
                ++target;
            }
        }
    }
}

/* honei/la/sparse_matrix_ell-impl.hh:186-207 (row lengths of a raw ELL
 * matrix as read from file: count slots until the first zero value). */
void FN(or_ell_row_length)(uint64_t rows, uint64_t stride, uint64_t threads,
                           uint64_t size, const T *Ax, uint64_t *Arl)
{
    for (uint64_t row = 0; row < rows; ++row) {
        uint64_t count = 0;
        for (uint64_t i = row * threads; i < size; i += stride) {
            int stop = 0;
            for (uint64_t t = 0; t < threads; ++t) {
                if (Ax[i + t] == (T)0) {
                    stop = 1;
                    break;
                }
                if (t == 0)
                    ++count;
            }
            if (stop)
                break;
        }
        Arl[row] = count;
    }
}

/* Product<tags::CPU>::value(y, SparseMatrixELL, x, row_start, row_end):
 * honei/la/product.hh:148-206. */
void FN(or_ell_product)(T *y, const uint64_t *Aj, const T *Ax, const uint64_t *Arl,
                        const T *x, uint64_t row_start, uint64_t row_end,
                        uint64_t stride, uint64_t threads)
{
    for (uint64_t row = row_start; row < row_end; ++row) {
        T sum = 0;
        uint64_t j = row * threads;
        for (uint64_t c = 0; c < Arl[row]; ++c, j += stride)
            for (uint64_t t = 0; t < threads; ++t)
                sum += Ax[j + t] * x[Aj[j + t]];
        y[row] = sum;
    }
}

/* Defect<tags::CPU>::value(r, rhs, SparseMatrixELL, x):
 * honei/math/defect.hh:375-397 (product into a temporary, then rhs - temp). */
void FN(or_ell_defect)(T *r, const T *rhs, const uint64_t *Aj, const T *Ax,
                       const uint64_t *Arl, const T *x, uint64_t rows,
                       uint64_t stride, uint64_t threads)
{
    T *temp = (T *)malloc(sizeof(T) * (rows ? rows : 1));
    FN(or_ell_product)(temp, Aj, Ax, Arl, x, 0, rows, stride, threads);
    for (uint64_t i = 0; i < rows; ++i)
        r[i] = rhs[i] - temp[i];
    free(temp);
}

/* SparseMatrixCSR<T>(SparseMatrix) with csr::blocksize = bs:
 * honei/la/sparse_matrix_csr-impl.hh:98-136.  The input is a row-wise
 * sparse matrix (rowptr/col/val, columns ascending per row; stored zeros
 * stay stored).  Each stored block starts at the column rounded down to a
 * multiple of bs and takes the row's next entries whose columns fall on its
 * bs consecutive positions; the other positions hold 0.  Outputs: Ar
 * (rows + 1 block offsets), Aj (block start columns), Ax (bs values per
 * block); *nblocks = Ar[rows], *used = stored entries.  Aj and Ax must hold
 * rowptr[rows] and rowptr[rows] * bs elements. */
void FN(or_csr_from_rows)(uint64_t rows, const uint64_t *rowptr, const uint64_t *col,
                          const T *val, uint64_t bs, uint64_t *Ar, uint64_t *Aj, T *Ax,
                          uint64_t *nblocks, uint64_t *used)
{
    uint64_t gi = 0, ue = 0, ax = 0;
    for (uint64_t r = 0; r < rows; ++r) {
        const uint64_t n = rowptr[r + 1] - rowptr[r];
        const uint64_t *c = col + rowptr[r];
        const T *v = val + rowptr[r];
        Ar[r] = gi;
        uint64_t i = 0;
        while (i < n) {
            const uint64_t start = c[i] - c[i] % bs;
            uint64_t taken = 0;
            Aj[gi] = start;
            for (uint64_t k = 0; k < bs; ++k) {
                if (i + taken < n && start + k == c[i + taken]) {
                    Ax[ax++] = v[i + taken];
                    ++ue;
                    ++taken;
                } else {
                    Ax[ax++] = (T)0;
                }
            }
            i += taken;
            ++gi;
        }
    }
    Ar[rows] = gi;
    *nblocks = gi;
    *used = ue;
}

/* Product<tags::CPU>::value(y, SparseMatrixCSR, x, row_start, row_end):
 * honei/la/product.hh:208-245.  Positions at or past the last column
 * (x has `columns` entries) are skipped; rows outside the range keep y. */
void FN(or_csr_product)(T *y, const uint64_t *Ar, const uint64_t *Aj, const T *Ax,
                        const T *x, uint64_t columns, uint64_t bs,
                        uint64_t row_start, uint64_t row_end)
{
    for (uint64_t row = row_start; row < row_end; ++row) {
        T sum = 0;
        for (uint64_t i = Ar[row]; i < Ar[row + 1]; ++i)
            for (uint64_t k = 0; k < bs; ++k)
                if (Aj[i] + k < columns)
                    sum += Ax[i * bs + k] * x[Aj[i] + k];
        y[row] = sum;
    }
}

/* Defect<tags::CPU>::value(r, rhs, SparseMatrixCSR, x, row_start, row_end):
 * honei/math/defect.hh:386-418 (the row sum of the product, then
 * r = rhs - sum). */
void FN(or_csr_defect)(T *r, const T *rhs, const uint64_t *Ar, const uint64_t *Aj,
                       const T *Ax, const T *x, uint64_t columns, uint64_t bs,
                       uint64_t row_start, uint64_t row_end)
{
    for (uint64_t row = row_start; row < row_end; ++row) {
        T sum = 0;
        for (uint64_t i = Ar[row]; i < Ar[row + 1]; ++i)
            for (uint64_t k = 0; k < bs; ++k)
                if (Aj[i] + k < columns)
                    sum += Ax[i * bs + k] * x[Aj[i] + k];
        r[row] = rhs[row] - sum;
    }
}

/* Product<tags::CPU>::value(y, BandedMatrixQx<Q1Type>, x):
 * honei/la/product.hh:370-403, summation order DD,LL,LD,LU,DL,DU,UL,UD,UU. */
void FN(or_q1_product)(T *y, const T *const *bands, const T *x, uint64_t n)
{
    const long size = (long)n;
    const long root = (long)sqrt((double)n);
    const T *ll = bands[0], *ld = bands[1], *lu = bands[2];
    const T *dl = bands[3], *dd = bands[4], *du = bands[5];
    const T *ul = bands[6], *ud = bands[7], *uu = bands[8];
    for (long i = 0; i < size; ++i) {
        T s = dd[i] * x[i];
        if (i - root - 1 >= 0) s += ll[i] * x[i - root - 1];
        if (i - root >= 0)     s += ld[i] * x[i - root];
        if (i - root + 1 >= 0) s += lu[i] * x[i - root + 1];
        if (i - 1 >= 0)        s += dl[i] * x[i - 1];
        if (i + 1 < size)      s += du[i] * x[i + 1];
        if (i + root - 1 < size) s += ul[i] * x[i + root - 1];
        if (i + root < size)   s += ud[i] * x[i + root];
        if (i + root + 1 < size) s += uu[i] * x[i + root + 1];
        y[i] = s;
    }
}

/* Defect<tags::CPU>::value(r, rhs, BandedMatrixQx<Q1Type>, x):
 * honei/math/defect.hh:193-197 = Difference(r, rhs, Product(A, x)). */
void FN(or_q1_defect)(T *r, const T *rhs, const T *const *bands, const T *x, uint64_t n)
{
    T *temp = (T *)malloc(sizeof(T) * (n ? n : 1));
    FN(or_q1_product)(temp, bands, x, n);
    for (uint64_t i = 0; i < n; ++i)
        r[i] = rhs[i] - temp[i];
    free(temp);
}

/* DotProduct<tags::CPU>: honei/la/dot_product.hh:72-87 (sequential sum). */
T FN(or_dot)(const T *x, const T *y, uint64_t n)
{
    T result = 0;
    for (uint64_t i = 0; i < n; ++i)
        result += x[i] * y[i];
    return result;
}

/* Norm<vnt_l_two, false/true, tags::CPU>: honei/la/norm.hh:290-308, 363-381
 * (squared norm = DotProduct(x, x); root -> sqrt in T). */
T FN(or_norm_l2)(const T *x, uint64_t n, int32_t root)
{
    T s = FN(or_dot)(x, x, n);
    return root ? (T)sqrt(s) : s;
}

/* ScaledSum<tags::CPU>, honei/la/scaled_sum.hh:73-89: x += b * y. */
void FN(or_scaled_sum)(T *x, const T *y, T b, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        x[i] += b * y[i];
}

/* scaled_sum.hh:91-111: r = b * y + x. */
void FN(or_scaled_sum3)(T *r, const T *x, const T *y, T b, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        r[i] = b * y[i] + x[i];
}

/* scaled_sum.hh:177-197: a += b * c (elementwise). */
void FN(or_scaled_sum_vv)(T *a, const T *b, const T *c, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        a[i] += b[i] * c[i];
}

/* Sum<tags::CPU>: x += y (honei/la/sum.hh). */
void FN(or_sum)(T *x, const T *y, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        x[i] += y[i];
}

/* Scale<tags::CPU>, honei/la/scale.hh:229-240: x *= a with a of type T. */
void FN(or_scale)(T *x, T a, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        x[i] *= a;
}

/* honei/la/scale.hh:111-121 with a double factor on a T vector (the overload
 * selected by Scale<Tag>::value(DenseVector<float>&, 0.7)): the product is
 * formed in double and rounded to T. */
void FN(or_scale_by_double)(T *x, double a, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        x[i] = (T)((double)x[i] * a);
}

/* Difference<tags::CPU>: x -= y (honei/la/difference.hh). */
void FN(or_difference)(T *x, const T *y, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        x[i] -= y[i];
}

/* Difference<tags::CPU>::value(r, x, y): r = x - y. */
void FN(or_difference3)(T *r, const T *x, const T *y, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        r[i] = x[i] - y[i];
}

/* ElementProduct<tags::CPU>: a *= b (honei/la/element_product.hh). */
void FN(or_element_product)(T *a, const T *b, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        a[i] *= b[i];
}

/* ElementProduct<tags::CPU>::value(r, a, b): r = a * b. */
void FN(or_element_product3)(T *r, const T *a, const T *b, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        r[i] = a[i] * b[i];
}

/* ElementInverse<tags::CPU>, honei/la/element_inverse.hh:66-80: x = 1/x
 * for x != 0, zeros are kept. */
void FN(or_element_inverse)(T *x, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) {
        if (x[i] == (T)0)
            continue;
        x[i] = (T)1 / x[i];
    }
}

/* ApplyDirichletBoundaries: honei/math/apply_dirichlet_boundaries.hh:31-80. */
void FN(or_apply_dirichlet)(T *grid, uint64_t N, const uint64_t *mask)
{
    uint64_t M = (uint64_t)(int)sqrt((double)N);
    if (mask[0] == 2) grid[0] = 0;
    if (mask[1] == 2) grid[M - 1] = 0;
    if (mask[2] == 2) grid[N - 1] = 0;
    if (mask[3] == 2) grid[N - M] = 0;
    if (mask[4] == 2)
        for (uint64_t i = 1; i < M - 1; i++) grid[i] = 0;
    if (mask[5] == 2)
        for (uint64_t i = 2 * M - 1; i < N - M; i += M) grid[i] = 0;
    if (mask[6] == 2)
        for (uint64_t i = N - M + 1; i < N - 1; i++) grid[i] = 0;
    if (mask[7] == 2)
        for (uint64_t i = M; i < N - M; i += M) grid[i] = 0;
}

/* Grid side of a (2^L+1)^2 vector, computed the way the transfer operators
 * do it: (unsigned long)sqrt((Prec_)n) (restriction.hh:48, prolongation.hh:53).
 * For the float instantiation the argument is rounded to float first. */
static uint64_t FN(or_side)(uint64_t n)
{
#if OR_IS_FLOAT
    return (uint64_t)sqrtf((float)n);
#else
    return (uint64_t)sqrt((double)n);
#endif
}

/* Restriction<Tag, NONE>: honei/math/restriction.hh:43-172 (full weighting,
 * scattered in the reference's loop order). */
void FN(or_restriction)(T *coarse, const T *fine, uint64_t n_fine, const uint64_t *mask)
{
    uint64_t n_x_fine = FN(or_side)(n_fine) - 1;
    uint64_t n_x_coarse = n_x_fine / 2;
    uint64_t n_y_fine = n_x_fine + 1;
    uint64_t n_y_coarse = n_x_coarse + 1;
    uint64_t i_step = 2 * n_y_fine;
    long i_aux, i_aux_1;

    for (uint64_t i = 1; i <= n_y_coarse; ++i)
        for (uint64_t j = 0, k = 0; j < n_y_coarse; ++j, k += 2)
            coarse[1 + n_y_coarse * (i - 1) - 1 + j] = fine[2 * n_y_fine * (i - 1) + k];

    for (uint64_t i = 1; i <= n_x_coarse; ++i) {
        i_aux = (long)(1 + n_y_fine + 2 * n_y_fine * (i - 1));
        i_aux_1 = (long)(1 + n_y_coarse * i);
        for (uint64_t j = 0; j < n_y_coarse; ++j)
            coarse[i_aux_1 - 1 + j] += 0.5 * fine[i_aux - 1 + j * 2];
        i_aux_1 = i_aux_1 - (long)n_y_coarse;
        for (uint64_t j = 0; j < n_y_coarse; ++j)
            coarse[i_aux_1 - 1 + j] += 0.5 * fine[i_aux - 1 + j * 2];
        for (uint64_t j = 0; j < n_y_coarse; ++j)
            coarse[i - 1 + j * n_y_coarse] += 0.5 * fine[2 * i - 1 + j * i_step];
        for (uint64_t j = 0; j < n_y_coarse; ++j)
            coarse[i + j * n_y_coarse] += 0.5 * fine[2 * i - 1 + j * i_step];
        ++i_aux;
        for (uint64_t j = 0; j < n_x_coarse; ++j)
            coarse[i_aux_1 - 1 + j] += 0.25 * fine[i_aux - 1 + j * 2];
        for (uint64_t j = 0; j < n_x_coarse; ++j)
            coarse[i_aux_1 + j] += 0.25 * fine[i_aux - 1 + j * 2];
        for (uint64_t j = 0; j < n_x_coarse; ++j)
            coarse[n_y_coarse * i + j] += 0.25 * fine[i_aux - 1 + j * 2];
        for (uint64_t j = 0; j < n_x_coarse; ++j)
            coarse[1 + n_y_coarse * i + j] += 0.25 * fine[i_aux - 1 + j * 2];
    }
    FN(or_apply_dirichlet)(coarse, n_y_coarse * n_y_coarse, mask);
}

/* Prolongation<Tag, NONE>: honei/math/prolongation.hh:43-139 (bilinear). */
void FN(or_prolongation)(T *fine, const T *coarse, uint64_t n_fine, const uint64_t *mask)
{
    uint64_t n_x_fine = FN(or_side)(n_fine) - 1;
    uint64_t n_x_coarse = n_x_fine / 2;
    uint64_t n_y_fine = n_x_fine + 1;
    uint64_t n_y_coarse = n_x_coarse + 1;
    uint64_t i_step = 2 * n_y_fine;

    for (uint64_t i = 0; i < n_fine; ++i)
        fine[i] = 0.0;
    for (uint64_t i = 1; i <= n_y_coarse; ++i)
        for (uint64_t j = 0, k = 0; j < n_y_coarse; ++j, k += 2)
            fine[i_step * (i - 1) + k] = coarse[i + n_x_coarse * (i - 1) - 1 + j];
    for (uint64_t i = 1; i <= n_x_coarse; ++i) {
        uint64_t iaux = 2 * i;
        for (uint64_t j = 0; j < n_y_coarse; ++j) {
            fine[iaux - 1 + j * i_step] += coarse[i - 1 + j * n_y_coarse] * 0.5;
            fine[iaux - 1 + j * i_step] += coarse[i + j * n_y_coarse] * 0.5;
        }
        iaux = n_y_fine + 1 + (i - 1) * i_step;
        for (uint64_t j = 0; j < n_y_coarse; ++j) {
            fine[iaux - 1 + j * 2] += fine[iaux - n_y_fine - 1 + j * 2] * 0.5;
            fine[iaux - 1 + j * 2] += fine[iaux + n_y_fine - 1 + j * 2] * 0.5;
        }
    }
    for (uint64_t i = 1; i <= n_x_coarse; ++i) {
        uint64_t iaux = n_x_fine + 3 + (i - 1) * i_step;
        for (uint64_t j = 0; j < n_x_coarse; ++j) {
            fine[iaux - 1 + j * 2] += fine[iaux - n_x_fine - 3 + j * 2] * 0.25;
            fine[iaux - 1 + j * 2] += fine[iaux - n_x_fine - 1 + j * 2] * 0.25;
            fine[iaux - 1 + j * 2] += fine[iaux + n_x_fine - 1 + j * 2] * 0.25;
            fine[iaux - 1 + j * 2] += fine[iaux + n_x_fine + 1 + j * 2] * 0.25;
        }
    }
    FN(or_apply_dirichlet)(fine, n_fine, mask);
}

/* Jacobi<Tag>::value(to_smooth, Q1, rhs, x, iters, omega, dinv):
 * honei/math/jacobi.hh:438-480 with jacobi_kernel jacobi.hh:117-139 when
 * to_smooth and x are the same vector: x += dinv * (rhs - A x), iters times. */
void FN(or_jacobi_q1)(T *x, const T *const *bands, const T *rhs, const T *dinv,
                      uint64_t n, uint64_t iters, T *temp)
{
    for (uint64_t it = 0; it < iters; ++it) {
        FN(or_q1_defect)(temp, rhs, bands, x, n);
        FN(or_scaled_sum_vv)(x, temp, dinv, n);
    }
}

/* Same for ELL: jacobi.hh:268-279, 505-523. */
void FN(or_jacobi_ell)(T *x, const uint64_t *Aj, const T *Ax, const uint64_t *Arl,
                       uint64_t stride, uint64_t threads, const T *rhs, const T *dinv,
                       uint64_t n, uint64_t iters, T *temp)
{
    for (uint64_t it = 0; it < iters; ++it) {
        FN(or_ell_defect)(temp, rhs, Aj, Ax, Arl, x, n, stride, threads);
        FN(or_scaled_sum_vv)(x, temp, dinv, n);
    }
}

/* FillMatrix<Tag, POISSON, DIRICHLET_0>: honei/math/fill_matrix.hh:37-76. */
void FN(or_fill_poisson_dirichlet0)(T *const *bands, uint64_t N)
{
    uint64_t M = (uint64_t)sqrt((double)N);
    for (uint64_t i = 1; i <= N; ++i) {
        int bnd = (i <= M || i > N - M || (i % M) == 1 || (i % M) == 0);
        for (int b = 0; b < 9; ++b)
            bands[b][i - 1] = bnd ? (T)0.0 : (T)(-1.0 / 3.0);
        bands[4][i - 1] = bnd ? (T)1.0 : (T)(8.0 / 3.0);
    }
}

/* FillMatrix<Tag, POISSON, DIRICHLET_NEUMANN>: fill_matrix.hh:79-131. */
void FN(or_fill_poisson_dirichlet_neumann)(T *const *bands, uint64_t N)
{
    uint64_t M = (uint64_t)sqrt((double)N);
    T *ll = bands[0], *ld = bands[1], *lu = bands[2], *dl = bands[3], *dd = bands[4];
    T *du = bands[5], *ul = bands[6], *ud = bands[7], *uu = bands[8];
    for (uint64_t i = 1; i <= N; ++i) {
        uint64_t k = i - 1;
        if (i <= M || i > N - M || (i % M) == 1) {
            dd[k] = (T)1.0;
            ll[k] = ld[k] = lu[k] = dl[k] = du[k] = ul[k] = ud[k] = uu[k] = (T)0.0;
        } else if ((i % M) == 0 && i > M && i < N) {
            dd[k] = (T)(4.0 / 3.0);
            ll[k] = (T)(-1.0 / 3.0);
            ld[k] = (T)(-1.0 / 6.0);
            dl[k] = (T)(-1.0 / 3.0);
            ul[k] = (T)(-1.0 / 3.0);
            ud[k] = (T)(-1.0 / 6.0);
            lu[k] = du[k] = uu[k] = (T)0.0;
        } else {
            dd[k] = (T)(8.0 / 3.0);
            ll[k] = ld[k] = lu[k] = dl[k] = du[k] = ul[k] = ud[k] = uu[k] = (T)(-1.0 / 3.0);
        }
    }
}

/* FillVector<Tag, POISSON, DIRICHLET_NEUMANN>: fill_vector.hh:72-158. */
void FN(or_fill_vector_dirichlet_neumann)(T *v, uint64_t N)
{
    uint64_t M = (uint64_t)sqrt((double)N);
    int level = 0;
    while (level <= 30 && (uint64_t)(((1ull << level) + 1) * ((1ull << level) + 1)) != N)
        ++level;
    for (uint64_t i = 1; i <= N; i++) {
        if (i <= M || i > N - M || (i % M) == 1)
            v[i - 1] = (T)2.0;
        else if ((i % M) == 0 && i > M && i < N)
            v[i - 1] = (T)(0.5 * (1.0 / pow(4.0, level)));
        else
            v[i - 1] = (T)(1.0 / pow(4.0, level));
    }
}

/* Generic matrix application used by the CG restatements. */
typedef struct FN(or_matrix) {
    int kind; /* 0 = Q1 bands, 1 = ELL */
    const T *const *bands;
    const uint64_t *Aj;
    const T *Ax;
    const uint64_t *Arl;
    uint64_t stride, threads;
} FN(or_matrix);

static void FN(or_apply)(const FN(or_matrix) *A, T *y, const T *x, uint64_t n)
{
    if (A->kind == 0)
        FN(or_q1_product)(y, A->bands, x, n);
    else
        FN(or_ell_product)(y, A->Aj, A->Ax, A->Arl, x, 0, n, A->stride, A->threads);
}

static void FN(or_defect_any)(const FN(or_matrix) *A, T *r, const T *b, const T *x, uint64_t n)
{
    if (A->kind == 0)
        FN(or_q1_defect)(r, b, A->bands, x, n);
    else
        FN(or_ell_defect)(r, b, A->Aj, A->Ax, A->Arl, x, n, A->stride, A->threads);
}

#if OR_IS_FLOAT
#define OR_EPS FLT_EPSILON
#else
#define OR_EPS DBL_EPSILON
#endif

/* CGSolver<Tag, methods::NONE>::value: honei/math/cg.hh:63-133. */
static int FN(or_cg_none)(const FN(or_matrix) *A, const T *b, T *x, uint64_t n,
                          uint64_t max_iters, T eps_relative, uint64_t *used_iters)
{
    T *p = (T *)malloc(sizeof(T) * n), *r = (T *)malloc(sizeof(T) * n), *v = (T *)malloc(sizeof(T) * n);
    T alpha, alpha_old, lambda, initial_defect, current_defect = 0, temp;
    uint64_t iterations = 0;
    int converged = 0;

    FN(or_defect_any)(A, r, b, x, n);
    memcpy(p, r, sizeof(T) * n);
    alpha = FN(or_norm_l2)(r, n, 0);
    initial_defect = (T)sqrt(alpha);
    *used_iters = 0;
    while (iterations < max_iters) {
        FN(or_apply)(A, v, p, n);
        temp = FN(or_dot)(v, p, n);
        lambda = alpha / (fabs(temp) > OR_EPS ? temp : (T)OR_EPS);
        FN(or_scaled_sum)(x, p, lambda, n);
        T mlambda = -lambda;
        FN(or_scaled_sum)(r, v, mlambda, n);
        alpha_old = alpha;
        alpha = FN(or_norm_l2)(r, n, 0);
        T talpha = alpha / (fabs(alpha_old) > OR_EPS ? alpha_old : (T)OR_EPS);
        FN(or_scale)(p, talpha, n);
        FN(or_sum)(p, r, n);
        ++iterations;
        current_defect = (T)sqrt(alpha);
        if (current_defect < eps_relative * initial_defect) {
            *used_iters = iterations;
            converged = 1;
            break;
        }
        if (current_defect < eps_relative) {
            *used_iters = iterations;
            converged = 1;
            break;
        }
        if (iterations == max_iters)
            *used_iters = iterations;
    }
    free(p); free(r); free(v);
    return converged ? 0 : 1;
}

int FN(or_cg_none_q1)(const T *const *bands, const T *b, T *x, uint64_t n,
                      uint64_t max_iters, T eps, uint64_t *used_iters)
{
    FN(or_matrix) A = {0, bands, 0, 0, 0, 0, 0};
    return FN(or_cg_none)(&A, b, x, n, max_iters, eps, used_iters);
}

int FN(or_cg_none_ell)(const uint64_t *Aj, const T *Ax, const uint64_t *Arl,
                       uint64_t stride, uint64_t threads, const T *b, T *x, uint64_t n,
                       uint64_t max_iters, T eps, uint64_t *used_iters)
{
    FN(or_matrix) A = {1, 0, Aj, Ax, Arl, stride, threads};
    return FN(or_cg_none)(&A, b, x, n, max_iters, eps, used_iters);
}

static T FN(or_sign)(T v) { return v / (T)fabs(v); }

/* The preconditioner of CGSolver<Tag, VAR>: z = Product(P, r), an
 * elementwise product for a DenseVector P, an ELL product for a
 * SparseMatrixELL P. */
static void FN(or_precon)(const T *Pv, const OR_CAT(or_ell, OR_S) *Pm, T *z, const T *r, uint64_t n)
{
    if (Pm)
        FN(or_ell_product)(z, Pm->Aj, Pm->Ax, Pm->Arl, r, 0, Pm->rows, Pm->stride, Pm->threads);
    else
        FN(or_element_product3)(z, Pv, r, n);
}

/* CGSolver<Tag, methods::VAR>::value: honei/math/cg.hh:141-219, with the
 * preconditioner applied by or_precon. */
static int FN(or_cg_var_any)(const FN(or_matrix) *A, const T *Pv, const OR_CAT(or_ell, OR_S) *Pm,
                             const T *b, T *x, uint64_t n, uint64_t max_iters, T eps_relative,
                             uint64_t *used_iters)
{
    T *p = (T *)malloc(sizeof(T) * n), *r = (T *)malloc(sizeof(T) * n);
    T *v = (T *)malloc(sizeof(T) * n), *z = (T *)malloc(sizeof(T) * n);
    T alpha, alpha_new, lambda, initial_defect, temp;
    uint64_t iterations = 0;
    int converged = 0;

    FN(or_defect_any)(A, r, b, x, n);
    FN(or_precon)(Pv, Pm, p, r, n);
    initial_defect = FN(or_norm_l2)(r, n, 1);
    alpha_new = FN(or_dot)(r, p, n);
    *used_iters = 0;
    while (iterations < max_iters) {
        FN(or_apply)(A, v, p, n);
        temp = FN(or_dot)(v, p, n);
        lambda = alpha_new / (fabs(temp) > OR_EPS ? temp : (T)OR_EPS * FN(or_sign)(temp));
        ++iterations;
        FN(or_scaled_sum)(x, p, lambda, n);
        T mlambda = -lambda;
        FN(or_scaled_sum)(r, v, mlambda, n);
        T current_defect = FN(or_norm_l2)(r, n, 1);
        if (current_defect < eps_relative * initial_defect) {
            *used_iters = iterations;
            converged = 1;
            break;
        }
        if (current_defect < eps_relative) {
            *used_iters = iterations;
            converged = 1;
            break;
        }
        if (iterations == max_iters) {
            *used_iters = iterations;
            break;
        }
        FN(or_precon)(Pv, Pm, z, r, n);
        alpha = alpha_new;
        alpha_new = FN(or_dot)(r, z, n);
        T talpha_new = alpha_new / (fabs(alpha) > OR_EPS ? alpha : (T)OR_EPS * FN(or_sign)(alpha));
        FN(or_scale)(p, talpha_new, n);
        FN(or_sum)(p, z, n);
    }
    free(p); free(r); free(v); free(z);
    return converged ? 0 : 1;
}

/* CGSolver<Tag, methods::VAR> with a DenseVector preconditioner P. */
int FN(or_cg_var_ell)(const uint64_t *Aj, const T *Ax, const uint64_t *Arl,
                      uint64_t stride, uint64_t threads, const T *P, const T *b, T *x,
                      uint64_t n, uint64_t max_iters, T eps_relative, uint64_t *used_iters)
{
    FN(or_matrix) A = {1, 0, Aj, Ax, Arl, stride, threads};
    return FN(or_cg_var_any)(&A, P, NULL, b, x, n, max_iters, eps_relative, used_iters);
}

/* CGSolver<Tag, methods::VAR> with a SparseMatrixELL preconditioner P
 * (rows(P) = columns(P) = rows(A)). */
int FN(or_cg_var_ell_pm)(const OR_CAT(or_ell, OR_S) *M, const OR_CAT(or_ell, OR_S) *P, const T *b, T *x,
                         uint64_t max_iters, T eps_relative, uint64_t *used_iters)
{
    FN(or_matrix) A = {1, 0, M->Aj, M->Ax, M->Arl, M->stride, M->threads};
    return FN(or_cg_var_any)(&A, NULL, P, b, x, M->rows, max_iters, eps_relative, used_iters);
}

/* ConjugateGradients<Tag, JAC>::value(ELL A, rhs, x, dd_inverted,
 * max_iters, used_iters, eps): honei/math/conjugate_gradients.hh:1110-1155
 * with the ELL cg_kernel :885-913.  Per step: beta = r.z; t = A d;
 * alpha = beta / (d.t) (eps-guarded); x += alpha d; r -= alpha t;
 * z = dd_inv * r; beta = (z.r) / beta (eps-guarded); d = z + beta d.
 * Stops when ||r|| < eps ||r0|| (no absolute test) or after max_iters;
 * used_iters is set by either.  Returns 0 when converged. */
int FN(or_pcg_jac_ell)(const uint64_t *Aj, const T *Ax, const uint64_t *Arl, uint64_t stride,
                       uint64_t threads, const T *dd_inv, const T *rhs, T *x, uint64_t n,
                       uint64_t max_iters, T eps_relative, uint64_t *used_iters)
{
    FN(or_matrix) A = {1, 0, Aj, Ax, Arl, stride, threads};
    T *r = (T *)malloc(sizeof(T) * n), *z = (T *)malloc(sizeof(T) * n);
    T *d = (T *)malloc(sizeof(T) * n), *t0 = (T *)malloc(sizeof(T) * n);
    int converged = 0;
    FN(or_defect_any)(&A, r, rhs, x, n);
    const T initial = FN(or_norm_l2)(r, n, 1);
    FN(or_element_product3)(z, dd_inv, r, n);
    memcpy(d, z, sizeof(T) * n);
    *used_iters = 0;
    for (uint64_t i = 0; i < max_iters; ++i) {
        T beta = FN(or_dot)(r, z, n);
        FN(or_apply)(&A, t0, d, n);
        T alpha = FN(or_dot)(d, t0, n);
        alpha = fabs(alpha) >= OR_EPS ? beta / alpha : beta / (T)OR_EPS;
        FN(or_scaled_sum)(x, d, alpha, n);
        FN(or_scaled_sum)(r, t0, -alpha, n);
        FN(or_element_product3)(z, dd_inv, r, n);
        const T zr = FN(or_dot)(z, r, n);
        beta = fabs(beta) >= OR_EPS ? zr / beta : zr / (T)OR_EPS;
        memcpy(t0, d, sizeof(T) * n);                    /* copy<Tag_>(d, temp_0) */
        FN(or_scaled_sum3)(d, z, t0, beta, n);           /* d = z + beta temp_0 */
        const T current = FN(or_norm_l2)(r, n, 1);
        if (current < initial * eps_relative) {
            *used_iters = i + 1;
            converged = 1;
            break;
        }
        if (i == max_iters - 1)
            *used_iters = i + 1;
    }
    free(r); free(z); free(d); free(t0);
    return converged ? 0 : 1;
}

/* One step of the "version 1" cg_kernel of ConjugateGradients<Tag, NONE>
 * (Q1: honei/math/conjugate_gradients.hh:108-151; ELL: :153-200, the same
 * steps on ELL products) on gradient g, iterate x and direction u;
 * energy and ng are work vectors. */
static void FN(or_cg_kernel_v1)(const FN(or_matrix) *A, const T *rhs, T *g, T *x, T *u, T *energy, T *ng,
                                uint64_t n)
{
    T upper = FN(or_norm_l2)(g, n, 0);
    FN(or_apply)(A, energy, u, n);
    T lower = FN(or_dot)(energy, u, n);
    memcpy(energy, u, sizeof(T) * n); /* u_c = copy(u); energy = u_c */
    if (fabs(lower) >= OR_EPS)
        FN(or_scale)(energy, (T)(upper / lower), n);
    else
        FN(or_scale)(energy, (T)(upper / (T)OR_EPS), n);
    FN(or_sum)(energy, x, n);
    lower = upper;
    FN(or_apply)(A, ng, energy, n);
    FN(or_difference)(ng, rhs, n);
    upper = FN(or_norm_l2)(ng, n, 0);
    if (fabs(lower) >= OR_EPS)
        FN(or_scale)(u, (T)(upper / lower), n);
    else
        FN(or_scale)(u, (T)(upper / (T)OR_EPS), n);
    FN(or_difference)(u, ng, n);
    memcpy(g, ng, sizeof(T) * n);
    memcpy(x, energy, sizeof(T) * n);
}

/* ConjugateGradients<Tag, NONE>::value(Q1, rhs, x, konv_rad):
 * honei/math/conjugate_gradients.hh:490-526 with cg_kernel :108-151.
 * max_iters is a safety cap the reference does not have (0 = none). */
static int FN(or_cg_konv_any)(const FN(or_matrix) *A, const T *rhs, T *x, uint64_t n,
                              double konv_rad, uint64_t max_iters, uint64_t *iters)
{
    T *g = (T *)malloc(sizeof(T) * n), *u = (T *)malloc(sizeof(T) * n);
    T *energy = (T *)malloc(sizeof(T) * n), *ng = (T *)malloc(sizeof(T) * n);
    uint64_t it = 0;
    int capped = 0;

    for (uint64_t i = 0; i < n; ++i)
        x[i] = 0;
    FN(or_apply)(A, g, x, n);
    FN(or_difference)(g, rhs, n);
    memcpy(u, g, sizeof(T) * n);
    FN(or_scale)(u, (T)-1., n);
    T norm_x = FN(or_norm_l2)(x, n, 0);
    T norm_x_last = (T)(norm_x - 1);
    while (norm_x - norm_x_last > konv_rad) {
        if (max_iters && it >= max_iters) {
            capped = 1;
            break;
        }
        FN(or_cg_kernel_v1)(A, rhs, g, x, u, energy, ng, n);
        ++it;
        norm_x_last = norm_x;
        norm_x = FN(or_norm_l2)(x, n, 0);
    }
    *iters = it;
    free(g); free(u); free(energy); free(ng);
    return capped;
}

/* ConjugateGradients<Tag, NONE>::value(A, rhs, x, konv_rad): Q1
 * conjugate_gradients.hh:490-526 with kernel :108-151; ELL :588-620 with
 * kernel :153-200 (the same steps on ELL products). */
int FN(or_cg_konv_q1)(const T *const *bands, const T *rhs, T *x, uint64_t n,
                      double konv_rad, uint64_t max_iters, uint64_t *iters)
{
    FN(or_matrix) A = {0, bands, NULL, NULL, NULL, 0, 0};
    return FN(or_cg_konv_any)(&A, rhs, x, n, konv_rad, max_iters, iters);
}

int FN(or_cg_konv_ell)(const OR_CAT(or_ell, OR_S) *M, const T *rhs, T *x, double konv_rad,
                       uint64_t max_iters, uint64_t *iters)
{
    FN(or_matrix) A = {1, NULL, M->Aj, M->Ax, M->Arl, M->stride, M->threads};
    return FN(or_cg_konv_any)(&A, rhs, x, M->rows, konv_rad, max_iters, iters);
}

/* ConjugateGradients<Tag, NONE>::value(Q1, rhs, x, unsigned long iters):
 * honei/math/conjugate_gradients.hh:526-569 with cg_kernel :108-151.
 * x = 0; g = A x - rhs; u = (copy of g) * -1; initial = ||g||^2 (the
 * squared Norm<vnt_l_two, false>).  Up to `iters` times: one cg_kernel
 * step, then the squared norm of the true defect rhs - A x, stopping when
 * it is below initial * 1e-8 (the product formed in double).  *used = the
 * steps taken; returns 0 when that test fired. */
int FN(or_cg_iters_q1)(const T *const *bands, const T *rhs, T *x, uint64_t n, uint64_t iters,
                       uint64_t *used)
{
    FN(or_matrix) A = {0, bands, NULL, NULL, NULL, 0, 0};
    T *g = (T *)malloc(sizeof(T) * n), *u = (T *)malloc(sizeof(T) * n);
    T *energy = (T *)malloc(sizeof(T) * n), *ng = (T *)malloc(sizeof(T) * n);
    T *r = (T *)malloc(sizeof(T) * n);
    int converged = 0;
    for (uint64_t i = 0; i < n; ++i)
        x[i] = 0;
    FN(or_apply)(&A, g, x, n);
    FN(or_difference)(g, rhs, n);
    memcpy(u, g, sizeof(T) * n);        /* g_c = copy(g) */
    FN(or_scale)(u, (T)-1., n);         /* Scale(g_c, -1); u = copy(g_c) */
    const T initial_norm = FN(or_norm_l2)(g, n, 0);
    uint64_t i = 0;
    while (i < iters) {
        ++i;
        FN(or_cg_kernel_v1)(&A, rhs, g, x, u, energy, ng, n);
        FN(or_q1_defect)(r, rhs, bands, x, n);
        const T current_norm = FN(or_norm_l2)(r, n, 0);
        if ((double)current_norm < (double)initial_norm * 1e-08) {
            converged = 1;
            break;
        }
    }
    *used = i;
    free(g); free(u); free(energy); free(ng); free(r);
    return converged ? 0 : 1;
}

/* ConjugateGradients<Tag, NONE>::value(ELL, rhs, x, unsigned long
 * max_iters): honei/math/conjugate_gradients.hh:571-586 with the
 * "version 2" cg_kernel :203-248.  CGSolver<NONE>'s steps from the given
 * x, without its epsilon guards (lambda = alpha / (v.p), p = p *
 * (alpha / alpha_old) + r), stopping when ||r|| < 1e-8 ||r0|| or
 * ||r|| < 1e-8 (both formed in double) or after max_iters.  *used = the
 * iterations run; returns 0 when a stopping test fired. */
int FN(or_cg_iters_ell)(const OR_CAT(or_ell, OR_S) *M, const T *rhs, T *x, uint64_t max_iters, uint64_t *used)
{
    FN(or_matrix) A = {1, NULL, M->Aj, M->Ax, M->Arl, M->stride, M->threads};
    const uint64_t n = M->rows;
    T *r = (T *)malloc(sizeof(T) * n), *p = (T *)malloc(sizeof(T) * n), *v = (T *)malloc(sizeof(T) * n);
    int converged = 0;
    FN(or_defect_any)(&A, r, rhs, x, n);
    memcpy(p, r, sizeof(T) * n);
    T alpha = FN(or_norm_l2)(r, n, 0);
    const T initial_defect = (T)sqrt(alpha);
    uint64_t iterations = 0;
    while (iterations < max_iters) {
        FN(or_apply)(&A, v, p, n);
        const T lambda = alpha / FN(or_dot)(v, p, n);
        FN(or_scaled_sum)(x, p, lambda, n);
        FN(or_scaled_sum)(r, v, -lambda, n);
        const T alpha_old = alpha;
        alpha = FN(or_norm_l2)(r, n, 0);
        FN(or_scale)(p, alpha / alpha_old, n);
        FN(or_sum)(p, r, n);
        ++iterations;
        const T current_defect = (T)sqrt(alpha);
        if ((double)current_defect < 1e-08 * (double)initial_defect || (double)current_defect < 1e-08) {
            converged = 1;
            break;
        }
    }
    *used = iterations;
    free(r); free(p); free(v);
    return converged ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* Operator-list multigrid: MGCycleCreation (honei/math/mg.hh:560-1060),     */
/* MGSolver (mg.hh:1064-1100), RISmoother (ri.hh:96-134), PROLMAT transfers */
/* (restriction.hh:174-185, prolongation.hh:141-155).                        */
/* ------------------------------------------------------------------------ */
typedef struct FN(or_mgop) {
    const or_mgop_params *p;
    const OR_CAT(or_ell, OR_S) *A, *prol, *res;
    const T *const *P;                     /* DenseVector preconditioners, or NULL */
    const OR_CAT(or_ell, OR_S) *Pm;        /* SparseMatrixELL preconditioners, or NULL */
    T **c, **d, **store, **t0, **t1;    /* per internal level */
    uint64_t used_coarse;
} FN(or_mgop);

static void FN(mgop_product)(const OR_CAT(or_ell, OR_S) *M, T *y, const T *x)
{
    FN(or_ell_product)(y, M->Aj, M->Ax, M->Arl, x, 0, M->rows, M->stride, M->threads);
}

static void FN(mgop_defect)(const OR_CAT(or_ell, OR_S) *M, T *r, const T *b, const T *x)
{
    FN(or_ell_defect)(r, b, M->Aj, M->Ax, M->Arl, x, M->rows, M->stride, M->threads);
}

/* RISmoother::value: iters times { t0 = b - A x; t1 = P * t0; x += t1 } */
static void FN(mgop_smooth)(FN(or_mgop) *g, uint64_t i, T *x, const T *b, uint64_t iters)
{
    const uint64_t n = g->A[i].rows;
    for (uint64_t k = 0; k < iters; ++k) {
        FN(mgop_defect)(&g->A[i], g->t0[i], b, x);
        FN(or_precon)(g->P ? g->P[i] : NULL, g->Pm ? &g->Pm[i] : NULL, g->t1[i], g->t0[i], n);
        FN(or_sum)(x, g->t1[i], n);
    }
}

/* SolverOperator on the coarsest level: CGSolver<NONE> or <VAR> */
static void FN(mgop_coarse)(FN(or_mgop) *g, uint64_t i, T *x, const T *b)
{
    const OR_CAT(or_ell, OR_S) *M = &g->A[i];
    uint64_t used = 0;
    if (g->p->coarse_var) {
        FN(or_matrix) Am = {1, 0, M->Aj, M->Ax, M->Arl, M->stride, M->threads};
        FN(or_cg_var_any)(&Am, g->P ? g->P[i] : NULL, g->Pm ? &g->Pm[i] : NULL, b, x, M->rows,
                          g->p->max_iters_coarse, (T)g->p->eps_relative, &used);
    }
    else
        FN(or_cg_none_ell)(M->Aj, M->Ax, M->Arl, M->stride, M->threads, b, x, M->rows,
                           g->p->max_iters_coarse, (T)g->p->eps_relative, &used);
    g->used_coarse = used;
}

/* One cycle of shape `shape` (0 V, 1 W, 2 F) on base level `level` with
 * iterate x and right-hand side b; coarser levels use c (iterate) and d. */
static void FN(mgop_cycle)(FN(or_mgop) *g, int shape, uint64_t level, T *x, const T *b)
{
    const uint64_t i = level - 1;
    if (level == g->p->min_level) {
        FN(mgop_coarse)(g, i, x, b);
        return;
    }
    FN(mgop_smooth)(g, i, x, b, g->p->n_pre_smooth);
    FN(mgop_defect)(&g->A[i], g->store[i], b, x);
    FN(mgop_product)(&g->res[i - 1], g->d[i - 1], g->store[i]);   /* Restriction<PROLMAT> */
    memset(g->c[i - 1], 0, sizeof(T) * g->A[i - 1].rows);         /* FillOperator */
    if (shape == 0) {
        FN(mgop_cycle)(g, 0, level - 1, g->c[i - 1], g->d[i - 1]);
    } else if (shape == 1) {
        if (level - 1 != g->p->min_level)
            FN(mgop_cycle)(g, 1, level - 1, g->c[i - 1], g->d[i - 1]);
        FN(mgop_cycle)(g, 1, level - 1, g->c[i - 1], g->d[i - 1]);
    } else {
        FN(mgop_cycle)(g, 2, level - 1, g->c[i - 1], g->d[i - 1]);
        FN(mgop_cycle)(g, 0, level - 1, g->c[i - 1], g->d[i - 1]);
    }
    FN(mgop_product)(&g->prol[i - 1], g->store[i], g->c[i - 1]);  /* Prolongation<PROLMAT> */
    FN(or_sum)(x, g->store[i], g->A[i].rows);                     /* SumOperator */
    FN(mgop_smooth)(g, i, x, b, g->p->n_post_smooth);
}

static int FN(or_mgop_solve_any)(const or_mgop_params *p, const OR_CAT(or_ell, OR_S) *A, const T *const *P,
                                 const OR_CAT(or_ell, OR_S) *Pm, const OR_CAT(or_ell, OR_S) *prol,
                                 const OR_CAT(or_ell, OR_S) *res, const T *b, T *x, uint64_t *used_iters,
                                 uint64_t *used_iters_coarse, T *final_norm)
{
    if (p->levels < 1 || p->min_level < 1 || p->min_level > p->levels || p->cycle < 0 || p->cycle > 2)
        return -1;
    const uint64_t L = p->levels, top = L - 1, n = A[top].rows;
    FN(or_mgop) g;
    g.p = p;
    g.A = A;
    g.P = P;
    g.Pm = Pm;
    g.prol = prol;
    g.res = res;
    g.used_coarse = 0;
    T **vec[5] = {NULL, NULL, NULL, NULL, NULL};
    for (int v = 0; v < 5; ++v) {
        vec[v] = (T **)calloc(L, sizeof(T *));
        for (uint64_t i = 0; i < L; ++i)
            vec[v][i] = (T *)calloc(A[i].rows ? A[i].rows : 1, sizeof(T));
    }
    g.c = vec[0]; g.d = vec[1]; g.store = vec[2]; g.t0 = vec[3]; g.t1 = vec[4];

    T *r = (T *)malloc(sizeof(T) * n);
    FN(mgop_defect)(&A[top], r, b, x);
    const T rnorm_initial = FN(or_norm_l2)(r, n, 1);
    T rnorm_current = (T)1e16;
    *used_iters = 0;
    for (uint64_t it = 0; it < p->max_iters; ++it) {
        FN(mgop_cycle)(&g, p->cycle, L, x, b);
        FN(mgop_defect)(&A[top], r, b, x);
        rnorm_current = FN(or_norm_l2)(r, n, 1);
        *used_iters = it + 1;
        if (rnorm_current < (T)p->eps_relative * rnorm_initial)
            break;
    }
    if (used_iters_coarse)
        *used_iters_coarse = g.used_coarse;
    if (final_norm)
        *final_norm = rnorm_current;
    free(r);
    for (int v = 0; v < 5; ++v) {
        for (uint64_t i = 0; i < L; ++i)
            free(vec[v][i]);
        free(vec[v]);
    }
    return 0;
}

/* MGSolver over the cycle with DenseVector preconditioners P[level]
 * (PreconFill<DenseVector>: "jac"). */
int FN(or_mgop_solve)(const or_mgop_params *p, const OR_CAT(or_ell, OR_S) *A, const T *const *P,
                      const OR_CAT(or_ell, OR_S) *prol, const OR_CAT(or_ell, OR_S) *res, const T *b, T *x,
                      uint64_t *used_iters, uint64_t *used_iters_coarse, T *final_norm)
{
    return FN(or_mgop_solve_any)(p, A, P, NULL, prol, res, b, x, used_iters, used_iters_coarse, final_norm);
}

/* The same with SparseMatrixELL preconditioners Pm[level] (PreconFill for a
 * matrix PreconContType_, mg.hh:315-335: the level's <A_i>_<suffix>.ell
 * file): RISmoother applies t1 = Pm t0 and CGSolver<VAR> z = Pm r. */
int FN(or_mgop_solve_pm)(const or_mgop_params *p, const OR_CAT(or_ell, OR_S) *A, const OR_CAT(or_ell, OR_S) *Pm,
                         const OR_CAT(or_ell, OR_S) *prol, const OR_CAT(or_ell, OR_S) *res, const T *b, T *x,
                         uint64_t *used_iters, uint64_t *used_iters_coarse, T *final_norm)
{
    return FN(or_mgop_solve_any)(p, A, NULL, Pm, prol, res, b, x, used_iters, used_iters_coarse, final_norm);
}

/* ------------------------------------------------------------------------ */
/* Multigrid<Tag, OuterTag, ProlType, JAC, CYCLE::V, FIXED|MIXED>: the      */
/* V-cycle kernel _multigrid_kernel (multigrid.hh:160-466 Q1 and 469-820 ELL */
/* FIXED, 1620-1933 Q1 and 1936-2335 ELL MIXED: the same steps on either    */
/* matrix kind), with structured (NONE) or matrix (PROLMAT) transfers.       */
/* ------------------------------------------------------------------------ */
typedef struct FN(or_mgl) {
    uint64_t n;
    FN(or_matrix) A;
    const T *dinv;
    const OR_CAT(or_ell, OR_S) *prol;   /* level-1 -> level (PROLMAT), or NULL */
    const OR_CAT(or_ell, OR_S) *res;    /* level -> level-1 (PROLMAT), or NULL */
    T *c, *d, *rhs, *x, *tmp;
} FN(or_mgl);

static void FN(or_mgl_jacobi)(FN(or_mgl) *l, T *x, const T *rhs, uint64_t iters)
{
    for (uint64_t it = 0; it < iters; ++it) {
        FN(or_defect_any)(&l->A, l->tmp, rhs, x, l->n);
        FN(or_scaled_sum_vv)(x, l->tmp, l->dinv, l->n);
    }
}

static void FN(or_mgl_product)(const OR_CAT(or_ell, OR_S) *M, T *y, const T *x)
{
    FN(or_ell_product)(y, M->Aj, M->Ax, M->Arl, x, 0, M->rows, M->stride, M->threads);
}

/* lv indexed by level; returns the V-cycles run (MGInfo::inner_iterations) */
static uint64_t FN(or_mg_inner)(FN(or_mgl) *lv, const or_mg_params *p, T tol, T adapt, T *defect_out,
                                T *initial_out)
{
    const uint64_t minl = p->min_level, maxl = p->max_level;
    FN(or_mgl) *top = &lv[maxl];
    if (p->initial_zero)
        memcpy(top->d, top->rhs, sizeof(T) * top->n);
    else
        FN(or_defect_any)(&top->A, top->d, top->rhs, top->x, top->n);
    T defect = FN(or_norm_l2)(top->d, top->n, 1);
    const T initial_defect = defect;
    if (initial_out) *initial_out = initial_defect;
    if (defect_out) *defect_out = defect;
    if (p->convergence_check && defect <= tol)
        return 0;
    uint64_t iter;
    for (iter = 1; iter <= p->n_max_iter; ++iter) {
        uint64_t L = maxl;
        int restriction_started = 1;
        while (minl != maxl) {
            FN(or_mgl) *l = &lv[L];
            if (restriction_started) {
                FN(or_element_product3)(l->c, l->dinv, l->d, l->n);
                FN(or_mgl_jacobi)(l, l->c, l->d, p->n_pre_smooth - 1);
                FN(or_sum)(l->x, l->c, l->n);
            } else {
                FN(or_element_product3)(l->x, l->dinv, l->d, l->n);
                FN(or_mgl_jacobi)(l, l->x, l->d, p->n_pre_smooth - 1);
            }
            FN(or_defect_any)(&l->A, l->d, l->rhs, l->x, l->n);
            --L;
            if (lv[L + 1].res)
                FN(or_mgl_product)(lv[L + 1].res, lv[L].d, lv[L + 1].d);
            else
                FN(or_restriction)(lv[L].d, lv[L + 1].d, lv[L + 1].n, p->macro_border_mask);
            memcpy(lv[L].rhs, lv[L].d, sizeof(T) * lv[L].n);
            if (L == minl)
                break;
            restriction_started = 0;
        }
        {
            uint64_t cg_iters;
            FN(or_cg_konv_any)(&lv[L].A, lv[L].d, lv[L].x, lv[L].n, OR_EPS, 0, &cg_iters);
        }
        if (minl == maxl) {
            FN(or_defect_any)(&lv[L].A, lv[L].d, lv[L].rhs, lv[L].x, lv[L].n);
            break;
        }
        while (1) {
            ++L;
            FN(or_mgl) *l = &lv[L];
            if (l->prol)
                FN(or_mgl_product)(l->prol, l->c, lv[L - 1].x);
            else
                FN(or_prolongation)(l->c, lv[L - 1].x, l->n, p->macro_border_mask);
            T alpha;
            if (adapt == (T)0) {
                T d1 = FN(or_dot)(l->d, l->c, l->n);
                FN(or_apply)(&l->A, l->tmp, l->c, l->n);
                T d2 = FN(or_dot)(l->tmp, l->c, l->n);
                alpha = d1 / d2;
            } else {
                alpha = adapt;
            }
            FN(or_scaled_sum)(l->x, l->c, alpha, l->n);
            FN(or_mgl_jacobi)(l, l->x, l->rhs, p->n_pre_smooth);
            FN(or_defect_any)(&l->A, l->d, l->rhs, l->x, l->n);
            if (L == maxl)
                break;
        }
        defect = FN(or_norm_l2)(top->d, top->n, 1);
        if (defect_out) *defect_out = defect;
        if (defect <= initial_defect * tol)
            break;
    }
    return (iter > p->n_max_iter) ? iter - 1 : iter;
}

/* Levels min..max of either kind (kind 0: a_q1[9 L + b]; kind 1: a_ell[L]),
 * dinv[L], PROLMAT matrices prol[L] / res[L] for L in (min, max] or NULL. */
static int FN(or_mg_levels)(FN(or_mgl) *lv, int kind, const T *const *a_q1, const OR_CAT(or_ell, OR_S) *a_ell,
                            const T *const *dinv, const OR_CAT(or_ell, OR_S) *prol,
                            const OR_CAT(or_ell, OR_S) *res, const or_mg_params *p)
{
    memset(lv, 0, sizeof(FN(or_mgl)) * 15);
    for (uint64_t L = p->min_level; L <= p->max_level; ++L) {
        FN(or_mgl) *l = &lv[L];
        if (kind == 0) {
            l->n = ((1ull << L) + 1) * ((1ull << L) + 1);
            FN(or_matrix) A = {0, a_q1 + 9 * L, NULL, NULL, NULL, 0, 0};
            l->A = A;
        } else {
            const OR_CAT(or_ell, OR_S) *M = &a_ell[L];
            l->n = M->rows;
            FN(or_matrix) A = {1, NULL, M->Aj, M->Ax, M->Arl, M->stride, M->threads};
            l->A = A;
        }
        l->dinv = dinv[L];
        if (L > p->min_level && prol && res) {
            l->prol = &prol[L];
            l->res = &res[L];
        }
        l->c = (T *)calloc(l->n, sizeof(T));
        l->d = (T *)calloc(l->n, sizeof(T));
        l->rhs = (T *)calloc(l->n, sizeof(T));
        l->x = (T *)calloc(l->n, sizeof(T));
        l->tmp = (T *)calloc(l->n, sizeof(T));
    }
    return 0;
}

static void FN(or_mg_free)(FN(or_mgl) *lv, const or_mg_params *p)
{
    for (uint64_t L = p->min_level; L <= p->max_level; ++L) {
        free(lv[L].c); free(lv[L].d); free(lv[L].rhs); free(lv[L].x); free(lv[L].tmp);
    }
}

/* Multigrid<..., FIXED>::value (multigrid.hh:790-872): x0 with b injected on
 * the first row, the last row and the first column, then one inner solve
 * in this precision. */
int FN(or_mg_fixed)(int kind, const T *const *a_q1, const OR_CAT(or_ell, OR_S) *a_ell, const T *const *dinv,
                    const OR_CAT(or_ell, OR_S) *prol, const OR_CAT(or_ell, OR_S) *res, const T *b, T *x,
                    const or_mg_params *p, uint64_t *iters, T *final_defect, T *initial_defect)
{
    if (p->max_level < p->min_level || p->max_level > 14 || p->n_pre_smooth < 1)
        return -1;
    FN(or_mgl) lv[15];
    FN(or_mg_levels)(lv, kind, a_q1, a_ell, dinv, prol, res, p);
    FN(or_mgl) *top = &lv[p->max_level];
    const uint64_t n = top->n;
    const uint64_t sq = (uint64_t)sqrt((double)n);
    memcpy(top->x, x, sizeof(T) * n);
    for (uint64_t i = 0; i < sq; ++i) top->x[i] = b[i];
    for (uint64_t i = n - sq; i < n; ++i) top->x[i] = b[i];
    for (uint64_t i = 0; i < n; i += sq) top->x[i] = b[i];
    memcpy(top->rhs, b, sizeof(T) * n);
    T fd = 0, id = 0;
    *iters = FN(or_mg_inner)(lv, p, (T)p->tolerance, (T)p->adapt_correction_factor, &fd, &id);
    memcpy(x, top->x, sizeof(T) * n);
    if (final_defect) *final_defect = fd;
    if (initial_defect) *initial_defect = id;
    FN(or_mg_free)(lv, p);
    return 0;
}

#undef OR_EPS
#undef T
#undef FN
#undef OR_CAT
#undef OR_CAT2
