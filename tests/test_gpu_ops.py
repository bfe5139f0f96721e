"""GPU parity of the linear-algebra and multigrid kernels against the oracle
(the C restatement of HONEI's tags::CPU path).  Products, defects, smoother
sweeps, transfers and elementwise ops are bit-exact; reductions are within
a stated tolerance (their summation tree differs from the sequential CPU
sum)."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import dinv_reference, load_bin, load_ell

pytestmark = pytest.mark.gpu

DTYPES = (np.float64, np.float32)


def _ell_dev(hip, E, dtype):
    """Raw reference-layout ELL arrays on the device (uint64 Aj / Arl)."""
    return (hip.DeviceVector.from_numpy(np.asarray(E["Aj"], np.uint64)),
            hip.DeviceVector.from_numpy(np.asarray(E["Ax"], dtype)),
            hip.DeviceVector.from_numpy(np.asarray(E["Arl"], np.uint64)))


def _raw_product(hip, E, x, dtype, rhs=None):
    from honei_amd._lib import lib, u64, vp
    Aj, Ax, Arl = _ell_dev(hip, E, dtype)
    xd = hip.DeviceVector.from_numpy(x.astype(dtype))
    yd = hip.DeviceVector.zeros(E["rows"], dtype)
    sfx = "double" if dtype == np.float64 else "float"
    if rhs is None:
        rc = getattr(lib(), f"hsolve_product_smell_dv_{sfx}")(
            vp(xd.ptr), vp(yd.ptr), vp(Aj.ptr), vp(Ax.ptr), vp(Arl.ptr), u64(0), u64(E["rows"]),
            u64(E["ncpr"]), u64(E["stride"]), u64(E.get("threads", 1)), None)
    else:
        rd = hip.DeviceVector.from_numpy(rhs.astype(dtype))
        rc = getattr(lib(), f"hsolve_defect_smell_dv_{sfx}")(
            vp(rd.ptr), vp(yd.ptr), vp(Aj.ptr), vp(Ax.ptr), vp(Arl.ptr), vp(xd.ptr), u64(0), u64(E["rows"]),
            u64(E["ncpr"]), u64(E["stride"]), u64(E.get("threads", 1)), None)
    assert rc == 0, lib().hsolve_last_error()
    return yd.get()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", ["A_6.ell", "A_3.ell", "prol_6.ell", "prol_2.ell"])
def test_ell_fixture_product_and_defect(hip, oracle, rng, name, dtype):
    F = load_ell("poisson/" + name)
    E = dict(F)
    E["Ax"] = F["Ax"].astype(dtype)
    E["Arl"] = oracle.ell_row_length(F["rows"], F["stride"], 1, E["Ax"])
    x = rng.uniform(-1, 1, F["cols"]).astype(dtype)
    ref = oracle.ell_product(E, x)
    # reference-layout entry point (uint64 indices)
    assert np.array_equal(_raw_product(hip, E, x, dtype), ref)
    # device handle (uint32 indices, packed row lengths; lengths derived by scan)
    M = hip.EllMatrix(F["rows"], F["cols"], F["stride"], F["ncpr"], 1, F["Aj"], E["Ax"], None)
    xd = hip.DeviceVector.from_numpy(x)
    yd = hip.DeviceVector.zeros(F["rows"], dtype)
    M.spmv(xd, yd)
    assert np.array_equal(yd.get(), ref)
    if F["rows"] == F["cols"]:
        rhs = rng.uniform(-1, 1, F["rows"]).astype(dtype)
        refd = oracle.ell_defect(E, rhs, x)
        assert np.array_equal(_raw_product(hip, E, x, dtype, rhs), refd)
        M.spmv(xd, yd, rhs=hip.DeviceVector.from_numpy(rhs))
        assert np.array_equal(yd.get(), refd)


def _known_answer_csr(size, dtype):
    """product_TEST.cc:773-845: SparseMatrix(size, size+3) with every linear
    element whose index is a multiple of 5 set to index/1.234."""
    cols = size + 3
    lin = np.arange(0, size * cols, 5)
    r, c = lin // cols, lin % cols
    v = (lin.astype(np.float64) / 1.234).astype(dtype)
    rowptr = np.zeros(size + 1, np.uint64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.uint64)
    return rowptr, c.astype(np.uint64), v, cols


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("threads", [1, 2, 4, 8, 16])
def test_ell_known_answer_threads(hip, oracle, dtype, threads):
    for size in (11, 22, 44, 88, 176, 352):
        rowptr, col, val, cols = _known_answer_csr(size, dtype)
        E = oracle.ell_from_csr(size, rowptr, col, val, threads)
        x = np.full(cols, 4, dtype)
        x[0], x[1] = 1, 2
        ref = oracle.ell_product(E, x)
        assert np.array_equal(_raw_product(hip, E, x, dtype), ref)
        M = hip.EllMatrix(size, cols, E["stride"], E["ncpr"], threads, E["Aj"], E["Ax"], E["Arl"])
        yd = hip.DeviceVector.zeros(size, dtype)
        M.spmv(hip.DeviceVector.from_numpy(x), yd)
        assert np.array_equal(yd.get(), ref)
        # the closed form: sum over stored entries of the row
        dense = np.zeros((size, cols), np.float64)
        dense[np.repeat(np.arange(size), np.diff(rowptr.astype(np.int64))), col.astype(np.int64)] = val
        tol = 1e-12 if dtype == np.float64 else 1e-5
        np.testing.assert_allclose(ref, dense @ x.astype(np.float64), rtol=tol, atol=tol * 1e3)


@pytest.mark.parametrize("dtype", DTYPES)
def test_ell_matrix_b_and_a_small(hip, oracle, dtype):
    for E in (P.matrix_b_ell(1 << 15, dtype=dtype), P.matrix_b_ell((1 << 15) - 7, dtype=dtype),
              P.q1_to_ell(P.matrix_a_bands(129, dtype=dtype))):
        n = E["rows"]
        x = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
        ref = oracle.ell_product(E, x)
        rhs = P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
        ref_d = oracle.ell_defect(E, rhs, x)
        M = hip.EllMatrix.from_dict(E)
        yd = hip.DeviceVector.zeros(n, dtype)
        xd, rd = hip.DeviceVector.from_numpy(x), hip.DeviceVector.from_numpy(rhs)
        # every kernel variant gives the same bits (48, 52: rows sorted by length per LDS block;
        # 80, 112, 116: the persistent LDS-window kernel)
        for variant in list(range(8)) + [16, 20, 22, 48, 52, 80, 112, 116]:
            M.set_variant(variant)
            M.spmv(xd, yd)
            assert np.array_equal(yd.get(), ref), variant
            M.spmv(xd, yd, rhs=rd)
            assert np.array_equal(yd.get(), ref_d), variant
        # the Jacobi sweep through the default and the LDS-window kernels
        dinv = (1.0 / (np.abs(x) + 1.0)).astype(dtype)
        ref_j = oracle.jacobi_ell(x, E, rhs, dinv, 1)
        dd = hip.DeviceVector.from_numpy(dinv)
        for variant in (0, 16, 48, 80, 112):
            M.set_variant(variant)
            M.jacobi(rd, dd, xd, yd)
            assert np.array_equal(yd.get(), ref_j), variant


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [9, 25, 4225, 66049])
def test_q1_product_defect_jacobi(hip, oracle, rng, n, dtype):
    bands = (load_bin(n)["bands"] if n <= 4225 else P.matrix_a_bands(int(round(np.sqrt(n)))))
    bands = [b.astype(dtype) for b in bands]
    A = hip.Q1Matrix(bands)
    x = rng.uniform(-1, 1, n).astype(dtype)
    rhs = rng.uniform(-1, 1, n).astype(dtype)
    xd, rd, yd = (hip.DeviceVector.from_numpy(x), hip.DeviceVector.from_numpy(rhs),
                  hip.DeviceVector.zeros(n, dtype))
    hip.q1_product(A, xd, yd)
    assert np.array_equal(yd.get(), oracle.q1_product(bands, x))
    hip.q1_defect(A, rd, xd, yd)
    assert np.array_equal(yd.get(), oracle.q1_defect(bands, rhs, x))
    dinv = dinv_reference(bands[4].astype(np.float32)).astype(dtype)
    dd = hip.DeviceVector.from_numpy(dinv)
    hip.q1_jacobi(A, rd, dd, xd, yd)
    assert np.array_equal(yd.get(), oracle.jacobi_q1(x, bands, rhs, dinv, 1))


@pytest.mark.parametrize("dtype", DTYPES)
def test_reductions(hip, rng, dtype):
    for n in (1, 7, 1000, 2048 * 256 * 4 + 13, 3_000_000):
        a = rng.uniform(-1, 1, n).astype(dtype)
        b = rng.uniform(-1, 1, n).astype(dtype)
        ad, bd = hip.DeviceVector.from_numpy(a), hip.DeviceVector.from_numpy(b)
        exact = float(np.dot(a.astype(np.longdouble), b.astype(np.longdouble)))
        sq = float(np.dot(a.astype(np.longdouble), a.astype(np.longdouble)))
        tol = 1e-12 if dtype == np.float64 else 1e-6
        scale = float(np.dot(np.abs(a).astype(np.float64), np.abs(b).astype(np.float64))) + 1e-300
        assert abs(hip.dot(ad, bd) - exact) <= tol * scale
        assert abs(hip.norm(ad) - sq) <= tol * sq
        assert abs(hip.norm(ad, root=True) - np.sqrt(sq)) <= tol * np.sqrt(sq)
        # deterministic: the same call twice gives the same bits
        assert hip.dot(ad, bd) == hip.dot(ad, bd)
    ones = hip.DeviceVector.from_numpy(np.ones(123457, dtype))
    assert hip.dot(ones, ones) == 123457          # dot_product_TEST.cc:33-60


@pytest.mark.parametrize("dtype", DTYPES)
def test_blas1_bit_exact(hip, oracle, rng, dtype):
    from honei_amd._lib import lib, u64, vp
    n = 100_003
    x = rng.uniform(-2, 2, n).astype(dtype)
    y = rng.uniform(-2, 2, n).astype(dtype)
    z = rng.uniform(-2, 2, n).astype(dtype)
    x[::17] = 0
    s = "double" if dtype == np.float64 else "float"
    L = lib()
    b = dtype(0.731)

    def dv(a):
        return hip.DeviceVector.from_numpy(a)

    r = hip.DeviceVector.zeros(n, dtype)
    xd, yd, zd = dv(x), dv(y), dv(z)
    assert getattr(L, f"hsolve_scaled_sum_three_{s}_s")(vp(r.ptr), vp(xd.ptr), vp(yd.ptr), b, u64(n), None) == 0
    assert np.array_equal(r.get(), oracle.scaled_sum3(x, y, b))
    t = dv(x)
    getattr(L, f"hsolve_scaled_sum_three_{s}_s")(vp(t.ptr), vp(t.ptr), vp(yd.ptr), b, u64(n), None)
    assert np.array_equal(t.get(), oracle.scaled_sum(x, y, b))
    t = dv(x)
    getattr(L, f"hsolve_scaled_sum_three_{s}")(vp(t.ptr), vp(yd.ptr), vp(zd.ptr), u64(n), None)
    assert np.array_equal(t.get(), oracle.scaled_sum_vv(x, y, z))
    t = dv(x)
    getattr(L, f"hsolve_sum_two_{s}")(vp(t.ptr), vp(yd.ptr), u64(n), None)
    assert np.array_equal(t.get(), oracle.sum(x, y))
    t = dv(x)
    getattr(L, f"hsolve_scale_one_{s}")(vp(t.ptr), b, u64(n), None)
    assert np.array_equal(t.get(), oracle.scale(x, b))
    t = dv(x)
    getattr(L, f"hsolve_difference_two_{s}")(vp(t.ptr), vp(yd.ptr), u64(n), None)
    assert np.array_equal(t.get(), oracle.difference(x, y))
    getattr(L, f"hsolve_difference_three_{s}")(vp(r.ptr), vp(xd.ptr), vp(yd.ptr), u64(n), None)
    assert np.array_equal(r.get(), oracle.difference3(x, y))
    getattr(L, f"hsolve_element_product_three_{s}")(vp(r.ptr), vp(xd.ptr), vp(yd.ptr), u64(n), None)
    assert np.array_equal(r.get(), oracle.element_product3(x, y))
    t = dv(x)
    getattr(L, f"hsolve_element_inverse_one_{s}")(vp(t.ptr), u64(n), None)
    assert np.array_equal(t.get(), oracle.element_inverse(x))


def test_convert(hip, oracle, rng):
    from honei_amd._lib import lib, u64, vp
    n = 54321
    d = rng.normal(0, 1e3, n)
    d[:5] = [0.1, 1e-40, -1e-45, 3.4e38, -7.0]
    dd = hip.DeviceVector.from_numpy(d)
    f = hip.DeviceVector.zeros(n, np.float32)
    assert lib().hsolve_convert_double_float(vp(dd.ptr), vp(f.ptr), u64(n), None) == 0
    assert np.array_equal(f.get(), oracle.convert_d2f(d))
    back = hip.DeviceVector.zeros(n, np.float64)
    assert lib().hsolve_convert_float_double(vp(f.ptr), vp(back.ptr), u64(n), None) == 0
    assert np.array_equal(back.get(), oracle.convert_f2d(f.get()))


MASKS = [(2,) * 8, (1,) * 8, (2, 1, 2, 1, 2, 1, 2, 1), (1, 1, 1, 1, 2, 2, 1, 1)]


@pytest.mark.parametrize("dtype", DTYPES)
def test_transfers_bit_exact(hip, oracle, rng, dtype):
    """restriction_TEST.cc / prolongation_TEST.cc: GPU equals CPU exactly, levels 1..10."""
    for L in range(1, 11):
        nf, nc = P.side_of_level(L) ** 2, P.side_of_level(L - 1) ** 2
        f = rng.uniform(-1, 1, nf).astype(dtype)
        c = rng.uniform(-1, 1, nc).astype(dtype)
        fd, cd = hip.DeviceVector.from_numpy(f), hip.DeviceVector.from_numpy(c)
        for mask in (MASKS if L <= 4 else MASKS[:1]):
            out_c = hip.DeviceVector.zeros(nc, dtype)
            hip.restriction(out_c, fd, mask)
            assert np.array_equal(out_c.get(), oracle.restriction(f, mask)), (L, mask)
            out_f = hip.DeviceVector.zeros(nf, dtype)
            hip.prolongation(out_f, cd, mask)
            assert np.array_equal(out_f.get(), oracle.prolongation(c, nf, mask)), (L, mask)
            x = rng.uniform(-1, 1, nf).astype(dtype)
            xd = hip.DeviceVector.from_numpy(x)
            alpha = dtype(0.75)
            hip.prolongation_axpy(xd, cd, alpha, mask)
            assert np.array_equal(xd.get(), oracle.scaled_sum(x, oracle.prolongation(c, nf, mask), alpha))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [9, 25, 81, 289])
def test_coarse_cg_bit_exact(hip, oracle, rng, n, dtype):
    F = load_bin(n)
    bands = [b.astype(dtype) for b in F["bands"]]
    side = int(round(np.sqrt(n)))
    rhs = rng.uniform(-1, 1, n).astype(dtype)
    rhs[P.boundary_mask(side)] = 0
    konv = float(np.finfo(dtype).eps)
    x_ref, it_ref, _ = oracle.cg_konv_q1(bands, rhs, konv)
    A = hip.Q1Matrix(bands)
    xd = hip.DeviceVector.zeros(n, dtype)
    it = hip.cg_konv_q1(A, hip.DeviceVector.from_numpy(rhs), xd, konv)
    assert it == it_ref
    assert np.array_equal(xd.get(), x_ref)


def _banded_ell_tail(rows, cols, maxoff, ncpr, dtype, seed=5):
    """An ELL matrix with rows > cols: rows below `cols` hold 1..ncpr-2
    entries at distinct columns within +-maxoff of the row, every row from
    `cols` on is empty (the LDS-window kernel sees whole blocks of them)."""
    g = np.random.default_rng(seed)
    stride = rows
    Aj = np.zeros(ncpr * stride, np.uint64)
    Ax = np.zeros(ncpr * stride, dtype)
    Arl = np.zeros(rows, np.uint64)
    for r in range(min(rows, cols)):
        lo, hi = max(0, r - maxoff), min(cols, r + maxoff + 1)
        k = int(g.integers(1, ncpr - 1))
        c = np.sort(g.choice(np.arange(lo, hi), size=min(k, hi - lo), replace=False))
        Arl[r] = c.size
        for j, cc in enumerate(c):
            Aj[j * stride + r] = cc
            Ax[j * stride + r] = dtype(g.uniform(0.5, 1.5))
    return dict(rows=rows, cols=cols, stride=stride, ncpr=ncpr, threads=1, Aj=Aj, Ax=Ax, Arl=Arl)


@pytest.mark.parametrize("dtype", DTYPES)
def test_ell_lds_empty_tail_and_zero_columns(hip, oracle, rng, dtype):
    """Handles the LDS-window kernel is chosen for, with blocks whose x
    window is empty: rows > cols with an empty tail of more than one LDS
    block (1024 rows FP64, 512 FP32), and a matrix with no columns."""
    E = _banded_ell_tail(4096, 1000, 30, 8, dtype)
    x = rng.uniform(-1, 1, E["cols"]).astype(dtype)
    ref = oracle.ell_product(E, x)
    assert np.all(ref[1000:] == 0)
    M = hip.EllMatrix.from_dict(E)
    xd = hip.DeviceVector.from_numpy(x)
    yd = hip.DeviceVector.from_numpy(np.full(E["rows"], 7, dtype))
    for variant in (0, 16, 20, 48, 52, 80, 112, 116):
        M.set_variant(variant)
        M.spmv(xd, yd)
        assert np.array_equal(yd.get(), ref), variant
    # no columns: every row empty, y = 0
    Z = dict(rows=2048, cols=0, stride=2048, ncpr=1, threads=1, Aj=np.zeros(2048, np.uint64),
             Ax=np.zeros(2048, dtype), Arl=np.zeros(2048, np.uint64))
    MZ = hip.EllMatrix.from_dict(Z)
    x1 = hip.DeviceVector.zeros(1, dtype)
    y0 = hip.DeviceVector.from_numpy(np.full(2048, 3, dtype))
    for variant in (0, 16, 48, 112):
        MZ.set_variant(variant)
        MZ.spmv(x1, y0)
        assert np.all(y0.get() == 0), variant


@pytest.mark.parametrize("dtype", DTYPES)
def test_ell_lds_window_alignment(hip, oracle, rng, dtype):
    """The LDS-window kernel loads x in 16-byte chunks from a window start
    rounded down to a 16-byte boundary (odd max offset: the rounding moves
    the start; a window end that is not a whole chunk: the tail loads).  An
    x that is not 16-byte aligned (a view at an element offset) takes the
    row kernels; every offset gives the oracle's bits."""
    E = _banded_ell_tail(5004, 4999, 31, 10, dtype, seed=11)
    n, c = E["rows"], E["cols"]
    x = rng.uniform(-1, 1, c).astype(dtype)
    ref = oracle.ell_product(E, x)
    M = hip.EllMatrix.from_dict(E)
    big = hip.DeviceVector.zeros(c + 8, dtype)
    for off in range(4):
        xv = big.view(off, c)
        host = np.zeros(c + 8, dtype)
        host[off:off + c] = x
        big.upload(host)
        for variant in (0, 16, 48):
            M.set_variant(variant)
            yd = hip.DeviceVector.from_numpy(np.full(n, 9, dtype))
            M.spmv(xv, yd)
            assert np.array_equal(yd.get(), ref), (off, variant)


@pytest.mark.parametrize("dtype", DTYPES)
def test_raw_launcher_empty_range_is_noop(hip, dtype):
    """row_end <= row_start: y unchanged, status OK."""
    from honei_amd._lib import lib, u64, vp
    E = P.q1_to_ell(P.matrix_a_bands(17, dtype=dtype))
    Aj, Ax, Arl = _ell_dev(hip, E, dtype)
    xd = hip.DeviceVector.from_numpy(np.ones(E["rows"], dtype))
    yd = hip.DeviceVector.from_numpy(np.full(E["rows"], 5, dtype))
    sfx = "double" if dtype == np.float64 else "float"
    for r0, r1 in ((0, 0), (10, 10), (20, 3)):
        rc = getattr(lib(), f"hsolve_product_smell_dv_{sfx}")(
            vp(xd.ptr), vp(yd.ptr), vp(Aj.ptr), vp(Ax.ptr), vp(Arl.ptr), u64(r0), u64(r1),
            u64(E["ncpr"]), u64(E["stride"]), u64(1), None)
        assert rc == 0
        assert np.all(yd.get() == 5)


@pytest.mark.parametrize("dtype", DTYPES)
def test_ell_spmv_rows_equals_whole_product(hip, oracle, dtype):
    """hsolve_ell_spmv_rows on [r0, r1): those rows of the whole product and
    defect bit for bit, every other entry of y untouched, for even and odd
    starts and ends, one row, the last row, an eighth of the rows (one
    rank's share of a row split) and through every kernel variant; an empty
    range is a no-op and row_end beyond the matrix is refused."""
    from honei_amd._lib import lib, u64, vp
    for E in (P.matrix_b_ell(1 << 15, dtype=dtype), P.matrix_b_ell((1 << 15) - 7, dtype=dtype),
              P.q1_to_ell(P.matrix_a_bands(129, dtype=dtype))):
        n = E["rows"]
        x = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
        rhs = P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
        ref, ref_d = oracle.ell_product(E, x), oracle.ell_defect(E, rhs, x)
        M = hip.EllMatrix.from_dict(E)
        xd, rd = hip.DeviceVector.from_numpy(x), hip.DeviceVector.from_numpy(rhs)
        eighth = n // 8
        ranges = ((0, n), (1, n - 1), (3, 1000), (2, 1001), (3 * eighth, 4 * eighth), (7 * eighth, n),
                  (n - 1, n), (n - 3, n), (17, 18))
        for variant in (0, 1, 2, 4, 16, 48, 80, 112):
            M.set_variant(variant)
            for r0, r1 in ranges:
                fill = np.full(n, 7, dtype)
                yd = hip.DeviceVector.from_numpy(fill)
                M.spmv_rows(xd, yd, r0, r1)
                got = yd.get()
                assert np.array_equal(got[r0:r1], ref[r0:r1]), (variant, r0, r1)
                assert np.all(got[:r0] == 7) and np.all(got[r1:] == 7), (variant, r0, r1)
                M.spmv_rows(xd, yd, r0, r1, rhs=rd)
                got = yd.get()
                assert np.array_equal(got[r0:r1], ref_d[r0:r1]), (variant, r0, r1)
                assert np.all(got[:r0] == 7) and np.all(got[r1:] == 7), (variant, r0, r1)
        yd = hip.DeviceVector.from_numpy(np.full(n, 5, dtype))
        for r0, r1 in ((0, 0), (10, 10), (20, 3)):
            M.spmv_rows(xd, yd, r0, r1)
            assert np.all(yd.get() == 5)
        rc = lib().hsolve_ell_spmv_rows(vp(M.handle), vp(xd.ptr), vp(yd.ptr), None, u64(0), u64(n + 1), None)
        assert rc == 2 and np.all(yd.get() == 5)
