"""GPU parity of the blocked-CSR product and defect (hsolve_product_csr_dv_*,
hsolve_defect_csr_dv_*, and the hsolve_csr handle) against the oracle, bit
for bit: the matrices of product_TEST.cc:917-957 for blocksizes 1..8, the
FEATFLOW2 ELL fixtures converted as SparseMatrixCSR(SparseMatrixELL) does,
and random matrices with empty rows and rows far longer than one LDS chunk."""
import numpy as np
import pytest

from honei_fixtures import load_ell

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32]


def _every_fifth(size, dtype):
    cols = size + 3
    idx = np.arange(size * cols)
    sel = idx[idx % 5 == 0]
    r, c = sel // cols, sel % cols
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=size))]).astype(np.uint64)
    return cols, rowptr, c.astype(np.uint64), (sel / 1.234).astype(dtype)


def _random_rows(rng, rows, cols, lengths, dtype):
    rowptr = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)
    col = np.concatenate([np.sort(rng.choice(cols, size=int(n), replace=False)) if n else np.zeros(0, np.int64)
                          for n in lengths]).astype(np.uint64)
    val = rng.standard_normal(col.size).astype(dtype)
    val[rng.random(col.size) < 0.05] = 0          # stored zeros stay stored
    return rowptr, col, val


def _raw(hip, M, x, rhs=None, row_start=0, row_end=None, y0=None):
    from honei_amd._lib import lib, u64, vp
    dtype = M["Ax"].dtype
    rows = M["rows"]
    row_end = rows if row_end is None else row_end
    Ar = hip.DeviceVector.from_numpy(M["Ar"])
    Aj = hip.DeviceVector.from_numpy(M["Aj"] if M["Aj"].size else np.zeros(1, np.uint64))
    Ax = hip.DeviceVector.from_numpy(M["Ax"] if M["Ax"].size else np.zeros(1, dtype))
    xd = hip.DeviceVector.from_numpy(x)
    yd = hip.DeviceVector.from_numpy(np.full(max(rows, 1), 4711, dtype) if y0 is None else y0)
    sfx = "double" if dtype == np.float64 else "float"
    if rhs is None:
        rc = getattr(lib(), f"hsolve_product_csr_dv_{sfx}")(
            vp(xd.ptr), vp(yd.ptr), vp(Aj.ptr), vp(Ax.ptr), vp(Ar.ptr), u64(row_start), u64(row_end),
            u64(M["bs"]), u64(M["cols"]), None)
    else:
        rd = hip.DeviceVector.from_numpy(rhs)
        rc = getattr(lib(), f"hsolve_defect_csr_dv_{sfx}")(
            vp(rd.ptr), vp(yd.ptr), vp(Aj.ptr), vp(Ax.ptr), vp(Ar.ptr), vp(xd.ptr), u64(rows), u64(M["bs"]),
            u64(M["cols"]), None)
    assert rc == 0, lib().hsolve_last_error()
    return yd.get()[:rows]


def _handle(hip, M, x, rhs=None, row_start=0, row_end=0, y0=None, variant=0):
    H = hip.CsrMatrix.from_dict(M)
    H.set_variant(variant)
    rows = M["rows"]
    xd = hip.DeviceVector.from_numpy(x)
    yd = hip.DeviceVector.from_numpy(np.full(max(rows, 1), 4711, M["Ax"].dtype) if y0 is None else y0)
    H.spmv(xd, yd, hip.DeviceVector.from_numpy(rhs) if rhs is not None else None, row_start, row_end)
    hip.synchronize()
    return yd.get()[:rows]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("bs", [1, 2, 4, 8, 3])
def test_csr_every_fifth(hip, oracle, rng, dtype, bs):
    """product_TEST.cc:917-957 (sizes 11..352, columns size + 3, blocks 1..8)."""
    for size in (11, 22, 44, 88, 176, 352):
        cols, rowptr, col, val = _every_fifth(size, dtype)
        M = oracle.csr_from_rows(size, cols, rowptr, col, val, bs)
        x = np.full(cols, 4, dtype)
        x[0], x[1] = 1, 2
        ref = oracle.csr_product(M, x, y=np.full(size, 4711, dtype))
        assert np.array_equal(_raw(hip, M, x), ref)
        assert np.array_equal(_handle(hip, M, x), ref)
        rhs = rng.standard_normal(size).astype(dtype)
        xr = rng.standard_normal(cols).astype(dtype)
        refd = oracle.csr_defect(M, rhs, xr)
        assert np.array_equal(_raw(hip, M, xr, rhs), refd)
        assert np.array_equal(_handle(hip, M, xr, rhs), refd)
        lo, hi = 3, size - 2
        y0 = np.full(size, -9, dtype)
        refp = oracle.csr_product(M, xr, lo, hi, y=y0)
        assert np.array_equal(_raw(hip, M, xr, None, lo, hi, y0), refp)
        assert np.array_equal(_handle(hip, M, xr, None, lo, hi, y0), refp)
        assert np.array_equal(_handle(hip, M, xr, rhs, lo, hi, y0), oracle.csr_defect(M, rhs, xr, lo, hi, r=y0))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("name", ["poisson/A_6.ell", "l2/area51_full_0.ell", "poisson/prol_6.ell"])
def test_csr_from_ell_fixtures(hip, oracle, rng, dtype, name):
    """defect_TEST.cc:376-420 (area51 defect through CSR) and the FEATFLOW2
    Poisson / prolongation matrices, blocksizes 1, 2 and 4."""
    F = load_ell(name)
    E = dict(F)
    E["Ax"] = F["Ax"].astype(dtype)
    for bs in (1, 2, 4):
        M = oracle.csr_from_ell(E, bs, cols=F["cols"])
        x = rng.standard_normal(F["cols"]).astype(dtype)
        x[::7] = -0.0
        rhs = rng.standard_normal(F["rows"]).astype(dtype)
        ref = oracle.csr_product(M, x)
        refd = oracle.csr_defect(M, rhs, x)
        assert np.array_equal(_raw(hip, M, x), ref)
        assert np.array_equal(_handle(hip, M, x), ref)
        assert np.array_equal(_raw(hip, M, x, rhs), refd)
        assert np.array_equal(_handle(hip, M, x, rhs), refd)
        # the CSR kernel itself when the handle also holds the ELL form
        assert np.array_equal(_handle(hip, M, x, variant=1), ref)
        assert np.array_equal(_handle(hip, M, x, rhs, variant=1), refd)
        if bs == 1 and name == "poisson/A_6.ell":   # at most 16 entries per row, at most 50% padding
            assert hip.CsrMatrix.from_dict(M).ell_form


@pytest.mark.parametrize("dtype", DTYPES)
def test_csr_ragged_and_long_rows(hip, oracle, rng, dtype):
    """Empty rows, a 20000-entry row (five LDS chunks), rows crossing chunk
    boundaries, a matrix with no entries at all, and the empty matrix."""
    rows, cols = 3000, 50000
    lengths = rng.integers(0, 40, rows)
    lengths[rng.random(rows) < 0.2] = 0
    lengths[1234] = 20000
    lengths[2999] = 5000
    rowptr, col, val = _random_rows(rng, rows, cols, lengths, dtype)
    x = rng.standard_normal(cols).astype(dtype)
    rhs = rng.standard_normal(rows).astype(dtype)
    for bs in (1, 3):
        M = oracle.csr_from_rows(rows, cols, rowptr, col, val, bs)
        ref = oracle.csr_product(M, x)
        assert np.array_equal(_raw(hip, M, x), ref)
        assert np.array_equal(_handle(hip, M, x), ref)
        assert np.array_equal(_handle(hip, M, x, rhs), oracle.csr_defect(M, rhs, x))
    empty = oracle.csr_from_rows(7, 9, np.zeros(8, np.uint64), np.zeros(0, np.uint64), np.zeros(0, dtype), 2)
    x9 = rng.standard_normal(9).astype(dtype)
    r7 = rng.standard_normal(7).astype(dtype)
    assert np.array_equal(_raw(hip, empty, x9), np.zeros(7, dtype))
    assert np.array_equal(_handle(hip, empty, x9, r7), r7)
    none = oracle.csr_from_rows(0, 4, np.zeros(1, np.uint64), np.zeros(0, np.uint64), np.zeros(0, dtype), 1)
    assert _handle(hip, none, rng.standard_normal(4).astype(dtype)).size == 0
