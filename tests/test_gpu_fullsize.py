"""The ELL SpMV / defect at the benchmark's full size: Matrix A, the Q1
9-point stencil on 4097 x 4097 nodes (16.8M rows, 151M nonzeros, the
9-point structured form with implied columns), and Matrix B (random
columns within +-2048, the LDS-window form) at 2^22 rows.  Rows sampled
across the matrix (both ends included) are recomputed on the host in
HONEI's order (slots ascending, summed from zero) and must match bit for
bit; the whole product must also be linear (A(x1 + x2) = A x1 + A x2 to
rounding)."""
import numpy as np
import pytest

from honei_amd import problems as P

pytestmark = pytest.mark.gpu


def _host_rows(E, x, rows):
    """Row sums of y = A x for the given rows: slot k ascending from zero;
    slots at or beyond a row's length add nothing."""
    stride, ncpr = E["stride"], E["ncpr"]
    Aj = E["Aj"].reshape(ncpr, stride)
    Ax = E["Ax"].reshape(ncpr, stride)
    ln = np.asarray(E["Arl"], np.int64)[rows]
    s = np.zeros(rows.size, E["Ax"].dtype)
    for k in range(ncpr):
        m = k < ln
        s[m] = s[m] + Ax[k, rows[m]] * x[Aj[k, rows[m]].astype(np.int64)]
    return s


def _sample(n, rng, count=20000):
    edge = np.r_[np.arange(5000), np.arange(n - 5000, n)]
    return np.unique(np.r_[edge, rng.integers(0, n, count)])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matrix_a_4097_sampled_rows_and_linearity(hip, dtype):
    E = P.q1_to_ell(P.matrix_a_bands(4097, dtype=dtype))
    n = E["rows"]
    A = hip.EllMatrix.from_dict(E)
    assert A.q1_side == 4097
    rng = np.random.default_rng(11)
    x1 = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    x2 = P.rng_uniform(P.SEED + 1, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    rhs = P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    d1, d2, dr = (hip.DeviceVector.from_numpy(v) for v in (x1, x2, rhs))
    y = hip.DeviceVector.zeros(n, dtype)
    A.spmv(d1, y)
    y1 = y.get()
    rows = _sample(n, rng)
    assert np.array_equal(y1[rows], _host_rows(E, x1, rows))
    A.spmv(d1, y, rhs=dr)
    assert np.array_equal(y.get()[rows], rhs[rows] - _host_rows(E, x1, rows))
    A.spmv(d2, y)
    y2 = y.get()
    A.spmv(hip.DeviceVector.from_numpy(x1 + x2), y)
    y12 = y.get().astype(np.float64)
    scale = np.abs(E["Ax"]).max() * 2 * 9
    tol = (1e-14 if dtype == np.float64 else 1e-5) * scale
    assert np.max(np.abs(y12 - (y1.astype(np.float64) + y2))) <= tol


def test_matrix_b_lds_window_sampled_rows(hip):
    E = P.matrix_b_ell(1 << 22)
    n = E["rows"]
    A = hip.EllMatrix.from_dict(E)
    x = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1)
    y = hip.DeviceVector.zeros(n)
    A.spmv(hip.DeviceVector.from_numpy(x), y)
    rows = _sample(n, np.random.default_rng(5))
    got = y.get()
    assert np.array_equal(got[rows], _host_rows(E, x, rows))
    A.set_variant(4)                      # scattered gathers, no LDS window
    A.spmv(hip.DeviceVector.from_numpy(x), y)
    assert np.array_equal(y.get(), got)


@pytest.mark.parametrize("which", ["a_f64", "a_f32", "b_f64"])
def test_csr_equals_ell_full_size(hip, which):
    """The same matrix as SparseMatrixCSR (blocksize 1): its rows hold the
    ELL slots in the same order, so Product and Defect must equal the ELL
    results bit for bit on every row (Matrix A 4097^2, Matrix B 2^22 rows);
    rows sampled across the matrix are also recomputed on the host."""
    if which == "b_f64":
        E = P.matrix_b_ell(1 << 22)
    else:
        E = P.q1_to_ell(P.matrix_a_bands(4097, dtype=np.float64 if which == "a_f64" else np.float32))
    n, dtype = E["rows"], E["Ax"].dtype
    A = hip.EllMatrix.from_dict(E)
    C = hip.CsrMatrix.from_dict(P.ell_to_csr(E))
    x = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    rhs = P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    dx, dr = hip.DeviceVector.from_numpy(x), hip.DeviceVector.from_numpy(rhs)
    ye, yc = hip.DeviceVector.zeros(n, dtype), hip.DeviceVector.zeros(n, dtype)
    assert C.ell_form
    A.spmv(dx, ye)
    ref = ye.get()
    for variant in (0, 1):                 # the handle's ELL form, then the CSR kernel itself
        C.set_variant(variant)
        yc.fill(0.0)
        C.spmv(dx, yc)
        got = yc.get()
        assert np.array_equal(got, ref), variant
        rows = _sample(n, np.random.default_rng(3))
        assert np.array_equal(got[rows], _host_rows(E, x, rows))
    A.spmv(dx, ye, rhs=dr)
    for variant in (0, 1):
        C.set_variant(variant)
        C.spmv(dx, yc, rhs=dr)
        assert np.array_equal(yc.get(), ye.get()), variant
