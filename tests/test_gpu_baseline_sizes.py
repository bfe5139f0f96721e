"""The benchmark workloads at their full sizes, against the oracle.

* FillMatrix<POISSON, DIRICHLET_0> on the device (the kernel that assembles
  every level of the multigrid hierarchies): bit-identical to the oracle's
  restatement of fill_matrix.hh:37-76, FP64 and FP32, up to 4097^2.
* MG-IR (FP32 V-cycles on levels 2..12, FP64 refinement) on Q1 Poisson
  4097^2 to relative residual 1e-10: outer count within +-1 of the oracle's
  Multigrid<MIXED> restatement, inner count within +-2 (two V-cycles per
  outer step, so one outer step more is two cycles more), solution within
  1e-6 relative, final defect below the tolerance.
* The 64-right-hand-side MG-IR batch on the 2049^2 variable-coefficient
  diffusion hierarchy (Neumann border mask): columns 0 and 63 against the
  oracle's Multigrid<MIXED> on that hierarchy (same bounds), and columns 0,
  31 and 63 equal to their own single solves bit for bit.
* Matrix B (2^24 rows, 5..13 entries per row within +-2048 of the
  diagonal), FP64 and FP32: sampled rows of the product and the defect
  bit-identical to HONEI's row sums."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import dinv_reference, ell_diag

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [2, 3, 5, 17, 129, 4097])
def test_fill_poisson_device_equals_fill_matrix(hip, oracle, side, dtype):
    n = side * side
    ref = oracle.fill_poisson_dirichlet0(n, dtype)
    A = hip.Q1Matrix([np.full(n, 7.0, dtype) for _ in range(9)])
    hip.fill_poisson_q1(A)
    for k in range(9):
        got = A.bands[k].get()
        assert got.dtype == dtype
        assert np.array_equal(got.view(np.uint8), np.asarray(ref[k], dtype).view(np.uint8)), (side, k)


def _q1_levels(maxl, minl):
    a32, d32 = {}, {}
    for L in range(minl, maxl + 1):
        n = P.side_of_level(L) ** 2
        a32[L] = [b.astype(np.float32) for b in P.poisson_q1_bands(n)]
        d32[L] = dinv_reference(a32[L][4])
    return a32, d32


def test_mg_ir_poisson_4097_levels_2_12(hip, oracle):
    maxl, minl, tol = 12, 2, 1e-10
    n = P.side_of_level(maxl) ** 2
    b = P.rhs_uniform(n)
    mg = hip.MGSolver(hip.mg_params(minl, maxl, tolerance=tol))      # device-assembled hierarchy
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    info = mg.solve(bd, xd)
    x = xd.get()
    del mg, bd, xd
    a32, d32 = _q1_levels(maxl, minl)
    ref = oracle.mg_mixed_q1(P.poisson_q1_bands(n), b, np.zeros(n), a32, d32, min_level=minl, max_level=maxl,
                             tolerance=tol)
    assert abs(info["outer"] - ref["outer"]) <= 1 and abs(info["inner"] - ref["inner"]) <= 2, (info, ref["outer"],
                                                                                                ref["inner"])
    assert _rel(x, ref["x"]) <= 1e-6
    assert info["final_defect"] < float(np.float32(tol)) * 1.0000001 * info["initial_defect"]


def test_mg_ir_diffusion_2049_64rhs(hip, oracle):
    maxl, minl, R = 11, 2, 64
    n = P.side_of_level(maxl) ** 2
    A64 = P.diffusion_ell(maxl)
    a32 = {L: P.diffusion_ell(L, dtype=np.float32) for L in range(minl, maxl + 1)}
    d32 = {L: dinv_reference(ell_diag(a32[L])) for L in a32}
    prm = hip.mg_params(minl, maxl, tolerance=1e-8, mask=P.DIFFUSION_MASK)
    mg = hip.MGSolver.from_ell(prm, A64, a32, d32, slots=R)
    cols = (0, 31, R - 1)
    bs = {j: P.rhs_diffusion(maxl, j) for j in range(R)}
    bd = [hip.DeviceVector.from_numpy(bs[j]) for j in range(R)]
    xd = [hip.DeviceVector.zeros(n) for _ in range(R)]
    infos = mg.solve_batch(bd, xd)
    xs = {j: xd[j].get() for j in cols}
    for j in range(R):
        assert infos[j]["final_defect"] < 1e-8 * infos[j]["initial_defect"], (j, infos[j])
    for j in cols:
        x1 = hip.DeviceVector.zeros(n)
        single = mg.solve(bd[j], x1)
        assert np.array_equal(x1.get(), xs[j]), j
        assert (single["outer"], single["inner"]) == (infos[j]["outer"], infos[j]["inner"])
    del mg, bd, xd
    for j in (0, R - 1):
        ref = oracle.mg_mixed(kind="ell", b=bs[j], x0=np.zeros(n), A64=A64, a32=a32, dinv32=d32, min_level=minl,
                              max_level=maxl, n_max_iter=2, tolerance=1e-8, mask=P.DIFFUSION_MASK)
        assert abs(infos[j]["outer"] - ref["outer"]) <= 1 and abs(infos[j]["inner"] - ref["inner"]) <= 2, \
            (j, infos[j], ref["outer"], ref["inner"])
        assert _rel(xs[j], ref["x"]) <= 1e-6, j


@pytest.fixture(scope="module")
def matrix_b_2e24():
    return P.matrix_b_ell(1 << 24)


def _host_rows(E, x, rows):
    stride, ncpr = E["stride"], E["ncpr"]
    Aj = E["Aj"].reshape(ncpr, stride)
    Ax = E["Ax"].reshape(ncpr, stride)
    ln = np.asarray(E["Arl"], np.int64)[rows]
    s = np.zeros(rows.size, E["Ax"].dtype)
    for k in range(ncpr):
        m = k < ln
        s[m] = s[m] + Ax[k, rows[m]] * x[Aj[k, rows[m]].astype(np.int64)]
    return s


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matrix_b_2e24_sampled_rows(hip, matrix_b_2e24, dtype):
    E = dict(matrix_b_2e24)
    E["Ax"] = E["Ax"].astype(dtype)
    n = E["rows"]
    A = hip.EllMatrix.from_dict(E)
    x = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    rhs = P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype)
    xd, rd = hip.DeviceVector.from_numpy(x), hip.DeviceVector.from_numpy(rhs)
    y = hip.DeviceVector.zeros(n, dtype)
    rng = np.random.default_rng(24)
    rows = np.unique(np.r_[np.arange(5000), np.arange(n - 5000, n), rng.integers(0, n, 40000)])
    ref = _host_rows(E, x, rows)
    A.spmv(xd, y)
    assert np.array_equal(y.get()[rows], ref)
    A.spmv(xd, y, rhs=rd)
    d_all = y.get()
    assert np.array_equal(d_all[rows], rhs[rows] - ref)
    # the persistent LDS-window kernel (sorted copy, XCD order): every row the same
    A.spmv(xd, y)
    y_all = y.get()
    A.set_variant(116)
    A.spmv(xd, y)
    assert np.array_equal(y.get(), y_all)
    A.spmv(xd, y, rhs=rd)
    assert np.array_equal(y.get(), d_all)
