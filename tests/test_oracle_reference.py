"""The oracle against HONEI itself: the C restatement (oracle/oracle.c) and the
reference's own tags::CPU templates compiled from its sources
(oracle/ref/Makefile -> oracle/_ref) must agree bit for bit, including
iteration counts.  Skipped where the reference sources are absent."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import dinv_reference, load_bin, load_ell, load_exp, q1_csr

DTYPES = (np.float64, np.float32)


@pytest.mark.parametrize("dtype", DTYPES)
def test_q1_product_defect_dot_norm(oracle, reference, rng, dtype):
    F = load_bin(4225)
    bands = [b.astype(dtype) for b in F["bands"]]
    x = rng.uniform(-1, 1, F["n"]).astype(dtype)
    rhs = F["b"].astype(dtype)
    assert np.array_equal(oracle.q1_product(bands, x), reference.q1_product(bands, x))
    assert np.array_equal(oracle.q1_defect(bands, rhs, x), reference.q1_defect(bands, rhs, x))
    assert oracle.dot(x, rhs) == reference.dot(x, rhs)
    assert oracle.norm(x) == reference.norm(x)
    assert oracle.norm(x, True) == reference.norm(x, True)


@pytest.mark.parametrize("dtype", DTYPES)
def test_transfers_levels_1_to_10(oracle, reference, rng, dtype):
    for L in range(1, 11):
        nf, nc = P.side_of_level(L) ** 2, P.side_of_level(L - 1) ** 2
        f = rng.uniform(-1, 1, nf).astype(dtype)
        c = rng.uniform(-1, 1, nc).astype(dtype)
        for mask in ((2,) * 8, (2, 1, 2, 1, 1, 2, 1, 2)):
            assert np.array_equal(oracle.restriction(f, mask), reference.restriction(f, mask))
            assert np.array_equal(oracle.prolongation(c, nf, mask), reference.prolongation(c, nf, mask))


@pytest.mark.parametrize("threads", [1, 2, 3, 4, 7, 16])
def test_ell_construction_and_product(oracle, reference, rng, threads):
    F = load_bin(4225)
    rp, col, val = q1_csr(F["bands"], F["n"])
    E1 = oracle.ell_from_csr(F["n"], rp, col, val, threads)
    E2 = reference.ell_from_csr(F["n"], F["n"], rp, col, val, threads)
    assert (E1["stride"], E1["ncpr"]) == (E2["stride"], E2["ncpr"])
    for k in ("Aj", "Ax", "Arl"):
        assert np.array_equal(E1[k], E2[k]), k
    x = rng.uniform(-1, 1, F["n"])
    y2, arl_scan = reference.ell_product(E1, x)
    assert np.array_equal(oracle.ell_product(E1, x), y2)
    # Arl as the raw-array constructor derives it (first zero value)
    assert np.array_equal(oracle.ell_row_length(F["n"], E1["stride"], threads, E1["Ax"]), arl_scan)
    rhs = rng.uniform(-1, 1, F["n"])
    assert np.array_equal(oracle.ell_defect(E1, rhs, x), reference.ell_defect(E1, rhs, x))


def test_cg_none_iterations(oracle, reference):
    n = 129 * 129
    bands = P.poisson_q1_bands(n)
    b = P.rhs_uniform(n)
    x1, it1, _ = oracle.cg_none_q1(bands, b, np.zeros(n))
    x2, it2 = reference.cg_none_q1(bands, b, np.zeros(n))
    assert it1 == it2 and np.array_equal(x1, x2)
    E = P.q1_to_ell(bands)
    x3, it3, _ = oracle.cg_none_ell(E, b, np.zeros(n))
    x4, it4 = reference.cg_none_ell(E, b, np.zeros(n))
    assert it3 == it4 and np.array_equal(x3, x4)


def test_pcg_fixture(oracle, reference):
    F = load_ell("poisson/A_6.ell")
    E = dict(F)
    E["Arl"] = oracle.ell_row_length(F["rows"], F["stride"], 1, F["Ax"])
    n = F["rows"]
    Aj = F["Aj"].reshape(F["ncpr"], F["stride"])[:, :n]
    Ax = F["Ax"].reshape(F["ncpr"], F["stride"])[:, :n]
    diag = np.where((Aj == np.arange(n)) & (Ax != 0), Ax, 0).sum(axis=0)
    Pv = 0.7 / diag
    rhs, init = load_exp("poisson/rhs_6"), load_exp("poisson/init_6")
    x1, it1, _ = oracle.cg_var_ell(E, Pv, rhs, init, 700, 1e-8)
    x2, it2 = reference.cg_var_ell(E, Pv, rhs, init, 700, 1e-8)
    assert it1 == it2 and np.array_equal(x1, x2)


@pytest.mark.parametrize("n", [25, 81, 289])
def test_coarse_cg_konv(oracle, reference, rng, n):
    bands = [b.astype(np.float32) for b in load_bin(n)["bands"]]
    side = int(round(np.sqrt(n)))
    rhs = rng.uniform(-1, 1, n).astype(np.float32)
    rhs[P.boundary_mask(side)] = 0
    eps = float(np.finfo(np.float32).eps)
    x1, _, _ = oracle.cg_konv_q1(bands, rhs, eps)
    assert np.array_equal(x1, reference.cg_konv_q1(bands, rhs, eps))


def test_jacobi(oracle, reference, rng):
    F = load_bin(1089)
    bands = [b.astype(np.float32) for b in F["bands"]]
    dinv = dinv_reference(bands[4])
    x = rng.uniform(-1, 1, F["n"]).astype(np.float32)
    rhs = F["b"].astype(np.float32)
    assert np.array_equal(oracle.jacobi_q1(x, bands, rhs, dinv, 3), reference.jacobi_q1(x, bands, rhs, dinv, 3))


@pytest.mark.parametrize("maxl,minl,tol", [(6, 1, 1e-8), (7, 2, 1e-10), (8, 1, 1e-8)])
def test_mg_mixed_counts_and_solution(oracle, reference, maxl, minl, tol):
    n = P.side_of_level(maxl) ** 2
    A64 = P.poisson_q1_bands(n)
    b = P.rhs_uniform(n)
    a32 = {L: [v.astype(np.float32) for v in P.poisson_q1_bands(P.side_of_level(L) ** 2)]
           for L in range(minl, maxl + 1)}
    d32 = {L: dinv_reference(a32[L][4]) for L in a32}
    r1 = oracle.mg_mixed_q1(A64, b, np.zeros(n), a32, d32, min_level=minl, max_level=maxl, tolerance=tol)
    r2 = reference.mg_mixed_q1(A64, b, np.zeros(n), a32, min_level=minl, max_level=maxl, tolerance=tol)
    assert (r1["outer"], r1["inner"]) == (r2["outer"], r2["inner"])
    assert np.array_equal(r1["x"], r2["x"])


def test_mg_mixed_bin_fixture(oracle, reference):
    F = load_bin(4225)
    a32 = {L: [v.astype(np.float32) for v in load_bin(P.side_of_level(L) ** 2)["bands"]] for L in range(1, 7)}
    d32 = {L: dinv_reference(a32[L][4]) for L in a32}
    r1 = oracle.mg_mixed_q1(F["bands"], F["b"], np.zeros(F["n"]), a32, d32, min_level=1, max_level=6)
    r2 = reference.mg_mixed_q1(F["bands"], F["b"], np.zeros(F["n"]), a32, min_level=1, max_level=6)
    assert (r1["outer"], r1["inner"]) == (r2["outer"], r2["inner"])
    assert np.array_equal(r1["x"], r2["x"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pcg_jac_matches_reference_and_featflow(oracle, reference, dtype):
    """ConjugateGradients<JAC> on ELL: the oracle is the compiled reference bit
    for bit (iterations included) on the area51 fixture of
    conjugate_gradients_TEST.cc:620-700 and on the diffusion matrix, and meets
    FEATFLOW2's solution within that test's eps (3e-5 FP64, 3e-3 FP32)."""
    F = load_ell("l2/area51_full_0.ell")
    E = dict(F)
    E["Arl"] = oracle.ell_row_length(F["rows"], F["stride"], 1, F["Ax"])
    E["Ax"] = F["Ax"].astype(dtype)
    n = F["rows"]
    Aj = F["Aj"].reshape(F["ncpr"], F["stride"])[:, :n]
    Ax = F["Ax"].reshape(F["ncpr"], F["stride"])[:, :n]
    diag = np.where((Aj == np.arange(n)) & (Ax != 0), Ax, 0).sum(axis=0).astype(dtype)
    dd = (dtype(1) / diag).astype(dtype)
    rhs = load_exp("l2/area51_rhs_0").astype(dtype)
    init = load_exp("l2/area51_init_0").astype(dtype)
    sol = load_exp("l2/area51_sol_0")
    eps = dtype(1e-8)
    x1, it1, st = oracle.pcg_jac_ell(E, dd, rhs, init, 10000, eps)
    x2, it2 = reference.pcg_jac_ell(E, dd, rhs, init, 10000, eps)
    assert it1 == it2 and np.array_equal(x1, x2)
    assert np.abs(x1.astype(np.float64) - sol).max() <= (3e-5 if dtype == np.float64 else 3e-3)
    D = P.diffusion_ell(5)
    D["Ax"] = D["Ax"].astype(dtype)
    ddd = (dtype(1) / np.where(D["Aj"].reshape(D["ncpr"], D["stride"]) == np.arange(D["stride"]),
                               D["Ax"].reshape(D["ncpr"], D["stride"]), 0).sum(axis=0)[:D["rows"]]).astype(dtype)
    b = P.rhs_diffusion(5, 0).astype(dtype)
    y1, j1, _ = oracle.pcg_jac_ell(D, ddd, b, np.zeros(D["rows"], dtype), 10000, dtype(1e-6))
    y2, j2 = reference.pcg_jac_ell(D, ddd, b, np.zeros(D["rows"], dtype), 10000, dtype(1e-6))
    assert j1 == j2 and np.array_equal(y1, y2)


def _every_fifth(size, dtype):
    """product_TEST.cc:917-957: size x (size + 3), element index i of the
    row-major matrix stored when i % 5 == 0, with value i / 1.234."""
    cols = size + 3
    idx = np.arange(size * cols)
    sel = idx[idx % 5 == 0]
    r, c = sel // cols, sel % cols
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=size))]).astype(np.uint64)
    return cols, rowptr, c.astype(np.uint64), (sel / 1.234).astype(dtype), r


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bs", [1, 2, 3, 4, 8])
def test_csr_matches_reference(oracle, reference, rng, dtype, bs):
    """SparseMatrixCSR construction (from rows and from ELL), Product and
    Defect on row ranges: the oracle is the compiled reference bit for bit,
    for csr::blocksize 1..8 (the blocks of product_TEST.cc:929) and the
    matrices of product_TEST.cc:917-957 and defect_TEST.cc:376-420."""
    for size in (11, 22, 44, 88, 176, 352):
        cols, rowptr, col, val, r = _every_fifth(size, dtype)
        A = oracle.csr_from_rows(size, cols, rowptr, col, val, bs)
        B = reference.csr_from_rows(size, cols, rowptr, col, val, bs)
        for k in ("Ar", "Aj", "Ax", "nblocks", "used"):
            assert np.array_equal(A[k], B[k]), k
        assert A["used"] == col.size
        x = np.full(cols, 4, dtype)
        x[0], x[1] = 1, 2
        y = oracle.csr_product(A, x, y=np.full(size, 4711, dtype))
        assert np.array_equal(y, reference.csr_product(B, x, y=np.full(size, 4711, dtype)))
        dense = np.zeros((size, cols))
        dense[r, col.astype(np.int64)] = val
        assert np.allclose(y, dense @ x.astype(np.float64), rtol=1e-5 if dtype == np.float32 else 1e-13)
        lo, hi = 2, size - 3
        rhs = rng.standard_normal(size).astype(dtype)
        xr = rng.standard_normal(cols).astype(dtype)
        y0 = np.full(size, 7, dtype)
        assert np.array_equal(oracle.csr_product(A, xr, lo, hi, y=y0), reference.csr_product(B, xr, lo, hi, y=y0))
        d = oracle.csr_defect(A, rhs, xr, lo, hi, r=y0)
        assert np.array_equal(d, reference.csr_defect(B, rhs, xr, lo, hi, r=y0))
        assert np.array_equal(d[:lo], y0[:lo]) and np.array_equal(d[hi:], y0[hi:])
    for name in ("poisson/A_6.ell", "l2/area51_full_0.ell", "poisson/prol_6.ell"):
        F = load_ell(name)
        E = dict(F)
        E["Ax"] = F["Ax"].astype(dtype)
        A = oracle.csr_from_ell(E, bs, cols=F["cols"])
        B = reference.csr_from_ell(E, bs, cols=F["cols"])
        for k in ("Ar", "Aj", "Ax", "nblocks", "used"):
            assert np.array_equal(A[k], B[k]), (name, k)
        x = rng.standard_normal(F["cols"]).astype(dtype)
        rhs = rng.standard_normal(F["rows"]).astype(dtype)
        y = oracle.csr_product(A, x)
        assert np.array_equal(y, reference.csr_product(B, x))
        assert np.array_equal(oracle.csr_defect(A, rhs, x), reference.csr_defect(B, rhs, x))
        if bs == 1:
            E["Arl"] = oracle.ell_row_length(F["rows"], F["stride"], 1, E["Ax"])
            assert np.array_equal(y, oracle.ell_product(E, x))
