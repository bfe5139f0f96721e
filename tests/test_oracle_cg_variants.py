"""The oracle's fixed-iteration ConjugateGradients<NONE> forms, CGSolver<VAR>
with a SparseMatrixELL preconditioner and the operator-list multigrid with
matrix preconditioners, pinned against the compiled reference (bit for bit),
HONEI's fixtures and what the mathematics fixes:

* ConjugateGradients<NONE>(Q1, rhs, x, unsigned long iters)
  (conjugate_gradients.hh:526-569): the case of
  poisson_cg_double_banded_q1_TEST.cc:50-72 (FillMatrix/FillVector
  DIRICHLET_NEUMANN on 5x5, 20 iterations) and the 4225.bin system; one
  step from x = 0 is the steepest-descent step (r.r / r.Ar) r.
* ConjugateGradients<NONE>(ELL, rhs, x, unsigned long max_iters)
  (:571-586, kernel :203-248): conjugate_gradients_TEST.cc:520-600 on
  area51 against FEATFLOW2's solution; in FP64 on a system where neither
  epsilon guard fires it equals CGSolver<NONE> with eps 1e-8 bit for bit.
* CGSolver<VAR> with a matrix P (cg.hh:141-219): the SPAI fixtures
  5pt_10x10_spai.ell and poisson_advanced/sort_0/A_4_spai.ell; a diagonal
  P as a matrix gives the bits of the same P as a DenseVector.
* MGSolver with PreconFill<SparseMatrixELL> ("spai", "sainv") smoothers and
  VAR coarse solves (mg.hh:315-335)."""
import os

import numpy as np
import pytest

from honei_fixtures import TESTDATA, ell_from_file, load_bin, load_ell, load_exp, load_mg_data

DTYPES = (np.float64, np.float32)
ADV = os.path.join(TESTDATA, "poisson_advanced", "sort_0")


def _ell(path, dtype, oracle):
    return ell_from_file(os.path.join(TESTDATA, path), dtype, oracle)


def _featflow_eps(dtype):
    """conjugate_gradients_TEST.cc:585-596: 3 * (m * sizeof(DT) + b)."""
    base_eps = 1 / 10 ** 3
    add_eps = base_eps / 10 ** 2
    m = (add_eps - base_eps) / 4
    b = base_eps - 4 * m
    return 3 * (m * np.dtype(dtype).itemsize + b)


# ---------------------------------------------------------------- Q1, iters
def test_cg_iters_q1_dirichlet_neumann_25(oracle, reference):
    n = 25
    bands = oracle.fill_poisson_dirichlet_neumann(n)
    b = oracle.fill_vector_dirichlet_neumann(n)
    x, used, st = oracle.cg_iters_q1(bands, b, 20)
    assert np.array_equal(x, reference.cg_iters_q1(bands, b, 20))
    assert 1 <= used <= 20
    if st == 0:   # the test fired: squared true defect below 1e-8 of the initial one
        r = oracle.q1_defect(bands, b, x)
        assert np.dot(r, r) < 1e-8 * np.dot(b, b)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("iters", [1, 3, 40, 2000])
def test_cg_iters_q1_matches_reference(oracle, reference, dtype, iters):
    F = load_bin(4225)
    bands = [v.astype(dtype) for v in F["bands"]]
    b = F["b"].astype(dtype)
    x, used, st = oracle.cg_iters_q1(bands, b, iters)
    assert np.array_equal(x, reference.cg_iters_q1(bands, b, iters))
    assert used <= iters
    if st != 0:
        assert used == iters


def test_cg_iters_q1_first_step_is_steepest_descent(oracle):
    """x = 0 and u = rhs, so one step gives x1 = (r.r / r.A r) r, r = rhs."""
    F = load_bin(1089)
    bands = F["bands"]
    b = F["b"]
    x, used, _ = oracle.cg_iters_q1(bands, b, 1)
    Ab = oracle.q1_product(bands, b)
    x1 = (b @ b) / (b @ Ab) * b
    assert used == 1
    np.testing.assert_allclose(x, x1, rtol=1e-12, atol=1e-15)


# --------------------------------------------------------------- ELL, iters
@pytest.mark.parametrize("dtype", DTYPES)
def test_cg_iters_ell_area51(oracle, reference, dtype):
    E = _ell("l2/area51_full_0.ell", dtype, oracle)
    rhs = load_exp("l2/area51_rhs_0").astype(dtype)
    init = load_exp("l2/area51_init_0").astype(dtype)
    sol = load_exp("l2/area51_sol_0")
    x, used, st = oracle.cg_iters_ell(E, rhs, init, 10000)
    assert np.array_equal(x, reference.cg_iters_ell(E, rhs, init, 10000))
    assert st == 0 and used < 10000
    assert np.max(np.abs(x.astype(np.float64) - sol)) <= _featflow_eps(dtype)


def test_cg_iters_ell_equals_cgsolver_without_guards(oracle, reference):
    """FP64: while |v.p| and alpha_old = ||r||^2 stay above DBL_EPSILON the
    fixed-iteration form and CGSolver<NONE>(eps 1e-8) take the same steps;
    near convergence ||r||^2 falls below DBL_EPSILON on this system, where
    CGSolver's guard replaces alpha_old and the two forms part."""
    E = _ell("poisson/A_6.ell", np.float64, oracle)
    rhs, init = load_exp("poisson/rhs_6"), load_exp("poisson/init_6")
    x1, u1, _ = oracle.cg_iters_ell(E, rhs, init, 100)
    x2, u2, _ = oracle.cg_none_ell(E, rhs, init, 100, 1e-8)
    assert u1 == u2 == 100 and np.array_equal(x1, x2)
    x1, u1, st1 = oracle.cg_iters_ell(E, rhs, init, 5000)
    x2, u2, _ = oracle.cg_none_ell(E, rhs, init, 5000, 1e-8)
    assert st1 == 0 and not np.array_equal(x1, x2)
    assert np.array_equal(x1, reference.cg_iters_ell(E, rhs, init, 5000))
    # a cap below the converged count stops there
    x3, u3, st3 = oracle.cg_iters_ell(E, rhs, init, 7)
    assert (u3, st3) == (7, 1)
    assert np.array_equal(x3, reference.cg_iters_ell(E, rhs, init, 7))


# ------------------------------------------------ CGSolver<VAR>, matrix P
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", ["5pt", "sort0_A4"])
def test_cg_var_matrix_preconditioner(oracle, reference, dtype, case):
    if case == "5pt":
        A = _ell("5pt_10x10.ell", dtype, oracle)
        Pm = _ell("5pt_10x10_spai.ell", dtype, oracle)
    else:
        A = _ell("poisson_advanced/sort_0/A_4.ell", dtype, oracle)
        Pm = _ell("poisson_advanced/sort_0/A_4_spai.ell", dtype, oracle)
    n = A["rows"]
    b = np.random.default_rng(3).uniform(-1, 1, n).astype(dtype)
    x0 = np.zeros(n, dtype)
    x, used, st = oracle.cg_var_ell_pm(A, Pm, b, x0, 500, dtype(1e-8))
    xr, ur = reference.cg_var_ell_pm(A, Pm, b, x0, 500, dtype(1e-8))
    assert used == ur and np.array_equal(x, xr)
    if dtype == np.float64:
        assert st == 0
        r = oracle.ell_defect(A, b, x)
        assert np.linalg.norm(r) < 1e-7 * np.linalg.norm(b)


def test_cg_var_diagonal_matrix_equals_vector(oracle):
    A = _ell("poisson/A_6.ell", np.float64, oracle)
    n = A["rows"]
    Aj = A["Aj"].reshape(A["ncpr"], A["stride"])[:, :n]
    Ax = A["Ax"].reshape(A["ncpr"], A["stride"])[:, :n]
    diag = np.where((Aj == np.arange(n)) & (Ax != 0), Ax, 0).sum(axis=0)
    Pv = 0.7 / diag
    stride = A["stride"]
    D = dict(rows=n, cols=n, stride=stride, ncpr=1, threads=1, Aj=np.zeros(stride, np.uint64),
             Ax=np.zeros(stride), Arl=np.ones(n, np.uint64))
    D["Aj"][:n] = np.arange(n, dtype=np.uint64)
    D["Ax"][:n] = Pv
    rhs, init = load_exp("poisson/rhs_6"), load_exp("poisson/init_6")
    x1, u1, _ = oracle.cg_var_ell(A, Pv, rhs, init, 700, 1e-8)
    x2, u2, _ = oracle.cg_var_ell_pm(A, D, rhs, init, 700, 1e-8)
    assert u1 == u2 and np.array_equal(x1, x2)


# ------------------------------------- operator-list MG, matrix preconditioners
def load_mg_data_pm(base, levels, dtype, oracle, suffix):
    d = load_mg_data(base, levels, dtype, oracle)
    d["Pm"] = [ell_from_file(os.path.join(base, f"A_{i + 1}_{suffix}.ell"), dtype, oracle) for i in range(levels)]
    return d


@pytest.mark.parametrize("suffix", ["spai", "sainv"])
@pytest.mark.parametrize("cycle", ["V", "W", "F"])
@pytest.mark.parametrize("coarse_var", [False, True])
def test_mgop_matrix_preconditioner_matches_reference(oracle, reference, suffix, cycle, coarse_var):
    d = load_mg_data_pm(ADV, 4, np.float64, oracle, suffix)
    kw = dict(cycle=cycle, coarse_var=coarse_var, max_iters=100, max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8)
    mine = oracle.mgop_solve_pm(d, d["b"], d["x"], **kw)
    ref = reference.mgop_solve_pm(ADV, d["b"].size, suffix, levels=4, **kw)
    assert (mine["used"], mine["used_coarse"]) == (ref["used"], ref["used_coarse"])
    # (the F cycle with a CGSolver<VAR> coarse solve diverges to NaN in the
    # reference too: same NaN pattern)
    assert np.array_equal(mine["x"], ref["x"], equal_nan=True)
    if cycle != "F" or not coarse_var:
        # the cycle converges to the fixture's solution
        sol = load_exp(os.path.join(ADV, "sol_4"))
        assert mine["used"] < 100
        assert np.max(np.abs(mine["x"] - sol)) < 1e-4
