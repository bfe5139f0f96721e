"""GPU parity of the remaining solver entry points of the boundary against
the oracle:

* ConjugateGradients<NONE>(Q1, rhs, x, unsigned long iters)
  (hsolve_cg_iters_q1_*): one workgroup with sequential sums, bit-exact;
* ConjugateGradients<NONE>(ELL, rhs, x, unsigned long max_iters)
  (hsolve_cg_iters_ell): iteration count within +-1, solution within 1e-6
  (FP64) / 1e-4 (FP32) relative, FEATFLOW2's area51 solution within the
  reference test's bound;
* CGSolver<VAR> with a SparseMatrixELL preconditioner (hsolve_pcg_ell_pmat)
  on the SPAI fixtures, same bounds;
* the multi-block loops of the solvers (grids above 512 blocks): CG, PCG and
  ConjugateGradients<JAC> at 513^2 (263,169 rows), same bounds."""
import os

import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import TESTDATA, ell_from_file, load_bin, load_exp

pytestmark = pytest.mark.gpu

DTYPES = (np.float64, np.float32)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _tol(dtype):
    return 1e-6 if dtype == np.float64 else 1e-4


def _ell(path, dtype, oracle):
    return ell_from_file(os.path.join(TESTDATA, path), dtype, oracle)


@pytest.mark.parametrize("dtype", DTYPES)
def test_cg_iters_q1_bit_exact(hip, oracle, dtype):
    # poisson_cg_double_banded_q1_TEST.cc:50-72: DIRICHLET_NEUMANN 5 x 5, 20 steps
    bands = [v.astype(dtype) for v in oracle.fill_poisson_dirichlet_neumann(25)]
    b = oracle.fill_vector_dirichlet_neumann(25).astype(dtype)
    cases = [(bands, b, 20)]
    F = load_bin(4225)
    b4 = F["b"].astype(dtype)
    bands4 = [v.astype(dtype) for v in F["bands"]]
    cases += [(bands4, b4, it) for it in (1, 3, 50, 3000)]
    for bb, rhs, iters in cases:
        x_ref, used_ref, _ = oracle.cg_iters_q1(bb, rhs, iters)
        A = hip.Q1Matrix(bb)
        xd = hip.DeviceVector.from_numpy(np.full(rhs.size, 5, dtype))      # x is set to 0 first
        used = hip.cg_iters_q1(A, hip.DeviceVector.from_numpy(rhs), xd, iters)
        assert used == used_ref, (rhs.size, iters, used, used_ref)
        assert np.array_equal(xd.get(), x_ref), (rhs.size, iters)


@pytest.mark.parametrize("dtype", DTYPES)
def test_cg_iters_ell_area51(hip, oracle, dtype):
    """conjugate_gradients_TEST.cc:520-600: value(A, rhs, x, 10000ul) from
    init against FEATFLOW2's solution within 3 (m sizeof(DT) + b)."""
    E = _ell("l2/area51_full_0.ell", dtype, oracle)
    rhs = load_exp("l2/area51_rhs_0").astype(dtype)
    init = load_exp("l2/area51_init_0").astype(dtype)
    sol = load_exp("l2/area51_sol_0")
    x_ref, it_ref, _ = oracle.cg_iters_ell(E, rhs, init, 10000)
    M = hip.EllMatrix.from_dict(E)
    xd = hip.DeviceVector.from_numpy(init)
    used, _ = hip.cg_iters_ell(M, hip.DeviceVector.from_numpy(rhs), xd, 10000)
    x = xd.get()
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(x, x_ref) <= _tol(dtype)
    m = (1e-5 - 1e-3) / 4
    eps = 3 * (m * np.dtype(dtype).itemsize + (1e-3 - 4 * m))
    assert np.max(np.abs(x.astype(np.float64) - sol)) <= eps
    # a cap below convergence: exactly that many steps, no error
    xd = hip.DeviceVector.from_numpy(init)
    used, _ = hip.cg_iters_ell(M, hip.DeviceVector.from_numpy(rhs), xd, 5)
    x5, it5, st5 = oracle.cg_iters_ell(E, rhs, init, 5)
    assert (used, it5, st5) == (5, 5, 1)
    assert _rel(xd.get(), x5) <= _tol(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", ["5pt", "sort0_A4"])
def test_pcg_matrix_preconditioner(hip, oracle, dtype, case):
    if case == "5pt":
        A = _ell("5pt_10x10.ell", dtype, oracle)
        Pm = _ell("5pt_10x10_spai.ell", dtype, oracle)
    else:
        A = _ell("poisson_advanced/sort_0/A_4.ell", dtype, oracle)
        Pm = _ell("poisson_advanced/sort_0/A_4_spai.ell", dtype, oracle)
    n = A["rows"]
    b = np.random.default_rng(3).uniform(-1, 1, n).astype(dtype)
    x0 = np.random.default_rng(4).uniform(-0.1, 0.1, n).astype(dtype)
    x_ref, it_ref, _ = oracle.cg_var_ell_pm(A, Pm, b, x0, 500, dtype(1e-8))
    MA, MP = hip.EllMatrix.from_dict(A), hip.EllMatrix.from_dict(Pm)
    xd = hip.DeviceVector.from_numpy(x0)
    used, _ = hip.pcg_ell_pmat(MA, MP, hip.DeviceVector.from_numpy(b), xd, 500, 1e-8, raise_on_cap=False)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(xd.get(), x_ref) <= _tol(dtype)


def test_solvers_above_512_blocks(hip, oracle):
    """513^2 rows: CGSolver<NONE> and <VAR> on Poisson, the fixed-iteration
    ConjugateGradients<NONE> and ConjugateGradients<JAC> on the level-9
    variable-coefficient diffusion matrix run their multi-block loops."""
    n = 513 * 513
    bands = P.poisson_q1_bands(n)
    E = P.q1_to_ell(bands)
    b = P.rhs_uniform(n)
    M = hip.EllMatrix.from_dict(E)
    bd = hip.DeviceVector.from_numpy(b)
    x_ref, it_ref, _ = oracle.cg_none_ell(E, b, np.zeros(n), 20000, 1e-8)
    xd = hip.DeviceVector.zeros(n)
    used, _ = hip.cg_ell(M, bd, xd, 20000, 1e-8)
    assert abs(used - it_ref) <= 1 and _rel(xd.get(), x_ref) <= 1e-6, (used, it_ref)
    x_ref, it_ref, _ = oracle.cg_iters_ell(E, b, np.zeros(n), 20000)
    xd = hip.DeviceVector.zeros(n)
    used, _ = hip.cg_iters_ell(M, bd, xd, 20000)
    assert abs(used - it_ref) <= 1 and _rel(xd.get(), x_ref) <= 1e-6, (used, it_ref)
    Pv = 0.7 / bands[4]
    x_ref, it_ref, _ = oracle.cg_var_ell(E, Pv, b, np.zeros(n), 20000, 1e-8)
    xd = hip.DeviceVector.zeros(n)
    used, _ = hip.pcg_ell(M, hip.DeviceVector.from_numpy(Pv), bd, xd, 20000, 1e-8)
    assert abs(used - it_ref) <= 1 and _rel(xd.get(), x_ref) <= 1e-6, (used, it_ref)
    D = P.diffusion_ell(9)
    m = D["rows"]
    Aj = D["Aj"].reshape(D["ncpr"], D["stride"])[:, :m]
    Ax = D["Ax"].reshape(D["ncpr"], D["stride"])[:, :m]
    dd = 1.0 / np.where((Aj == np.arange(m)) & (Ax != 0), Ax, 0).sum(axis=0)
    bdiff = P.rhs_diffusion(9, 3)
    x_ref, it_ref, _ = oracle.pcg_jac_ell(D, dd, bdiff, np.zeros(m), 40000, 1e-8)
    MD = hip.EllMatrix.from_dict(D)
    xd = hip.DeviceVector.zeros(m)
    used, _ = hip.pcg_jac_ell(MD, hip.DeviceVector.from_numpy(dd), hip.DeviceVector.from_numpy(bdiff), xd, 40000,
                              1e-8)
    assert abs(used - it_ref) <= 1 and _rel(xd.get(), x_ref) <= 1e-6, (used, it_ref)
