"""The device solvers above 131,072 rows, where they leave the one-workgroup
small-grid form for the multi-block loop (a reduction kernel, a finish
kernel and the update kernels per step): plain CG and CGSolver<VAR> on the
513^2 Poisson ELL matrix and ConjugateGradients<JAC> on the 513^2
variable-coefficient diffusion matrix, each against the oracle.

Tolerances (FP64): iteration counts within +-1 of the oracle, solutions
within 1e-6 relative, final defect below the tolerance."""
import numpy as np
import pytest

from honei_amd import problems as P

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _diag(E):
    n = E["rows"]
    Aj = E["Aj"].reshape(E["ncpr"], E["stride"])[:, :n]
    Ax = E["Ax"].reshape(E["ncpr"], E["stride"])[:, :n]
    return np.where((Aj == np.arange(n)) & (Ax != 0), Ax, 0).sum(axis=0)


@pytest.fixture(scope="module")
def poisson_513():
    n = 513 * 513
    return P.q1_to_ell(P.poisson_q1_bands(n)), P.rhs_uniform(n)


def test_cg_ell_poisson_513(hip, oracle, poisson_513):
    E, b = poisson_513
    n = E["rows"]
    assert n > 131072
    x_ref, it_ref, st = oracle.cg_none_ell(E, b, np.zeros(n), 20000, 1e-8)
    assert st == 0
    M = hip.EllMatrix.from_dict(E)
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    used, rel_def = hip.cg_ell(M, bd, xd, 20000, 1e-8)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert rel_def < 1e-8
    assert _rel(xd.get(), x_ref) <= 1e-6


def test_pcg_ell_var_poisson_513(hip, oracle, poisson_513):
    """CGSolver<VAR> with the Jacobi(0.7) vector preconditioner (cg.hh:141-219)."""
    E, b = poisson_513
    n = E["rows"]
    Pv = 0.7 / _diag(E)
    x_ref, it_ref, _ = oracle.cg_var_ell(E, Pv, b, np.zeros(n), 20000, 1e-8)
    M = hip.EllMatrix.from_dict(E)
    xd = hip.DeviceVector.zeros(n)
    used, rel_def = hip.pcg_ell(M, hip.DeviceVector.from_numpy(Pv), hip.DeviceVector.from_numpy(b), xd, 20000, 1e-8)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert rel_def < 1e-8
    assert _rel(xd.get(), x_ref) <= 1e-6


def test_pcg_jac_diffusion_513(hip, oracle):
    """ConjugateGradients<JAC> on ELL (conjugate_gradients.hh:1110-1155) on the 513^2
    diffusion matrix, right-hand side 3 of the batch."""
    D = P.diffusion_ell(9)
    n = D["rows"]
    assert n > 131072
    dd = 1.0 / _diag(D)
    b = P.rhs_diffusion(9, 3)
    x_ref, it_ref, _ = oracle.pcg_jac_ell(D, dd, b, np.zeros(n), 20000, 1e-8)
    M = hip.EllMatrix.from_dict(D)
    xd = hip.DeviceVector.zeros(n)
    used, defect = hip.pcg_jac_ell(M, hip.DeviceVector.from_numpy(dd), hip.DeviceVector.from_numpy(b), xd, 20000,
                                   1e-8)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert defect <= 1e-8
    assert _rel(xd.get(), x_ref) <= 1e-6
