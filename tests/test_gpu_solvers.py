"""GPU parity of the solvers (CGSolver NONE/VAR, MG-IR MIXED) against the
oracle and against HONEI's own fixtures.

Tolerances (floating point): iteration counts within +-1 of the oracle;
solutions within 1e-6 relative (FP64) of the oracle's; fixture reference
solutions within the reference tests' own bounds."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import dinv_reference, load_bin, load_ell, load_exp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_cg_q1_poisson_257(hip, oracle):
    """Config: plain FP64 CG on Q1 Dirichlet 257^2, b ~ U[-1,1] (boundary 0), x0 = 0, eps 1e-8."""
    n = 257 * 257
    bands = P.poisson_q1_bands(n)
    b = P.rhs_uniform(n)
    x_ref, it_ref, st = oracle.cg_none_q1(bands, b, np.zeros(n), 10000, 1e-8)
    assert st == 0
    A = hip.Q1Matrix(bands)
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    used, rel_def = hip.cg_q1(A, bd, xd, 10000, 1e-8)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(xd.get(), x_ref) <= 1e-6
    assert rel_def < 1e-8
    # true residual of the returned solution
    r = oracle.q1_defect(bands, b, xd.get())
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-7


def test_cg_ell_poisson_257(hip, oracle):
    n = 257 * 257
    E = P.q1_to_ell(P.poisson_q1_bands(n))
    b = P.rhs_uniform(n)
    x_ref, it_ref, st = oracle.cg_none_ell(E, b, np.zeros(n), 10000, 1e-8)
    M = hip.EllMatrix.from_dict(E)
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    used, rel_def = hip.cg_ell(M, bd, xd, 10000, 1e-8)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(xd.get(), x_ref) <= 1e-6


@pytest.mark.parametrize("side,seed", [(33, 0), (65, 0), (65, 3), (129, 1)])
def test_cg_fp32_q1(hip, oracle, side, seed):
    """Pure FP32 CGSolver<NONE> (eps 1e-5): the oracle sums its dot products
    sequentially in float (dot_product.hh:72-87), the device in a fixed tree
    in double; iteration counts within +-1, solutions within 1e-4."""
    n = side * side
    bands = [b.astype(np.float32) for b in P.poisson_q1_bands(n)]
    b = P.rhs_uniform(n, seed=P.SEED + seed, dtype=np.float32)
    x_ref, it_ref, _ = oracle.cg_none_q1(bands, b, np.zeros(n, np.float32), 2000, np.float32(1e-5))
    A = hip.Q1Matrix(bands)
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n, np.float32)
    used, _ = hip.cg_q1(A, bd, xd, 2000, 1e-5, raise_on_cap=False)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(xd.get().astype(np.float64), x_ref.astype(np.float64)) <= 1e-4
    E = P.q1_to_ell(bands)
    x_ref, it_ref, _ = oracle.cg_none_ell(E, b, np.zeros(n, np.float32), 2000, np.float32(1e-5))
    M = hip.EllMatrix.from_dict(E)
    xd = hip.DeviceVector.zeros(n, np.float32)
    used, _ = hip.cg_ell(M, bd, xd, 2000, 1e-5, raise_on_cap=False)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(xd.get().astype(np.float64), x_ref.astype(np.float64)) <= 1e-4


def test_pcg_ell_fixture_A6(hip, oracle):
    """cg_TEST.cc:359-440: CGSolver<VAR>, Jacobi(0.7) preconditioner, poisson/A_6.ell,
    rhs_6 / init_6, 700 iterations max, eps 1e-8, compared with FEATFLOW2 sol_6 at 3e-5."""
    F = load_ell("poisson/A_6.ell")
    E = dict(F)
    E["Arl"] = oracle.ell_row_length(F["rows"], F["stride"], 1, F["Ax"])
    rhs, init, sol = load_exp("poisson/rhs_6"), load_exp("poisson/init_6"), load_exp("poisson/sol_6")
    n = F["rows"]
    diag = np.zeros(n)
    Aj = F["Aj"].reshape(F["ncpr"], F["stride"])
    Ax = F["Ax"].reshape(F["ncpr"], F["stride"])
    for k in range(F["ncpr"]):
        hit = (Aj[k, :n] == np.arange(n)) & (Ax[k, :n] != 0)   # padding slots hold (col 0, 0.0)
        diag[hit] = Ax[k, :n][hit]
    Pv = 0.7 / diag
    x_ref, it_ref, _ = oracle.cg_var_ell(E, Pv, rhs, init, 700, 1e-8)
    M = hip.EllMatrix.from_dict(E)
    xd = hip.DeviceVector.from_numpy(init)
    used, _ = hip.pcg_ell(M, hip.DeviceVector.from_numpy(Pv), hip.DeviceVector.from_numpy(rhs), xd, 700, 1e-8,
                          raise_on_cap=False)
    x = xd.get()
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(x, x_ref) <= 1e-6
    assert np.max(np.abs(x - sol)) <= 3e-5


def _levels(maxl, minl=1, from_bin=False):
    a32, d32 = {}, {}
    for L in range(minl, maxl + 1):
        n = P.side_of_level(L) ** 2
        bands = load_bin(n)["bands"] if from_bin else P.poisson_q1_bands(n)
        a32[L] = [b.astype(np.float32) for b in bands]
        d32[L] = dinv_reference(a32[L][4])
    return a32, d32


def test_mg_mixed_fixture_4225(hip, oracle):
    """poisson_mg_mixed_banded_TEST.cc:243-427: levels 1..6 from the .bin files,
    n_max_iter 2, 2+2 sweeps, tol 1e-8; |ref_sol - x| <= 1e-4."""
    F = load_bin(4225)
    a32, d32 = _levels(6, 1, from_bin=True)
    ref = oracle.mg_mixed_q1(F["bands"], F["b"], np.zeros(F["n"]), a32, d32, min_level=1, max_level=6)
    mg = hip.MGSolver(hip.mg_params(1, 6), F["bands"], a32, d32)
    bd, xd = hip.DeviceVector.from_numpy(F["b"]), hip.DeviceVector.zeros(F["n"])
    info = mg.solve(bd, xd)
    x = xd.get()
    assert info["outer"] == ref["outer"] and info["inner"] == ref["inner"], (info, ref["outer"], ref["inner"])
    assert _rel(x, ref["x"]) <= 1e-6
    assert np.max(np.abs(x - F["ref_sol"])) <= 1e-4


@pytest.mark.parametrize("maxl,minl", [(7, 1), (9, 2), (10, 2), (6, 4), (7, 5)])
def test_mg_mixed_poisson_levels(hip, oracle, maxl, minl):
    """MG-IR on Q1 Poisson against the oracle's Multigrid<MIXED>.  The
    level pairs cover the engine's bottom-levels kernel with one to five
    levels below level 6 ((7, 1), (9, 2), (6, 4)) and a coarsest level too
    large for it ((7, 5): the coarse CG in global memory)."""
    n = P.side_of_level(maxl) ** 2
    A64 = P.poisson_q1_bands(n)
    b = P.rhs_uniform(n)
    a32, d32 = _levels(maxl, minl)
    ref = oracle.mg_mixed_q1(A64, b, np.zeros(n), a32, d32, min_level=minl, max_level=maxl, tolerance=1e-10)
    mg = hip.MGSolver(hip.mg_params(minl, maxl, tolerance=1e-10))   # device-assembled hierarchy
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    info = mg.solve(bd, xd)
    assert abs(info["outer"] - ref["outer"]) <= 1 and abs(info["inner"] - ref["inner"]) <= 2, (info, ref)
    assert _rel(xd.get(), ref["x"]) <= 1e-6
    assert info["final_defect"] < (1e-10 * 1.0000001) * info["initial_defect"] or \
        info["final_defect"] < float(np.float32(1e-10)) * info["initial_defect"]


def test_mg_value_codes_same_bits(hip):
    """Poisson bands hold two values per band, so the hierarchy keeps value
    codes and the V-cycle and outer-defect kernels read them; with
    hsolve_mg_set_value_codes(0) the same handle reads the plain bands.  Both
    give the same counts, defects and iterate bits (device-assembled and
    host-uploaded hierarchies)."""
    for maxl, minl, host in ((9, 2, False), (10, 2, False), (9, 2, True)):
        n = P.side_of_level(maxl) ** 2
        b = P.rhs_uniform(n)
        if host:
            a32, d32 = _levels(maxl, minl)
            mg = hip.MGSolver(hip.mg_params(minl, maxl, tolerance=1e-10, max_outer=40), P.poisson_q1_bands(n),
                              a32, d32)
        else:
            mg = hip.MGSolver(hip.mg_params(minl, maxl, tolerance=1e-10, max_outer=40))
        bd = hip.DeviceVector.from_numpy(b)
        runs = []
        for enable in (True, False, True):
            mg.set_value_codes(enable)
            xd = hip.DeviceVector.zeros(n)
            info = mg.solve(bd, xd)
            runs.append((info["outer"], info["inner"], info["final_defect"], xd.get()))
        for r in runs[1:]:
            assert r[:3] == runs[0][:3], (maxl, host, r[:3], runs[0][:3])
            assert np.array_equal(r[3], runs[0][3]), (maxl, host)


def _scaled_diagonal(bands, seed):
    """Q1 bands with the diagonal scaled node by node by 1 + k 2^-20, k in
    0..31: 32 distinct diagonal values (FP32 too), more than a value code
    holds.  (Steps of 2^-10 make the reference's MG-IR diverge: the oracle
    ends in NaN.)"""
    g = np.random.default_rng(seed)
    out = [np.array(x, copy=True) for x in bands]
    f = 1.0 + g.integers(0, 32, out[4].size) * 2.0 ** -20
    out[4] = (out[4] * f).astype(out[4].dtype)
    return out


def test_mg_plain_bands_when_values_exceed_codes(hip, oracle):
    """A finest level whose diagonal holds more than 16 values keeps the plain
    bands, FP32 and FP64 (the coarser Poisson levels keep their codes: the
    hierarchy holds fewer device bytes than the all-Poisson one), and the
    solve still meets the oracle's Multigrid<MIXED>.  Both sides cap the outer
    steps, so a solve that did not converge would end and fail the test."""
    maxl, minl = 9, 2
    n = P.side_of_level(maxl) ** 2
    A64 = _scaled_diagonal(P.poisson_q1_bands(n), 1)
    a32, d32 = {}, {}
    for L in range(minl, maxl + 1):
        nl = P.side_of_level(L) ** 2
        a32[L] = [x.astype(np.float32) for x in (A64 if L == maxl else P.poisson_q1_bands(nl))]
        d32[L] = dinv_reference(a32[L][4])
    b = P.rhs_uniform(n)
    ref = oracle.mg_mixed_q1(A64, b, np.zeros(n), a32, d32, min_level=minl, max_level=maxl, tolerance=1e-10,
                             max_outer=40)
    assert ref["status"] == 0 and ref["outer"] < 40, ref["outer"]
    prm = hip.mg_params(minl, maxl, tolerance=1e-10, max_outer=40)
    mg = hip.MGSolver(prm, A64, a32, d32)
    pa32, pd32 = _levels(maxl, minl)
    coded = hip.MGSolver(prm, P.poisson_q1_bands(n), pa32, pd32)
    assert mg.device_bytes < coded.device_bytes
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    info = mg.solve(bd, xd)
    assert abs(info["outer"] - ref["outer"]) <= 1 and abs(info["inner"] - ref["inner"]) <= 2, (info, ref)
    assert _rel(xd.get(), ref["x"]) <= 1e-6


def _jac_inverse(E, dtype):
    n = E["rows"]
    Aj = E["Aj"].reshape(E["ncpr"], E["stride"])[:, :n]
    Ax = E["Ax"].reshape(E["ncpr"], E["stride"])[:, :n]
    diag = np.where((Aj == np.arange(n)) & (Ax != 0), Ax, 0).sum(axis=0).astype(dtype)
    return (dtype(1) / diag).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pcg_jac_area51(hip, oracle, dtype):
    """conjugate_gradients_TEST.cc:620-700: ConjugateGradients<JAC> on l2/area51_full_0.ell
    with eps 1e-8 against FEATFLOW2's solution (3e-5 FP64, 3e-3 FP32)."""
    F = load_ell("l2/area51_full_0.ell")
    E = dict(F)
    E["Arl"] = oracle.ell_row_length(F["rows"], F["stride"], 1, F["Ax"])
    E["Ax"] = F["Ax"].astype(dtype)
    dd = _jac_inverse(E, dtype)
    rhs = load_exp("l2/area51_rhs_0").astype(dtype)
    init = load_exp("l2/area51_init_0").astype(dtype)
    sol = load_exp("l2/area51_sol_0")
    x_ref, it_ref, _ = oracle.pcg_jac_ell(E, dd, rhs, init, 10000, dtype(1e-8))
    M = hip.EllMatrix.from_dict(E)
    xd = hip.DeviceVector.from_numpy(init)
    used, _ = hip.pcg_jac_ell(M, hip.DeviceVector.from_numpy(dd), hip.DeviceVector.from_numpy(rhs), xd, 10000,
                              1e-8, raise_on_cap=False)
    x = xd.get().astype(np.float64)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert _rel(x, x_ref.astype(np.float64)) <= (1e-6 if dtype == np.float64 else 1e-4)
    assert np.max(np.abs(x - sol)) <= (3e-5 if dtype == np.float64 else 3e-3)


def test_pcg_jac_diffusion_level7(hip, oracle):
    """ConjugateGradients<JAC> on the variable-coefficient diffusion matrix (129^2 nodes):
    counts within +-1 of the oracle and the solution within 1e-6 relative."""
    D = P.diffusion_ell(7)
    dd = _jac_inverse(D, np.float64)
    b = P.rhs_diffusion(7, 3)
    x_ref, it_ref, _ = oracle.pcg_jac_ell(D, dd, b, np.zeros(D["rows"]), 20000, 1e-8)
    M = hip.EllMatrix.from_dict(D)
    xd = hip.DeviceVector.from_numpy(np.zeros(D["rows"]))
    used, defect = hip.pcg_jac_ell(M, hip.DeviceVector.from_numpy(dd), hip.DeviceVector.from_numpy(b), xd, 20000,
                                   1e-8)
    assert abs(used - it_ref) <= 1, (used, it_ref)
    assert defect <= 1e-8
    assert _rel(xd.get(), x_ref) <= 1e-6
