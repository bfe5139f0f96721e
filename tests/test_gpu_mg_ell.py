"""GPU parity of the device MG-IR engine on ELL matrices
(Multigrid<..., NONE | PROLMAT, JAC, CYCLE::V, MIXED>, multigrid.hh:1936-2543)
against the oracle's restatement, and of batched solves against single ones.

Tolerances: outer / inner iteration counts equal to the oracle's (+-1);
solutions within 1e-6 relative (FP64 result of the FP32-MG / FP64-IR solve).
A batch of right-hand sides solved in the same launches gives each solve the
same bits as solving it alone."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import bilinear_prolmat_csr, csr_transpose, dinv_reference, ell_diag, ell_of_csr

pytestmark = pytest.mark.gpu


def _hier(oracle, minl, maxl, prolmat):
    a, d, pr, rs = {}, {}, {}, {}
    for L in range(minl, maxl + 1):
        n = P.side_of_level(L) ** 2
        E = P.q1_to_ell([b.astype(np.float32) for b in P.poisson_q1_bands(n)])
        a[L] = E
        d[L] = dinv_reference(ell_diag(E))
        if prolmat and L > minl:
            csr = bilinear_prolmat_csr(L)
            pr[L] = ell_of_csr(oracle, *csr, np.float32)
            rs[L] = ell_of_csr(oracle, *csr_transpose(*csr), np.float32)
    return a, d, (pr or None), (rs or None)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prolmat", [False, True])
@pytest.mark.parametrize("maxl,minl", [(6, 1), (8, 2)])
def test_mg_ell_mixed_matches_oracle(hip, oracle, prolmat, maxl, minl):
    n = P.side_of_level(maxl) ** 2
    a32, d32, pr, rs = _hier(oracle, minl, maxl, prolmat)
    A64 = P.q1_to_ell(P.poisson_q1_bands(n))
    b = P.rhs_uniform(n)
    ref = oracle.mg_mixed(kind="ell", b=b, x0=np.zeros(n), A64=A64, a32=a32, dinv32=d32, prol32=pr, res32=rs,
                          min_level=minl, max_level=maxl, n_max_iter=2, tolerance=1e-8)
    mg = hip.MGSolver.from_ell(hip.mg_params(minl, maxl, tolerance=1e-8), A64, a32, d32, pr, rs)
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    info = mg.solve(bd, xd)
    assert abs(info["outer"] - ref["outer"]) <= 1 and abs(info["inner"] - ref["inner"]) <= 2, (info, ref)
    assert _rel(xd.get(), ref["x"]) <= 1e-6
    assert info["final_defect"] < 1e-8 * info["initial_defect"]


def test_mg_ell_batch_equals_single(hip, oracle):
    maxl, minl = 7, 2
    n = P.side_of_level(maxl) ** 2
    a32, d32, _, _ = _hier(oracle, minl, maxl, False)
    A64 = P.q1_to_ell(P.poisson_q1_bands(n))
    mg = hip.MGSolver.from_ell(hip.mg_params(minl, maxl, tolerance=1e-8), A64, a32, d32, slots=4)
    bs = [hip.DeviceVector.from_numpy(P.rhs_uniform(n, seed=P.SEED + j)) for j in range(4)]
    xs = [hip.DeviceVector.zeros(n) for _ in range(4)]
    infos = mg.solve_batch(bs, xs)
    for j in (0, 3):
        x1 = hip.DeviceVector.zeros(n)
        single = mg.solve(bs[j], x1)
        assert np.array_equal(x1.get(), xs[j].get())
        assert (single["outer"], single["inner"]) == (infos[j]["outer"], infos[j]["inner"])
    ref = oracle.mg_mixed(kind="ell", b=P.rhs_uniform(n, seed=P.SEED + 2), x0=np.zeros(n), A64=A64, a32=a32,
                          dinv32=d32, min_level=minl, max_level=maxl, n_max_iter=2, tolerance=1e-8)
    assert _rel(xs[2].get(), ref["x"]) <= 1e-6


def test_mg_ell_value_codes_same_bits(hip, oracle):
    """The Poisson ELL levels kept as bands hold two values per band, so the
    hierarchy keeps value codes for its levels of 65^2 and up; the solve with
    them and with the plain bands (hsolve_mg_set_value_codes) gives the same
    counts, defects and iterate bits, single and batched."""
    maxl, minl = 8, 2
    n = P.side_of_level(maxl) ** 2
    a32, d32, _, _ = _hier(oracle, minl, maxl, False)
    A64 = P.q1_to_ell(P.poisson_q1_bands(n))
    mg = hip.MGSolver.from_ell(hip.mg_params(minl, maxl, tolerance=1e-8, max_outer=40), A64, a32, d32, slots=2)
    bs = [hip.DeviceVector.from_numpy(P.rhs_uniform(n, seed=P.SEED + j)) for j in range(2)]
    runs = []
    for enable in (True, False):
        mg.set_value_codes(enable)
        x1 = hip.DeviceVector.zeros(n)
        single = mg.solve(bs[0], x1)
        xs = [hip.DeviceVector.zeros(n) for _ in range(2)]
        infos = mg.solve_batch(bs, xs)
        runs.append(((single["outer"], single["inner"], single["final_defect"]), x1.get(),
                     [(i["outer"], i["inner"]) for i in infos], [x.get() for x in xs]))
    assert runs[0][0] == runs[1][0] and runs[0][2] == runs[1][2]
    assert np.array_equal(runs[0][1], runs[1][1])
    for a, b in zip(runs[0][3], runs[1][3]):
        assert np.array_equal(a, b)
