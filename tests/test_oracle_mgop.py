"""The oracle's operator-list multigrid (MGCycleCreation V/W/F, MGSolver,
RISmoother, PROLMAT transfers, IRSolver) pinned against HONEI's fixtures
and against the reference's own templates compiled from source.

Pins: mg_TEST.cc:331-420 (W cycle, CG coarse, RI/jac, poisson/ levels 1..6:
|x - sol_6| <= 8e-3), ir_TEST.cc:30-100 (FP32 V cycle, CG<VAR> coarse inside
FP64 IR, poisson_advanced/sort_0 levels 1..4: |x - sol_4| <= 8e-3), and
bit-identical iterates and iteration counts against the compiled reference
for every cycle shape and both coarse solvers."""
import os

import numpy as np
import pytest

from honei_fixtures import TESTDATA, honei_eps_8e3, load_exp, load_mg_data

POISSON = os.path.join(TESTDATA, "poisson")
ADV = os.path.join(TESTDATA, "poisson_advanced", "sort_0")


def test_mgop_w_cycle_meets_fixture_solution(oracle):
    d = load_mg_data(POISSON, 6, np.float64, oracle)
    r = oracle.mgop_solve(d, d["b"], d["x"], cycle="W", coarse_var=False, max_iters=100, max_iters_coarse=100,
                          n_pre=4, n_post=4, eps=1e-8)
    ref = load_exp(os.path.join(POISSON, "sol_6"))
    assert r["used"] < 100
    assert np.max(np.abs(r["x"] - ref)) <= honei_eps_8e3()
    # the solve meets its own stopping test
    A = d["A"][-1]
    res = oracle.ell_defect(A, d["b"], r["x"])
    r0 = oracle.ell_defect(A, d["b"], d["x"])
    assert np.linalg.norm(res) < 1e-8 * np.linalg.norm(r0)


@pytest.mark.parametrize("cycle", ["V", "W", "F"])
@pytest.mark.parametrize("coarse_var", [False, True])
def test_mgop_matches_compiled_reference(oracle, reference, cycle, coarse_var):
    d = load_mg_data(POISSON, 6, np.float64, oracle)
    kw = dict(cycle=cycle, coarse_var=coarse_var, max_iters=100, max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8)
    mine = oracle.mgop_solve(d, d["b"], d["x"], **kw)
    ref = reference.mgop_solve(POISSON, d["b"].size, levels=6, **kw)
    assert (mine["used"], mine["used_coarse"]) == (ref["used"], ref["used_coarse"])
    assert np.array_equal(mine["x"], ref["x"])


def test_mgop_min_level_and_uneven_smoothing(oracle, reference):
    """Coarse level 2, 2 pre / 3 post sweeps, tighter eps: still the reference's bits."""
    d = load_mg_data(POISSON, 6, np.float64, oracle)
    kw = dict(cycle="V", coarse_var=True, min_level=2, max_iters=30, max_iters_coarse=50, n_pre=2, n_post=3,
              eps=1e-10)
    mine = oracle.mgop_solve(d, d["b"], d["x"], **kw)
    ref = reference.mgop_solve(POISSON, d["b"].size, levels=6, **kw)
    assert (mine["used"], mine["used_coarse"]) == (ref["used"], ref["used_coarse"])
    assert np.array_equal(mine["x"], ref["x"])


def test_ir_mgop_meets_fixture_and_matches_reference(oracle, reference):
    d32 = load_mg_data(ADV, 4, np.float32, oracle)
    d64 = load_mg_data(ADV, 4, np.float64, oracle)
    A64 = d64["A"][-1]
    r = oracle.ir_mgop(A64, d64["b"], d64["x"], d32, d32["x"], max_iters_inner=8, max_iters_coarse=100,
                       n_pre=4, n_post=4, eps_inner=1e-8, eps=1e-8, max_iters=100)
    sol = load_exp(os.path.join(ADV, "sol_4"))
    assert np.max(np.abs(r["x"] - sol)) <= honei_eps_8e3()
    ref = reference.ir_mgop(ADV, d64["b"].size, levels=4, max_iters_inner=8, max_iters_coarse=100, n_pre=4,
                            n_post=4, eps_inner=1e-8, eps=1e-8, max_iters=100)
    assert (r["used"], r["inner_used"], r["coarse_used"]) == (ref["used"], ref["inner_used"], ref["coarse_used"])
    assert np.array_equal(r["x"], ref["x"])
