"""The oracle's Multigrid<..., JAC, CYCLE::V, FIXED | MIXED> on Q1 bands and
on ELL matrices, with structured (NONE) or matrix (PROLMAT) transfers,
pinned bit for bit against the reference's own templates compiled from
source, and against HONEI's .bin fixtures with the tolerances of
poisson_mg_float_banded_TEST.cc (1e-2) and poisson_mg_double_banded_TEST.cc
(1e-4)."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import (bilinear_prolmat_csr, csr_transpose, dinv_reference, ell_diag, ell_of_csr,
                            load_bin)


def _dinv(dd, dtype):
    if np.dtype(dtype) == np.float32:
        return dinv_reference(dd.astype(np.float32))
    inv = np.where(dd != 0, 1.0 / np.where(dd != 0, dd, 1.0), 0.0)
    return inv * 0.7


def _hierarchy(oracle, minl, maxl, dtype, kind, prolmat=False):
    a, dinv, prol, res = {}, {}, {}, {}
    for L in range(minl, maxl + 1):
        n = P.side_of_level(L) ** 2
        bands = [b.astype(dtype) for b in P.poisson_q1_bands(n)]
        if kind == "q1":
            a[L] = bands
            dinv[L] = _dinv(bands[4], dtype)
        else:
            E = P.q1_to_ell(bands)
            a[L] = E
            dinv[L] = _dinv(ell_diag(E), dtype)
        if prolmat and L > minl:
            csr = bilinear_prolmat_csr(L)
            prol[L] = ell_of_csr(oracle, *csr, dtype)
            res[L] = ell_of_csr(oracle, *csr_transpose(*csr), dtype)
    return a, dinv, (prol or None), (res or None)


@pytest.mark.parametrize("prolmat", [False, True])
def test_mg_mixed_ell_matches_reference(oracle, reference, prolmat):
    maxl, minl = 5, 1
    n = P.side_of_level(maxl) ** 2
    a32, d32, pr, rs = _hierarchy(oracle, minl, maxl, np.float32, "ell", prolmat)
    A64 = P.q1_to_ell(P.poisson_q1_bands(n))
    b = P.rhs_uniform(n)
    kw = dict(kind="ell", b=b, x0=np.zeros(n), A64=A64, a32=a32, dinv32=d32, prol32=pr, res32=rs,
              min_level=minl, max_level=maxl, n_max_iter=2, tolerance=1e-8)
    mine = oracle.mg_mixed(**kw)
    ref = reference.mg_mixed(**kw)
    assert mine["status"] == 0 and mine["outer"] > 0
    assert np.array_equal(mine["x"], ref["x"])
    assert mine["final_defect"] < 1e-8 * mine["initial_defect"]


@pytest.mark.parametrize("kind", ["q1", "ell"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mg_fixed_matches_reference(oracle, reference, kind, dtype):
    maxl, minl = 5, 1
    n = P.side_of_level(maxl) ** 2
    a, d, _, _ = _hierarchy(oracle, minl, maxl, dtype, kind)
    b = P.rhs_uniform(n).astype(dtype)
    tol = 1e-2 if dtype == np.float32 else 1e-8
    kw = dict(kind=kind, b=b, x0=np.zeros(n, dtype), a=a, dinv=d, min_level=minl, max_level=maxl, n_max_iter=16,
              tolerance=tol, n_pre=2, convergence_check=True)
    mine = oracle.mg_fixed(**kw)
    ref = reference.mg_fixed(**kw)
    assert mine["status"] == 0 and 0 < mine["iters"] <= 16
    assert np.array_equal(mine["x"], ref["x"])


def test_mg_fixed_ell_prolmat_matches_reference(oracle, reference):
    maxl, minl = 4, 1
    n = P.side_of_level(maxl) ** 2
    a, d, pr, rs = _hierarchy(oracle, minl, maxl, np.float64, "ell", prolmat=True)
    b = P.rhs_uniform(n)
    kw = dict(kind="ell", b=b, x0=np.zeros(n), a=a, dinv=d, prol=pr, res=rs, min_level=minl, max_level=maxl,
              n_max_iter=50, tolerance=1e-8, n_pre=2, initial_zero=False, convergence_check=True)
    assert np.array_equal(oracle.mg_fixed(**kw)["x"], reference.mg_fixed(**kw)["x"])


@pytest.mark.parametrize("n,dtype,n_pre,tol,conv,bound", [(81, np.float32, 4, 1e-2, True, 1e-2),
                                                           (1089, np.float64, 2, 1e-8, False, 1e-4)])
def test_mg_fixed_fixture(oracle, n, dtype, n_pre, tol, conv, bound):
    """poisson_mg_float_banded_TEST.cc (81.bin, FP32) and
    poisson_mg_double_banded_TEST.cc (1089.bin, FP64): levels from the .bin
    hierarchy, n_max_iter 16, initial_zero, |ref_sol - x| within the tests' bound."""
    F = load_bin(n)
    maxl = {81: 3, 1089: 5}[n]
    a, d = {}, {}
    for L in range(1, maxl + 1):
        bands = [x.astype(dtype) for x in load_bin(P.side_of_level(L) ** 2)["bands"]]
        a[L] = bands
        d[L] = _dinv(bands[4], dtype)
    r = oracle.mg_fixed(kind="q1", b=F["b"].astype(dtype), x0=np.zeros(n, dtype), a=a, dinv=d, min_level=1,
                        max_level=maxl, n_max_iter=16, tolerance=tol, n_pre=n_pre, convergence_check=conv)
    assert np.max(np.abs(r["x"].astype(np.float64) - F["ref_sol"])) <= bound
