"""The allocation-free CG entry point (hsolve_cg_ell_noalloc) against
hsolve_cg_ell, both called through the C-ABI: the same iteration count, the
same final relative defect and the same iterate bit for bit (the two run
the same kernels on the same grid; only where the workspace lives differs).

Two solves in a row share the device's fixed scratch array, the second on a
smaller matrix, so a solve that read what the previous one left behind
would show here.  A matrix above the scratch's capacity is refused with
HSOLVE_ERR_SIZE before anything runs."""
import ctypes as C

import numpy as np
import pytest

from honei_amd import problems as P
from honei_amd._lib import lib, u64, vp

pytestmark = pytest.mark.gpu

HSOLVE_OK, HSOLVE_ERR_SIZE = 0, 2


def _cg(hip, fn, M, b, x0, max_iters=10000, eps=1e-8):
    bd = hip.DeviceVector.from_numpy(b)
    xd = hip.DeviceVector.from_numpy(x0)
    used, defect = u64(), C.c_double()
    rc = getattr(lib(), fn)(vp(M.handle), vp(bd.ptr), vp(xd.ptr), u64(max_iters), float(eps), C.byref(used),
                            C.byref(defect), None)
    return rc, used.value, defect.value, xd.get()


def _poisson(side, dtype, seed):
    n = side * side
    E = P.q1_to_ell(P.poisson_q1_bands(n))
    E["Ax"] = E["Ax"].astype(dtype)
    b = P.rhs_uniform(n, seed=seed).astype(dtype)
    return E, b


@pytest.mark.parametrize("dtype", (np.float64, np.float32))
def test_noalloc_cg_equals_cg_ell(hip, dtype):
    # 257^2 first, then 65^2: the second solve runs on a scratch array that
    # still holds the first solve's vectors and scalars
    eps = 1e-8 if dtype == np.float64 else 1e-5
    for side, seed, x0_kind in ((257, 11, "zero"), (65, 12, "random"), (129, 13, "zero")):
        E, b = _poisson(side, dtype, seed)
        n = b.size
        x0 = np.zeros(n, dtype) if x0_kind == "zero" else \
            np.random.default_rng(seed).uniform(-1, 1, n).astype(dtype)
        M = hip.EllMatrix.from_dict(E)
        ref = _cg(hip, "hsolve_cg_ell", M, b, x0, eps=eps)
        got = _cg(hip, "hsolve_cg_ell_noalloc", M, b, x0, eps=eps)
        assert ref[0] == HSOLVE_OK and got[0] == HSOLVE_OK, (side, ref[0], got[0])
        assert got[1] == ref[1], (side, got[1], ref[1])
        assert got[2] == ref[2], (side, got[2], ref[2])
        assert np.array_equal(got[3], ref[3]), side
        del M


def test_noalloc_cg_iteration_cap(hip):
    # a cap below convergence: both report HSOLVE_ERR_NOT_CONVERGED after
    # exactly that many steps, with the same iterate
    E, b = _poisson(129, np.float64, 21)
    M = hip.EllMatrix.from_dict(E)
    x0 = np.zeros(b.size)
    ref = _cg(hip, "hsolve_cg_ell", M, b, x0, max_iters=7)
    got = _cg(hip, "hsolve_cg_ell_noalloc", M, b, x0, max_iters=7)
    assert ref[0] == 5 and got[0] == 5
    assert got[1] == ref[1] == 7
    assert np.array_equal(got[3], ref[3])


def test_noalloc_cg_refuses_above_capacity(hip):
    # diagonal FP64 matrix of 2^21 + 2^18 rows: its workspace exceeds 64 MiB
    n = (1 << 21) + (1 << 18)
    E = dict(rows=n, cols=n, stride=n, ncpr=1, threads=1,
             Aj=np.arange(n, dtype=np.uint64), Ax=np.full(n, 2.0), Arl=np.ones(n, np.uint64))
    M = hip.EllMatrix.from_dict(E)
    rc, _, _, _ = _cg(hip, "hsolve_cg_ell_noalloc", M, np.ones(n), np.zeros(n))
    assert rc == HSOLVE_ERR_SIZE
    assert b"scratch" in lib().hsolve_last_error()
