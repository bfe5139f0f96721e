"""GPU parity of the many-right-hand-side kernels on one ELL matrix
(column-blocked multivectors, row i's nrhs entries contiguous).

* hsolve_ell_spmm: every column bit-identical to the oracle's Product /
  Defect (product.hh:148-206, defect.hh:357-384) on that column, for the
  register-tile path (<= 16 slots per row) and the long-row path, 16-bit
  offset and 32-bit column indices, FP64 and FP32.
* hsolve_cg_ell_multi: every column against the oracle's CGSolver<VAR> /
  <NONE> (cg.hh:63-219) on that column: iteration counts within +-1,
  solutions within 1e-6 relative (FP64).  The per-column reductions sum in
  a different tree than the single-vector solver, so agreement with it is
  to the same tolerance.
* hsolve_rhs_uniform_multi: bit-identical to the host generator.
The variable-coefficient diffusion problem (64 x 64 log-normal tiles,
Dirichlet left/right, Neumann bottom/top) is the matrix of the solver tests;
at 513^2 with 64 right-hand sides the true residual of every column is
checked in FP64 on the host."""
import numpy as np
import pytest

from honei_amd import problems as P
from honei_fixtures import ell_diag

pytestmark = pytest.mark.gpu


def _mv(cols):
    return np.ascontiguousarray(np.stack(cols, axis=1).ravel())


def _col(mv, R, j):
    return mv.reshape(-1, R)[:, j]


_MATS = {
    "diffusion33_f64": lambda: P.diffusion_ell(5),
    "diffusion33_f32": lambda: P.diffusion_ell(5, dtype=np.float32),
    "matB_f64_int16": lambda: P.matrix_b_ell(5000, halfwidth=300),
    "matB_f64_u32": lambda: P.matrix_b_ell(70000, halfwidth=40000),
    "matB_long_f64": lambda: P.matrix_b_ell(3000, halfwidth=500, min_len=14, max_len=24),
    "matB_long_f32": lambda: P.matrix_b_ell(3000, halfwidth=500, min_len=14, max_len=24, dtype=np.float32),
}


@pytest.mark.parametrize("R", [1, 5, 64])
@pytest.mark.parametrize("name", sorted(_MATS))
def test_spmm_columns_match_oracle(hip, oracle, name, R):
    E = _MATS[name]()
    dt = E["Ax"].dtype
    n = E["rows"]
    A = hip.EllMatrix.from_dict(E)
    xs = [P.rng_uniform(P.SEED + j, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dt) for j in range(R)]
    rs = [P.rng_uniform(P.SEED + 100 + j, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dt)
          for j in range(R)]
    X = hip.DeviceVector.from_numpy(_mv(xs))
    RHS = hip.DeviceVector.from_numpy(_mv(rs))
    Y = hip.DeviceVector.zeros(n * R, dt)
    A.spmm(R, X, Y)
    y = Y.get()
    A.spmm(R, X, Y, rhs=RHS)
    d = Y.get()
    for j in sorted({0, R // 2, R - 1}):
        assert np.array_equal(_col(y, R, j), oracle.ell_product(E, xs[j])), (name, R, j)
        assert np.array_equal(_col(d, R, j), oracle.ell_defect(E, rs[j], xs[j])), (name, R, j)


@pytest.mark.parametrize("var", [True, False])
def test_cg_multi_matches_oracle(hip, oracle, var):
    level, R = 5, 6
    n = P.side_of_level(level) ** 2
    E = P.diffusion_ell(level)
    Pv = 1.0 / ell_diag(E)
    bs = [P.rhs_diffusion(level, j) for j in range(R)]
    x0s = [np.zeros(n)] * (R - 1) + [P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1)]
    x0s[-1][P.dirichlet_left_right(P.side_of_level(level))] = 0.0
    A = hip.EllMatrix.from_dict(E)
    B = hip.DeviceVector.from_numpy(_mv(bs))
    X = hip.DeviceVector.from_numpy(_mv(x0s))
    Pd = hip.DeviceVector.from_numpy(Pv) if var else None
    used, defect = hip.cg_ell_multi(A, Pd, B, X, R, 10000, 1e-8)
    x = X.get()
    for j in range(R):
        if var:
            xr, it, st = oracle.cg_var_ell(E, Pv, bs[j], x0s[j], 10000, 1e-8)
        else:
            xr, it, st = oracle.cg_none_ell(E, bs[j], x0s[j], 10000, 1e-8)
        assert st == 0
        assert abs(int(used[j]) - it) <= 1, (j, used[j], it)
        xj = _col(x, R, j)
        assert np.linalg.norm(xj - xr) <= 1e-6 * np.linalg.norm(xr), j
        assert defect[j] < 1e-8


def test_cg_multi_equals_single_column(hip):
    level, R = 6, 3
    n = P.side_of_level(level) ** 2
    E = P.diffusion_ell(level)
    Pv = 1.0 / ell_diag(E)
    bs = [P.rhs_diffusion(level, j) for j in range(R)]
    A = hip.EllMatrix.from_dict(E)
    Pd = hip.DeviceVector.from_numpy(Pv)
    X = hip.DeviceVector.zeros(n * R)
    used, _ = hip.cg_ell_multi(A, Pd, hip.DeviceVector.from_numpy(_mv(bs)), X, R, 10000, 1e-8)
    x = X.get()
    for j in range(R):
        b1, x1 = hip.DeviceVector.from_numpy(bs[j]), hip.DeviceVector.zeros(n)
        it, _ = hip.pcg_ell(A, Pd, b1, x1, 10000, 1e-8)
        assert abs(int(used[j]) - it) <= 1
        ref = x1.get()
        assert np.linalg.norm(_col(x, R, j) - ref) <= 1e-6 * np.linalg.norm(ref)


def test_cg_multi_cap_reports_not_converged(hip):
    level, R = 5, 2
    n = P.side_of_level(level) ** 2
    E = P.diffusion_ell(level)
    A = hip.EllMatrix.from_dict(E)
    B = hip.DeviceVector.from_numpy(_mv([P.rhs_diffusion(level, j) for j in range(R)]))
    X = hip.DeviceVector.zeros(n * R)
    used, _ = hip.cg_ell_multi(A, None, B, X, R, 7, 1e-8, raise_on_cap=False)
    assert list(used) == [7, 7]
    with pytest.raises(RuntimeError):
        hip.cg_ell_multi(A, None, B, hip.DeviceVector.zeros(n * R), R, 7, 1e-8)


@pytest.mark.parametrize("layout", ["blocked", "separate"])
def test_rhs_uniform_multi_matches_host(hip, layout):
    level, R = 6, 5
    side = P.side_of_level(level)
    n = side * side
    buf = hip.DeviceVector.zeros(n * R)
    rs, cs = (R, 1) if layout == "blocked" else (1, n)
    hip.rhs_uniform_multi(buf, side, R, rs, cs, mask=P.DIFFUSION_MASK)
    got = buf.get()
    for j in range(R):
        col = _col(got, R, j) if layout == "blocked" else got[j * n:(j + 1) * n]
        assert np.array_equal(col, P.rhs_diffusion(level, j))


def test_pcg_64_rhs_true_residual(hip):
    """513^2 diffusion, 64 right-hand sides in one batched solve: every
    column's true FP64 residual ||b - A x|| / ||b|| <= 1e-8 (+ the recursive
    residual's drift, bounded by 2e-8 here)."""
    level, R = 9, 64
    side = P.side_of_level(level)
    n = side * side
    E = P.diffusion_ell(level)
    A = hip.EllMatrix.from_dict(E)
    Pd = hip.DeviceVector.from_numpy(1.0 / ell_diag(E))
    B = hip.DeviceVector.zeros(n * R)
    hip.rhs_uniform_multi(B, side, R, R, 1, mask=P.DIFFUSION_MASK)
    X = hip.DeviceVector.zeros(n * R)
    used, defect = hip.cg_ell_multi(A, Pd, B, X, R, 20000, 1e-8)
    Rm = hip.DeviceVector.zeros(n * R)
    A.spmm(R, X, Rm, rhs=B)
    r = Rm.get().reshape(n, R)
    b = B.get().reshape(n, R)
    rel = np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)
    assert rel.max() <= 2e-8, rel.max()
    assert (defect < 1e-8).all()
    assert used.min() > 100 and used.max() - used.min() < 0.2 * used.max()


def _diffusion_hierarchy(minl, maxl):
    from honei_fixtures import dinv_reference
    a32 = {L: P.diffusion_ell(L, dtype=np.float32) for L in range(minl, maxl + 1)}
    d32 = {L: dinv_reference(ell_diag(a32[L])) for L in a32}
    return P.diffusion_ell(maxl), a32, d32


def test_mg_ir_diffusion_batch_matches_oracle(hip, oracle):
    """MG-IR (FP32 V-cycles, FP64 refinement) on the diffusion hierarchy with
    the Neumann border mask, 4 right-hand sides in the same ticks: each
    against the oracle's Multigrid<..., MIXED> (counts +-1, 1e-6), and the
    batch bit-identical to single solves."""
    maxl, minl, R = 7, 2, 4
    n = P.side_of_level(maxl) ** 2
    A64, a32, d32 = _diffusion_hierarchy(minl, maxl)
    prm = hip.mg_params(minl, maxl, tolerance=1e-8, mask=P.DIFFUSION_MASK)
    mg = hip.MGSolver.from_ell(prm, A64, a32, d32, slots=R)
    bs = [P.rhs_diffusion(maxl, j) for j in range(R)]
    bd = [hip.DeviceVector.from_numpy(b) for b in bs]
    xd = [hip.DeviceVector.zeros(n) for _ in range(R)]
    infos = mg.solve_batch(bd, xd)
    for j in (0, R - 1):
        ref = oracle.mg_mixed(kind="ell", b=bs[j], x0=np.zeros(n), A64=A64, a32=a32, dinv32=d32, min_level=minl,
                              max_level=maxl, n_max_iter=2, tolerance=1e-8, mask=P.DIFFUSION_MASK)
        assert abs(infos[j]["outer"] - ref["outer"]) <= 1 and abs(infos[j]["inner"] - ref["inner"]) <= 2
        x = xd[j].get()
        assert np.linalg.norm(x - ref["x"]) <= 1e-6 * np.linalg.norm(ref["x"])
        x1 = hip.DeviceVector.zeros(n)
        single = mg.solve(bd[j], x1)
        assert np.array_equal(x1.get(), x)
        assert (single["outer"], single["inner"]) == (infos[j]["outer"], infos[j]["inner"])


@pytest.mark.parametrize("R", [3, 64])
@pytest.mark.parametrize("level", [4, 6])
def test_spmm_stencil_window_equals_generic(hip, oracle, level, R):
    """The 3 x 3 window path for 9-point structured matrices (chosen at
    upload) gives the same bits as the generic tile path and the oracle."""
    E = P.diffusion_ell(level)
    n = E["rows"]
    A = hip.EllMatrix.from_dict(E)
    G = hip.EllMatrix.from_dict(E)
    G.set_variant(1)                        # an explicit variant selects the generic path
    assert A.q1_side == P.side_of_level(level)
    assert hip.EllMatrix.from_dict(P.matrix_b_ell(4096, halfwidth=100)).q1_side == 0
    xs = [P.rng_uniform(P.SEED + 7 * j, P.S_X, np.arange(n, dtype=np.uint64), -1, 1) for j in range(R)]
    X = hip.DeviceVector.from_numpy(_mv(xs))
    RHS = hip.DeviceVector.from_numpy(_mv(xs[::-1]))
    Y1, Y2 = hip.DeviceVector.zeros(n * R), hip.DeviceVector.zeros(n * R)
    for rhs in (None, RHS):
        A.spmm(R, X, Y1, rhs=rhs)
        G.spmm(R, X, Y2, rhs=rhs)
        y1 = Y1.get()
        assert np.array_equal(y1, Y2.get())
        j = R - 1
        ref = oracle.ell_product(E, xs[j]) if rhs is None else oracle.ell_defect(E, xs[::-1][j], xs[j])
        assert np.array_equal(_col(y1, R, j), ref)
