"""Readers for the HONEI test-data formats kept under tests/golden/honei_testdata
(copied data files of honei/math/testdata).

* FEAST Q1 ``.bin``: int32 n, then 12 float64 arrays of n: DD, LL, LD, LU, DL,
  DU, UL, UD, UU, b, ana_sol, ref_sol (honei/math/poisson_mg_mixed_banded_TEST.cc:67-118).
* ELL binary ``.ell``: 5 x uint64 header size, rows, columns, stride,
  num_cols_per_row, then uint64 Aj[size] and float64 Ax[size]
  (honei/math/matrix_io.hh:468-583).
* EXP text vectors (``rhs_6``, ``sol_6``): '#'-lines skipped, one value per
  line read with atof (honei/math/vector_io.hh:52-116).
* DV binary vectors: uint64 size + float64[size] (vector_io.hh:205-250).
"""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TESTDATA = os.path.join(GOLDEN, "honei_testdata")

# Band order of BandedMatrixQx (honei/la/band_type.hh): LL LD LU DL DD DU UL UD UU
BIN_ORDER = ("DD", "LL", "LD", "LU", "DL", "DU", "UL", "UD", "UU")
BAND_ORDER = ("LL", "LD", "LU", "DL", "DD", "DU", "UL", "UD", "UU")


def load_bin(n_or_path):
    path = n_or_path if isinstance(n_or_path, str) else os.path.join(TESTDATA, f"{n_or_path}.bin")
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.int32)[0])
        arrs = np.frombuffer(f.read(), np.float64)
    assert arrs.size >= 12 * n, (path, arrs.size, n)
    arrs = arrs[: 12 * n].reshape(12, n).copy()
    named = dict(zip(BIN_ORDER, arrs[:9]))
    bands = [named[k].copy() for k in BAND_ORDER]
    return dict(n=n, bands=bands, b=arrs[9].copy(), ana_sol=arrs[10].copy(), ref_sol=arrs[11].copy())


def load_ell(path):
    path = path if os.path.isabs(path) else os.path.join(TESTDATA, path)
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(40), np.uint64)
        size, rows, cols, stride, ncpr = (int(v) for v in hdr)
        Aj = np.frombuffer(f.read(8 * size), np.uint64).copy()
        Ax = np.frombuffer(f.read(8 * size), np.float64).copy()
    return dict(size=size, rows=rows, cols=cols, stride=stride, ncpr=ncpr, threads=1, Aj=Aj, Ax=Ax)


def load_exp(path):
    path = path if os.path.isabs(path) else os.path.join(TESTDATA, path)
    vals = []
    with open(path, "r") as f:
        lines = f.read().split("\n")
    # std::getline loop: the fragment after the last newline is dropped
    for line in lines[:-1]:
        if "#" in line:
            continue
        s = line.strip()
        try:
            vals.append(float(s) if s else 0.0)
        except ValueError:
            vals.append(0.0)
    return np.asarray(vals, np.float64)


def load_dv(path):
    path = path if os.path.isabs(path) else os.path.join(TESTDATA, path)
    with open(path, "rb") as f:
        size = int(np.frombuffer(f.read(8), np.uint64)[0])
        return np.frombuffer(f.read(8 * size), np.float64).copy()


def level_of(n):
    for L in range(0, 16):
        if ((1 << L) + 1) ** 2 == n:
            return L
    raise ValueError(n)


def side_of(L):
    return (1 << L) + 1


def dinv_reference(dd32):
    """diags_inverted as the MG tests build it: ElementInverse then
    Scale(v, 0.7) with a double factor (poisson_mg_mixed_banded_TEST.cc:411-419)."""
    inv = dd32.copy()
    nz = inv != 0
    inv[nz] = np.float32(1.0) / inv[nz]
    return (inv.astype(np.float64) * 0.7).astype(np.float32)


def csr_from_dense(D):
    rows, cols = D.shape
    rowptr = [0]
    col, val = [], []
    for r in range(rows):
        nz = np.nonzero(D[r])[0]
        col.extend(nz.tolist())
        val.extend(D[r, nz].tolist())
        rowptr.append(len(col))
    return np.asarray(rowptr, np.uint64), np.asarray(col, np.uint64), np.asarray(val, D.dtype)


def q1_dense(bands, n):
    """Dense matrix of a Q1 banded matrix, with the product's index guards."""
    m = int(np.sqrt(n))
    offs = (-m - 1, -m, -m + 1, -1, 0, 1, m - 1, m, m + 1)
    D = np.zeros((n, n), bands[0].dtype)
    for b, o in zip(bands, offs):
        for i in range(n):
            j = i + o
            if 0 <= j < n:
                D[i, j] += b[i]
    return D


def q1_csr(bands, n, drop_zeros=True):
    """CSR (ascending columns) of a Q1 banded matrix, as SparseMatrix holds it."""
    m = int(np.sqrt(n))
    offs = np.array((-m - 1, -m, -m + 1, -1, 0, 1, m - 1, m, m + 1), np.int64)
    B = np.stack(bands)  # 9 x n
    i = np.arange(n, dtype=np.int64)
    J = i[None, :] + offs[:, None]
    keep = (J >= 0) & (J < n)
    if drop_zeros:
        keep &= B != 0
    # columns ascend with the offset order for every row
    counts = keep.sum(axis=0)
    rowptr = np.zeros(n + 1, np.uint64)
    rowptr[1:] = np.cumsum(counts)
    col = J.T[keep.T].astype(np.uint64)
    val = B.T[keep.T]
    return rowptr, col, val


# --------------------------------------------------------------------------
# Operator-list multigrid data as MGUtil::load_data builds it
# (honei/math/mg.hh:449-545): A_i.ell, the "jac" preconditioner
# P = damping / diag(A) (PreconFill, mg.hh:254-283), prol_i.ell and
# res = Transposition(prol) (transposition.hh:30-43), rhs / init vectors.
# --------------------------------------------------------------------------
def ell_from_file(path, dtype, oracle):
    """MatrixIO<ELL>::read_matrix(path, DT): values converted to DT, row
    lengths derived by the raw constructor's first-zero scan."""
    E = load_ell(path)
    Ax = E["Ax"].astype(dtype)
    Arl = oracle.ell_row_length(E["rows"], E["stride"], 1, Ax)
    return dict(rows=E["rows"], cols=E["cols"], stride=E["stride"], ncpr=E["ncpr"], threads=1,
                Aj=E["Aj"], Ax=Ax, Arl=Arl)


def ell_entries(E):
    """(row, col, val) of the stored entries, row by row in slot order."""
    rows = np.arange(E["rows"])
    rr, cc, vv = [], [], []
    for k in range(E["ncpr"]):
        m = k < E["Arl"]
        idx = rows[m] + k * E["stride"]
        rr.append(rows[m])
        cc.append(E["Aj"][idx].astype(np.int64))
        vv.append(E["Ax"][idx])
    return np.concatenate(rr), np.concatenate(cc), np.concatenate(vv)


def ell_diag(E):
    r, c, v = ell_entries(E)
    d = np.zeros(E["rows"], E["Ax"].dtype)
    m = r == c
    d[r[m]] = v[m]
    return d


def ell_transpose(E, oracle):
    """SparseMatrixELL of the transpose, each row's entries in ascending column order."""
    r, c, v = ell_entries(E)
    order = np.lexsort((r, c))
    nrows = E["cols"]
    rowptr = np.zeros(nrows + 1, np.uint64)
    np.add.at(rowptr, c[order] + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.uint64)
    T = oracle.ell_from_csr(nrows, rowptr, r[order].astype(np.uint64), v[order])
    T["cols"] = E["rows"]
    return T


def load_mg_data(base, levels, dtype, oracle, damping=0.7):
    """A[i], P[i] for internal levels i < levels; prol[i], res[i] (level i -> i+1)
    for i < levels - 1; b = rhs_<levels>, x = init_<levels>."""
    base = base if os.path.isabs(base) else os.path.join(TESTDATA, base)
    A, P, prol, res = [], [], [], []
    for i in range(levels):
        Ai = ell_from_file(os.path.join(base, f"A_{i + 1}.ell"), dtype, oracle)
        A.append(Ai)
        P.append((np.full(Ai["rows"], dtype(damping), dtype) / ell_diag(Ai)).astype(dtype))
        if i < levels - 1:
            Pi = ell_from_file(os.path.join(base, f"prol_{i + 2}.ell"), dtype, oracle)
            prol.append(Pi)
            res.append(ell_transpose(Pi, oracle))
    rhs, init = os.path.join(base, f"rhs_{levels}"), os.path.join(base, f"init_{levels}")
    b = load_exp(rhs).astype(dtype) if os.path.exists(rhs) else None
    x = load_exp(init).astype(dtype) if os.path.exists(init) else None
    return dict(A=A, P=P, prol=prol, res=res, b=b, x=x, base=base)


def honei_eps_8e3():
    """The tolerance of mg_TEST.cc / ir_TEST.cc: 8 * (m * sizeof(double) + b)."""
    base_eps = 1 / 10 ** 1
    add_eps = base_eps / 10 ** 2
    m = (add_eps - base_eps) / 4.0
    b = base_eps - 4.0 * m
    return (m * 8 + b) * 8.0


def bilinear_prolmat_csr(L, dirichlet=True):
    """Natural-order bilinear prolongation (level L-1 -> L) as CSR with
    ascending columns: fine node (R, S) takes h(R - 2i) h(S - 2j) from coarse
    node (i, j), h(0) = 1, h(+-1) = 1/2; with `dirichlet` the rows of
    boundary fine nodes are empty."""
    sf, sc = (1 << L) + 1, (1 << (L - 1)) + 1
    h = {0: 1.0, 1: 0.5, -1: 0.5}
    rowptr, col, val = [0], [], []
    for R in range(sf):
        for S in range(sf):
            if not (dirichlet and (R in (0, sf - 1) or S in (0, sf - 1))):
                for i in range(max(0, (R - 1) // 2), min(sc, R // 2 + 2)):
                    for j in range(max(0, (S - 1) // 2), min(sc, S // 2 + 2)):
                        dr, ds = R - 2 * i, S - 2 * j
                        if dr in h and ds in h:
                            col.append(i * sc + j)
                            val.append(h[dr] * h[ds])
            rowptr.append(len(col))
    return (sf * sf, sc * sc, np.asarray(rowptr, np.uint64), np.asarray(col, np.uint64),
            np.asarray(val, np.float64))


def csr_transpose(rows, cols, rowptr, col, val):
    r = np.repeat(np.arange(rows, dtype=np.uint64), np.diff(rowptr.astype(np.int64)))
    order = np.lexsort((r, col))
    tptr = np.zeros(cols + 1, np.uint64)
    np.add.at(tptr, col.astype(np.int64) + 1, 1)
    return cols, rows, np.cumsum(tptr).astype(np.uint64), r[order], val[order]


def ell_of_csr(oracle, rows, cols, rowptr, col, val, dtype):
    E = oracle.ell_from_csr(rows, rowptr, col, np.asarray(val, dtype))
    E["cols"] = cols
    return E
