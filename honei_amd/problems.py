"""Synthetic problem generators (host side, numpy): the inputs of the
benchmarks and tests, all drawn from one SplitMix64 counter RNG so that every
consumer sees bit-identical data.

Default seed 0x484F4E4549 ("HONEI"); every array has its own stream id.
Uniform doubles are (bits >> 11) * 2^-53 mapped to [lo, hi).
"""
from __future__ import annotations

import numpy as np

SEED = 0x484F4E4549
GAMMA = np.uint64(0x9E3779B97F4A7C15)

# stream ids
S_RHS = 1
S_X = 2
S_MATA = 3
S_MATB_LEN = 4
S_MATB_COL = 5
S_MATB_VAL = 6
S_KTILE = 7


def _mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def rng_bits(seed, stream, index):
    """SplitMix64 output number `index` of the stream keyed by (seed, stream)."""
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) + GAMMA * np.uint64(stream + 1))
        idx = np.asarray(index, np.uint64)
        return _mix64(key + GAMMA * (idx + np.uint64(1)))


def rng_uniform(seed, stream, index, lo=0.0, hi=1.0):
    u = (rng_bits(seed, stream, index) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def rng_normal(seed, stream, index):
    """N(0,1) by Box-Muller from the counters 2i and 2i+1."""
    idx = np.asarray(index, np.uint64)
    u1 = rng_uniform(seed, stream, idx * np.uint64(2))
    u2 = rng_uniform(seed, stream, idx * np.uint64(2) + np.uint64(1))
    u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def side_of_level(level):
    return (1 << level) + 1


def boundary_mask(side):
    m = np.zeros((side, side), bool)
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = True
    return m.ravel()


def rhs_uniform(n, seed=SEED, stream=S_RHS, zero_boundary=True, dtype=np.float64):
    """b ~ U[-1, 1]; with zero_boundary the Dirichlet nodes of the square grid get 0
    (as in the FEAST .bin fixtures), which keeps the Q1 system symmetric."""
    b = rng_uniform(seed, stream, np.arange(n, dtype=np.uint64), -1.0, 1.0)
    if zero_boundary:
        side = int(round(np.sqrt(n)))
        b[boundary_mask(side)] = 0.0
    return b.astype(dtype)


def poisson_q1_bands(n, dtype=np.float64):
    """FillMatrix<POISSON, DIRICHLET_0> (honei/math/fill_matrix.hh:37-76)."""
    M = int(round(np.sqrt(n)))
    i = np.arange(1, n + 1, dtype=np.int64)
    bnd = (i <= M) | (i > n - M) | (i % M == 1) | (i % M == 0)
    off = np.where(bnd, 0.0, -1.0 / 3.0).astype(dtype)
    dd = np.where(bnd, 1.0, 8.0 / 3.0).astype(dtype)
    return [off.copy(), off.copy(), off.copy(), off.copy(), dd, off.copy(), off.copy(), off.copy(), off.copy()]


def q1_offsets(n):
    m = int(round(np.sqrt(n)))
    return np.array((-m - 1, -m, -m + 1, -1, 0, 1, m - 1, m, m + 1), np.int64)


def q1_to_ell(bands, dtype=None):
    """SparseMatrixELL (threads = 1) of a Q1 banded matrix, as built from a
    SparseMatrix: ascending columns, zeros dropped, stride 32-aligned
    (honei/la/sparse_matrix_ell-impl.hh:78-117)."""
    n = bands[0].size
    dtype = dtype or bands[0].dtype
    offs = q1_offsets(n)
    B = np.stack([np.asarray(b, dtype) for b in bands])
    i = np.arange(n, dtype=np.int64)
    J = i[None, :] + offs[:, None]
    keep = (J >= 0) & (J < n) & (B != 0)
    arl = keep.sum(axis=0).astype(np.uint64)
    ncpr = max(int(arl.max()), 1)
    stride = 32 * ((n + 31) // 32)
    Aj = np.zeros((ncpr, stride), np.uint64)
    Ax = np.zeros((ncpr, stride), dtype)
    slot = np.cumsum(keep, axis=0) - 1       # target slot of each kept band entry
    for b in range(9):
        rows = np.nonzero(keep[b])[0]
        s = slot[b, rows]
        Aj[s, rows] = J[b, rows].astype(np.uint64)
        Ax[s, rows] = B[b, rows]
    return dict(rows=n, cols=n, stride=stride, ncpr=ncpr, threads=1, Aj=Aj.ravel(), Ax=Ax.ravel(), Arl=arl)


def matrix_a_bands(side, seed=SEED, dtype=np.float64):
    """Matrix A: the Q1 Dirichlet stencil with every coefficient multiplied by
    a U[0.5, 1.5] factor (counter = row*9 + band), then made strictly
    diagonally dominant: DD <- DD_scaled + sum |off-diagonal|."""
    n = side * side
    bands = poisson_q1_bands(n, np.float64)
    rows = np.arange(n, dtype=np.uint64)
    out = []
    for b in range(9):
        f = rng_uniform(seed, S_MATA, rows * np.uint64(9) + np.uint64(b), 0.5, 1.5)
        out.append(bands[b] * f)
    off = sum(np.abs(out[b]) for b in range(9) if b != 4)
    out[4] = out[4] + off
    return [o.astype(dtype) for o in out]


def matrix_b_ell(n, seed=SEED, dtype=np.float64, halfwidth=2048, min_len=5, max_len=13):
    """Matrix B: n rows, row length (diagonal included) U{min_len..max_len},
    off-diagonal columns distinct and uniform within +-halfwidth of the
    diagonal (window clipped to [0, n)), values N(0,1), diagonal 1 + sum|off|;
    columns ascending in each row; ELL threads = 1, stride 32-aligned."""
    rows = np.arange(n, dtype=np.int64)
    lens = (min_len + (rng_bits(seed, S_MATB_LEN, rows.astype(np.uint64)) % np.uint64(max_len - min_len + 1))
            ).astype(np.int64)
    K = max_len - 1
    lo = np.maximum(rows - halfwidth, 0)
    hi = np.minimum(rows + halfwidth, n - 1)
    W = hi - lo  # candidates excluding the diagonal
    cols = np.full((n, K), -1, np.int64)
    need = np.arange(K)[None, :] < (lens - 1)[:, None]
    rnd = 0
    todo = need.copy()
    while todo.any():
        r_idx, k_idx = np.nonzero(todo)
        ctr = (r_idx.astype(np.uint64) * np.uint64(K) + k_idx.astype(np.uint64)) * np.uint64(64) + np.uint64(rnd)
        u = (rng_bits(seed, S_MATB_COL, ctr) % W[r_idx].astype(np.uint64)).astype(np.int64)
        c = lo[r_idx] + u
        c = np.where(c >= r_idx, c + 1, c)   # skip the diagonal
        cols[r_idx, k_idx] = c
        # duplicates inside a row: keep the first occurrence, redraw the others
        srt = np.sort(np.where(need, cols, np.iinfo(np.int64).max), axis=1)
        dup_rows = np.nonzero(((srt[:, 1:] == srt[:, :-1]) & (srt[:, 1:] != np.iinfo(np.int64).max)).any(axis=1))[0]
        todo[:] = False
        if dup_rows.size:
            sub = cols[dup_rows]
            for k in range(1, K):
                same = (sub[:, :k] == sub[:, k:k + 1]).any(axis=1) & need[dup_rows, k]
                todo[dup_rows[same], k] = True
        rnd += 1
        if rnd > 60:
            raise RuntimeError("matrix_b_ell: column draw did not settle")
    vals = rng_normal(seed, S_MATB_VAL, (rows.astype(np.uint64)[:, None] * np.uint64(K)
                                        + np.arange(K, dtype=np.uint64)[None, :]))
    vals = np.where(need, vals, 0.0)
    diag = 1.0 + np.abs(vals).sum(axis=1)
    # assemble rows: K off-diagonals + diagonal, sorted by column
    allc = np.concatenate([np.where(need, cols, np.iinfo(np.int64).max), rows[:, None]], axis=1)
    allv = np.concatenate([vals, diag[:, None]], axis=1)
    order = np.argsort(allc, axis=1, kind="stable")
    allc = np.take_along_axis(allc, order, axis=1)
    allv = np.take_along_axis(allv, order, axis=1)
    ncpr = max_len
    stride = 32 * ((n + 31) // 32)
    valid = allc != np.iinfo(np.int64).max
    Aj = np.zeros((ncpr, stride), np.uint64)
    Ax = np.zeros((ncpr, stride), dtype)
    Aj[:, :n] = np.where(valid, allc, 0).T.astype(np.uint64)
    Ax[:, :n] = np.where(valid, allv, 0.0).T.astype(dtype)
    return dict(rows=n, cols=n, stride=stride, ncpr=ncpr, threads=1, Aj=Aj.ravel(), Ax=Ax.ravel(),
                Arl=lens.astype(np.uint64))


# -- variable-coefficient diffusion -div(k grad u) = f on the unit square ----
# Q1 elements, k constant per cell.  Dirichlet (u = 0) on the left and right
# edges, natural Neumann on the bottom and top edges.

DIFFUSION_TILES = 64
# macro border mask (corners 0..3, then bottom, right, top, left edges): 2 =
# Dirichlet, 1 = Neumann
DIFFUSION_MASK = (2, 2, 2, 2, 1, 2, 1, 2)


def ktile_field(seed=SEED, tiles=DIFFUSION_TILES):
    """k = exp(N(0,1)) on a tiles x tiles piecewise-constant tiling of the
    square, tile (ty, tx) from counter ty * tiles + tx of stream S_KTILE."""
    idx = np.arange(tiles * tiles, dtype=np.uint64)
    return np.exp(rng_normal(seed, S_KTILE, idx)).reshape(tiles, tiles)


def cell_coefficients(level, seed=SEED, tiles=DIFFUSION_TILES):
    """k per cell of the 2^level x 2^level cell grid: the tile value where a
    cell lies inside one tile, else the arithmetic mean of the tiles it covers."""
    cells = 1 << level
    kt = ktile_field(seed, tiles)
    if cells >= tiles:
        rep = cells // tiles
        return np.repeat(np.repeat(kt, rep, axis=0), rep, axis=1)
    f = tiles // cells
    return kt.reshape(cells, f, cells, f).mean(axis=(1, 3))


def dirichlet_left_right(side):
    m = np.zeros((side, side), bool)
    m[:, 0] = m[:, -1] = True
    return m.ravel()


def diffusion_q1_bands(level, seed=SEED, tiles=DIFFUSION_TILES, dtype=np.float64, cells=None):
    """The 9 Q1 bands (LL, LD, LU, DL, DD, DU, UL, UD, UU) of -div(k grad u)
    on the side 2^level + 1 grid.  The bilinear element matrix of a square
    cell is k * (2/3 on the diagonal, -1/6 to the two edge neighbours, -1/3
    to the opposite corner).  Rows of Dirichlet nodes are identity rows and
    their columns are eliminated from the other rows (u = 0 there), so the
    matrix is symmetric positive definite.  `cells` overrides the per-cell
    coefficients (2^level x 2^level, row = y)."""
    side = side_of_level(level)
    K = np.zeros((side + 1, side + 1))
    K[1:-1, 1:-1] = cell_coefficients(level, seed, tiles) if cells is None else cells
    SW, SE, NW, NE = K[:-1, :-1], K[:-1, 1:], K[1:, :-1], K[1:, 1:]
    third, sixth = 1.0 / 3.0, 1.0 / 6.0
    B = [-third * SW, -sixth * (SW + SE), -third * SE,
         -sixth * (SW + NW), (2.0 / 3.0) * ((SW + SE) + (NW + NE)), -sixth * (SE + NE),
         -third * NW, -sixth * (NW + NE), -third * NE]
    B = [np.ascontiguousarray(b).ravel() for b in B]
    n = side * side
    dmask = dirichlet_left_right(side)
    offs = q1_offsets(n)
    i = np.arange(n, dtype=np.int64)
    for b in range(9):
        if b == 4:
            continue
        j = np.clip(i + offs[b], 0, n - 1)
        B[b][dmask | dmask[j]] = 0.0
    B[4][dmask] = 1.0
    return [b.astype(dtype) for b in B]


def diffusion_ell(level, seed=SEED, tiles=DIFFUSION_TILES, dtype=np.float64):
    """SparseMatrixELL (threads = 1, ascending columns) of diffusion_q1_bands."""
    return q1_to_ell(diffusion_q1_bands(level, seed, tiles, np.float64), dtype)


def rhs_diffusion(level, j, seed=SEED, stream=S_RHS):
    """Right-hand side j of a batch: U[-1, 1] from seed + j, 0 on the
    Dirichlet (left and right edge) nodes."""
    side = side_of_level(level)
    b = rng_uniform(seed + j, stream, np.arange(side * side, dtype=np.uint64), -1.0, 1.0)
    b[dirichlet_left_right(side)] = 0.0
    return b


def ell_to_csr(E):
    """Blocked CSR with blocksize 1 of an ELL matrix with threads = 1 whose
    rows keep their columns ascending (what q1_to_ell and matrix_b_ell
    build): the slots holding a non-zero value, row by row, as
    SparseMatrixCSR(SparseMatrixELL) stores them; a dict rows/cols/bs/Ar/Aj/Ax."""
    assert E.get("threads", 1) == 1
    n, stride, ncpr = E["rows"], E["stride"], E["ncpr"]
    Aj = np.asarray(E["Aj"]).reshape(ncpr, stride)[:, :n].T
    Ax = np.asarray(E["Ax"]).reshape(ncpr, stride)[:, :n].T
    keep = Ax != 0
    counts = keep.sum(axis=1)
    Ar = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=Ar[1:])
    return dict(rows=n, cols=E.get("cols", n), bs=1, Ar=Ar, Aj=Aj[keep].astype(np.uint64), Ax=Ax[keep],
                nblocks=int(Ar[n]), used=int(Ar[n]))


def csr_spmv_bytes(n, z, s):
    """Algorithmic bytes of y = A x from a CSR matrix with 32-bit offsets and
    columns: z(s + 4) + n(2s + 4)."""
    return z * (s + 4) + n * (2 * s + 4)


def ell_nnz(E):
    return int(np.asarray(E["Arl"], np.uint64).sum()) * int(E.get("threads", 1))


def ell_spmv_bytes(n, z, s):
    """Algorithmic bytes of y = A x with 32-bit indices: z(s+4) + n(2s+4)."""
    return z * (s + 4) + n * (2 * s + 4)


def ell_defect_bytes(n, z, s):
    return z * (s + 4) + n * (3 * s + 4)
