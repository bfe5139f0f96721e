"""ctypes front end of the CPU oracle (liboracle.so) and, where it was built,
of the compiled HONEI reference (oracle/_ref/libhoneiref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py.  The product (honei_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhoneiref.so")

u64 = C.c_uint64
vp = C.c_void_p
BAND_NAMES = ("LL", "LD", "LU", "DL", "DD", "DU", "UL", "UD", "UU")


def _p(a):
    return a.ctypes.data_as(vp) if a is not None else None


def _pp(arrays):
    arr = (vp * len(arrays))(*[a.ctypes.data for a in arrays])
    return arr


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


class MGParams(C.Structure):
    _fields_ = [
        ("min_level", u64), ("max_level", u64), ("n_max_iter", u64),
        ("initial_zero", C.c_int32), ("convergence_check", C.c_int32),
        ("tolerance", C.c_double), ("n_pre_smooth", u64),
        ("adapt_correction_factor", C.c_double), ("macro_border_mask", u64 * 8),
        ("max_outer", u64),
    ]


class MGOpParams(C.Structure):
    """or_mgop_params: operator-list MG configuration (mg.hh MGUtil::configure)."""
    _fields_ = [
        ("cycle", C.c_int32), ("coarse_var", C.c_int32), ("levels", u64), ("min_level", u64),
        ("max_iters", u64), ("max_iters_coarse", u64), ("n_pre_smooth", u64), ("n_post_smooth", u64),
        ("eps_relative", C.c_double),
    ]


class OrEll(C.Structure):
    """or_ell_f64 / or_ell_f32 (same layout; Ax points at the value type)."""
    _fields_ = [("rows", u64), ("cols", u64), ("stride", u64), ("threads", u64),
                ("Aj", vp), ("Ax", vp), ("Arl", vp)]


def _ells(mats):
    arr = (OrEll * max(len(mats), 1))()
    for i, E in enumerate(mats):
        arr[i] = OrEll(E["rows"], E.get("cols", E["rows"]), E["stride"], E["threads"],
                       E["Aj"].ctypes.data, E["Ax"].ctypes.data, E["Arl"].ctypes.data)
    return arr


CYCLES = {"V": 0, "W": 1, "F": 2}


class Oracle:
    """The C restatement of HONEI's tags::CPU path."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run oracle/Makefile)")
        self.lib = C.CDLL(path)
        L = self.lib
        for s, T in (("f64", C.c_double), ("f32", C.c_float)):
            getattr(L, f"or_dot_{s}").restype = T
            getattr(L, f"or_norm_l2_{s}").restype = T
            for fn in ("or_scaled_sum", "or_scaled_sum3", "or_scale"):
                f = getattr(L, f"{fn}_{s}")
                if fn == "or_scaled_sum":
                    f.argtypes = [vp, vp, T, u64]
                elif fn == "or_scaled_sum3":
                    f.argtypes = [vp, vp, vp, T, u64]
                else:
                    f.argtypes = [vp, T, u64]
            getattr(L, f"or_scale_by_double_{s}").argtypes = [vp, C.c_double, u64]
            getattr(L, f"or_cg_none_q1_{s}").argtypes = [vp, vp, vp, u64, u64, T, vp]
            getattr(L, f"or_cg_none_ell_{s}").argtypes = [vp, vp, vp, u64, u64, vp, vp, u64, u64, T, vp]
            getattr(L, f"or_cg_var_ell_{s}").argtypes = [vp, vp, vp, u64, u64, vp, vp, vp, u64, u64, T, vp]
            getattr(L, f"or_pcg_jac_ell_{s}").argtypes = [vp, vp, vp, u64, u64, vp, vp, vp, u64, u64, T, vp]
            getattr(L, f"or_cg_konv_q1_{s}").argtypes = [vp, vp, vp, u64, C.c_double, u64, vp]
            getattr(L, f"or_cg_iters_q1_{s}").argtypes = [vp, vp, vp, u64, u64, vp]
            getattr(L, f"or_cg_iters_ell_{s}").argtypes = [vp, vp, vp, u64, vp]
            getattr(L, f"or_cg_var_ell_pm_{s}").argtypes = [vp, vp, vp, vp, u64, T, vp]
        L.or_rng_bits.restype = u64
        L.or_rng_bits.argtypes = [u64, u64, u64]
        L.or_rng_uniform.restype = C.c_double
        L.or_rng_uniform.argtypes = [u64, u64, u64, C.c_double, C.c_double]
        L.or_mg_mixed_q1.argtypes = [vp, vp, vp, vp, vp, C.POINTER(MGParams), vp, vp, vp, vp]
        for s in ("f64", "f32"):
            getattr(L, f"or_mgop_solve_{s}").argtypes = [C.POINTER(MGOpParams), vp, vp, vp, vp, vp, vp, vp, vp, vp]
            getattr(L, f"or_mgop_solve_pm_{s}").argtypes = [C.POINTER(MGOpParams), vp, vp, vp, vp, vp, vp, vp, vp,
                                                            vp]
        L.or_ir_mgop.argtypes = [C.POINTER(MGOpParams), vp, vp, vp, vp, vp, vp, vp, vp, C.c_double, u64,
                                 vp, vp, vp]

    # -- ELL ---------------------------------------------------------------
    def ell_from_csr(self, rows, rowptr, col, val, threads=1):
        rowptr = np.ascontiguousarray(rowptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint64)
        val = np.ascontiguousarray(val)
        stride, ncpr = u64(), u64()
        self.lib.or_ell_shape(u64(rows), _p(rowptr), u64(threads), C.byref(stride), C.byref(ncpr))
        size = stride.value * ncpr.value
        Aj = np.zeros(size, np.uint64)
        Ax = np.zeros(size, val.dtype)
        Arl = np.zeros(rows, np.uint64)
        getattr(self.lib, f"or_ell_from_csr_{_sfx(val.dtype)}")(
            u64(rows), _p(rowptr), _p(col), _p(val), u64(threads), stride, ncpr, _p(Aj), _p(Ax), _p(Arl))
        return dict(rows=rows, stride=stride.value, ncpr=ncpr.value, threads=threads,
                    Aj=Aj, Ax=Ax, Arl=Arl)

    def ell_row_length(self, rows, stride, threads, Ax):
        Arl = np.zeros(rows, np.uint64)
        getattr(self.lib, f"or_ell_row_length_{_sfx(Ax.dtype)}")(
            u64(rows), u64(stride), u64(threads), u64(Ax.size), _p(Ax), _p(Arl))
        return Arl

    def ell_product(self, E, x, row_start=0, row_end=None):
        rows = E["rows"]
        row_end = rows if row_end is None else row_end
        y = np.zeros(rows, x.dtype)
        getattr(self.lib, f"or_ell_product_{_sfx(x.dtype)}")(
            _p(y), _p(E["Aj"]), _p(E["Ax"]), _p(E["Arl"]), _p(x), u64(row_start), u64(row_end),
            u64(E["stride"]), u64(E["threads"]))
        return y

    def ell_defect(self, E, rhs, x):
        r = np.zeros(E["rows"], x.dtype)
        getattr(self.lib, f"or_ell_defect_{_sfx(x.dtype)}")(
            _p(r), _p(rhs), _p(E["Aj"]), _p(E["Ax"]), _p(E["Arl"]), _p(x), u64(E["rows"]),
            u64(E["stride"]), u64(E["threads"]))
        return r

    # -- CSR ---------------------------------------------------------------
    def csr_from_rows(self, rows, cols, rowptr, col, val, bs=1):
        """SparseMatrixCSR(SparseMatrix) with csr::blocksize bs
        (sparse_matrix_csr-impl.hh:98-136); a dict rows/cols/bs/Ar/Aj/Ax/nblocks/used."""
        rowptr = np.ascontiguousarray(rowptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint64)
        val = np.ascontiguousarray(val)
        cap = max(int(rowptr[rows]), 1)
        Ar = np.zeros(rows + 1, np.uint64)
        Aj = np.zeros(cap, np.uint64)
        Ax = np.zeros(cap * bs, val.dtype)
        nb, used = u64(), u64()
        getattr(self.lib, f"or_csr_from_rows_{_sfx(val.dtype)}")(
            u64(rows), _p(rowptr), _p(col), _p(val), u64(bs), _p(Ar), _p(Aj), _p(Ax), C.byref(nb), C.byref(used))
        n = nb.value
        return dict(rows=rows, cols=cols, bs=bs, Ar=Ar, Aj=Aj[:n].copy(), Ax=Ax[:n * bs].copy(), nblocks=n,
                    used=used.value)

    @staticmethod
    def ell_rows(E, cols=None):
        """SparseMatrix(SparseMatrixELL) (sparse_matrix.hh:157-186): every slot
        with a non-zero value, per row, columns ascending (a later slot with
        the same column overwrites)."""
        rows, stride, thr = E["rows"], E["stride"], E.get("threads", 1)
        Aj, Ax = E["Aj"], E["Ax"]
        rowptr, col, val = [0], [], []
        for r in range(rows):
            d = {}
            for i in range(r * thr, Aj.size, stride):
                for t in range(thr):
                    if Ax[i + t] != 0:
                        d[int(Aj[i + t])] = Ax[i + t]
            for c in sorted(d):
                col.append(c)
                val.append(d[c])
            rowptr.append(len(col))
        return (np.asarray(rowptr, np.uint64), np.asarray(col, np.uint64), np.asarray(val, Ax.dtype))

    def csr_from_ell(self, E, bs=1, cols=None):
        """SparseMatrixCSR(SparseMatrixELL) (sparse_matrix_csr-impl.hh:85-96)."""
        rowptr, col, val = self.ell_rows(E)
        return self.csr_from_rows(E["rows"], cols if cols is not None else E.get("cols", E["rows"]),
                                  rowptr, col, val, bs)

    def csr_product(self, M, x, row_start=0, row_end=None, y=None):
        """Product<tags::CPU>(y, SparseMatrixCSR, x, row_start, row_end) (product.hh:208-245)."""
        y = np.zeros(M["rows"], x.dtype) if y is None else y.copy()
        row_end = M["rows"] if row_end is None else row_end
        getattr(self.lib, f"or_csr_product_{_sfx(x.dtype)}")(
            _p(y), _p(M["Ar"]), _p(M["Aj"]), _p(M["Ax"]), _p(x), u64(M["cols"]), u64(M["bs"]),
            u64(row_start), u64(row_end))
        return y

    def csr_defect(self, M, rhs, x, row_start=0, row_end=None, r=None):
        """Defect<tags::CPU>(r, rhs, SparseMatrixCSR, x) (defect.hh:386-418)."""
        r = np.zeros(M["rows"], x.dtype) if r is None else r.copy()
        row_end = M["rows"] if row_end is None else row_end
        getattr(self.lib, f"or_csr_defect_{_sfx(x.dtype)}")(
            _p(r), _p(rhs), _p(M["Ar"]), _p(M["Aj"]), _p(M["Ax"]), _p(x), u64(M["cols"]), u64(M["bs"]),
            u64(row_start), u64(row_end))
        return r

    # -- Q1 ----------------------------------------------------------------
    def q1_product(self, bands, x):
        y = np.zeros_like(x)
        getattr(self.lib, f"or_q1_product_{_sfx(x.dtype)}")(_p(y), _pp(bands), _p(x), u64(x.size))
        return y

    def q1_defect(self, bands, rhs, x):
        r = np.zeros_like(x)
        getattr(self.lib, f"or_q1_defect_{_sfx(x.dtype)}")(_p(r), _p(rhs), _pp(bands), _p(x), u64(x.size))
        return r

    def fill_poisson_dirichlet0(self, n, dtype=np.float64):
        bands = [np.zeros(n, dtype) for _ in range(9)]
        getattr(self.lib, f"or_fill_poisson_dirichlet0_{_sfx(dtype)}")(_pp(bands), u64(n))
        return bands

    def fill_poisson_dirichlet_neumann(self, n, dtype=np.float64):
        bands = [np.zeros(n, dtype) for _ in range(9)]
        getattr(self.lib, f"or_fill_poisson_dirichlet_neumann_{_sfx(dtype)}")(_pp(bands), u64(n))
        return bands

    def fill_vector_dirichlet_neumann(self, n, dtype=np.float64):
        v = np.zeros(n, dtype)
        getattr(self.lib, f"or_fill_vector_dirichlet_neumann_{_sfx(dtype)}")(_p(v), u64(n))
        return v

    # -- BLAS-1 ------------------------------------------------------------
    def dot(self, x, y):
        return getattr(self.lib, f"or_dot_{_sfx(x.dtype)}")(_p(x), _p(y), u64(x.size))

    def norm(self, x, root=False):
        return getattr(self.lib, f"or_norm_l2_{_sfx(x.dtype)}")(_p(x), u64(x.size), C.c_int32(int(root)))

    def _inplace(self, name, x, *args):
        x = x.copy()
        getattr(self.lib, f"{name}_{_sfx(x.dtype)}")(_p(x), *args)
        return x

    def scaled_sum(self, x, y, b):
        return self._inplace("or_scaled_sum", x, _p(y), b, u64(x.size))

    def scaled_sum3(self, x, y, b):
        r = np.zeros_like(x)
        getattr(self.lib, f"or_scaled_sum3_{_sfx(x.dtype)}")(_p(r), _p(x), _p(y), b, u64(x.size))
        return r

    def scaled_sum_vv(self, a, b, c):
        return self._inplace("or_scaled_sum_vv", a, _p(b), _p(c), u64(a.size))

    def sum(self, x, y):
        return self._inplace("or_sum", x, _p(y), u64(x.size))

    def scale(self, x, a):
        return self._inplace("or_scale", x, a, u64(x.size))

    def scale_by_double(self, x, a):
        return self._inplace("or_scale_by_double", x, C.c_double(a), u64(x.size))

    def difference(self, x, y):
        return self._inplace("or_difference", x, _p(y), u64(x.size))

    def difference3(self, x, y):
        r = np.zeros_like(x)
        getattr(self.lib, f"or_difference3_{_sfx(x.dtype)}")(_p(r), _p(x), _p(y), u64(x.size))
        return r

    def element_product(self, a, b):
        return self._inplace("or_element_product", a, _p(b), u64(a.size))

    def element_product3(self, a, b):
        r = np.zeros_like(a)
        getattr(self.lib, f"or_element_product3_{_sfx(a.dtype)}")(_p(r), _p(a), _p(b), u64(a.size))
        return r

    def element_inverse(self, x):
        return self._inplace("or_element_inverse", x, u64(x.size))

    def convert_d2f(self, x):
        out = np.zeros(x.size, np.float32)
        self.lib.or_convert_d2f(_p(out), _p(np.ascontiguousarray(x, np.float64)), u64(x.size))
        return out

    def convert_f2d(self, x):
        out = np.zeros(x.size, np.float64)
        self.lib.or_convert_f2d(_p(out), _p(np.ascontiguousarray(x, np.float32)), u64(x.size))
        return out

    # -- transfers / smoother ------------------------------------------------
    def restriction(self, fine, mask=(2,) * 8):
        m = np.asarray(mask, np.uint64)
        side_f = int(round(np.sqrt(fine.size)))
        side_c = (side_f - 1) // 2 + 1
        coarse = np.zeros(side_c * side_c, fine.dtype)
        getattr(self.lib, f"or_restriction_{_sfx(fine.dtype)}")(_p(coarse), _p(fine), u64(fine.size), _p(m))
        return coarse

    def prolongation(self, coarse, n_fine, mask=(2,) * 8):
        m = np.asarray(mask, np.uint64)
        fine = np.zeros(n_fine, coarse.dtype)
        getattr(self.lib, f"or_prolongation_{_sfx(coarse.dtype)}")(_p(fine), _p(coarse), u64(n_fine), _p(m))
        return fine

    def apply_dirichlet(self, grid, mask=(2,) * 8):
        m = np.asarray(mask, np.uint64)
        return self._inplace("or_apply_dirichlet", grid, u64(grid.size), _p(m))

    def jacobi_q1(self, x, bands, rhs, dinv, iters):
        x = x.copy()
        tmp = np.zeros_like(x)
        getattr(self.lib, f"or_jacobi_q1_{_sfx(x.dtype)}")(
            _p(x), _pp(bands), _p(rhs), _p(dinv), u64(x.size), u64(iters), _p(tmp))
        return x

    def jacobi_ell(self, x, E, rhs, dinv, iters):
        x = x.copy()
        tmp = np.zeros_like(x)
        getattr(self.lib, f"or_jacobi_ell_{_sfx(x.dtype)}")(
            _p(x), _p(E["Aj"]), _p(E["Ax"]), _p(E["Arl"]), u64(E["stride"]), u64(E["threads"]),
            _p(rhs), _p(dinv), u64(x.size), u64(iters), _p(tmp))
        return x

    # -- solvers -------------------------------------------------------------
    def cg_none_q1(self, bands, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = u64()
        st = getattr(self.lib, f"or_cg_none_q1_{_sfx(b.dtype)}")(
            _pp(bands), _p(b), _p(x), u64(b.size), u64(max_iters), eps, C.byref(used))
        return x, used.value, st

    def cg_none_ell(self, E, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = u64()
        st = getattr(self.lib, f"or_cg_none_ell_{_sfx(b.dtype)}")(
            _p(E["Aj"]), _p(E["Ax"]), _p(E["Arl"]), u64(E["stride"]), u64(E["threads"]),
            _p(b), _p(x), u64(b.size), u64(max_iters), eps, C.byref(used))
        return x, used.value, st

    def cg_var_ell(self, E, P, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = u64()
        st = getattr(self.lib, f"or_cg_var_ell_{_sfx(b.dtype)}")(
            _p(E["Aj"]), _p(E["Ax"]), _p(E["Arl"]), u64(E["stride"]), u64(E["threads"]),
            _p(P), _p(b), _p(x), u64(b.size), u64(max_iters), eps, C.byref(used))
        return x, used.value, st

    def pcg_jac_ell(self, E, dd_inv, b, x0, max_iters=10000, eps=1e-8):
        """ConjugateGradients<JAC> on ELL (conjugate_gradients.hh:1110-1155)."""
        x = x0.copy()
        used = u64()
        st = getattr(self.lib, f"or_pcg_jac_ell_{_sfx(b.dtype)}")(
            _p(E["Aj"]), _p(E["Ax"]), _p(E["Arl"]), u64(E["stride"]), u64(E["threads"]),
            _p(dd_inv), _p(b), _p(x), u64(b.size), u64(max_iters), eps, C.byref(used))
        return x, used.value, st

    def cg_konv_q1(self, bands, rhs, konv_rad, max_iters=0):
        x = np.zeros_like(rhs)
        it = u64()
        st = getattr(self.lib, f"or_cg_konv_q1_{_sfx(rhs.dtype)}")(
            _pp(bands), _p(rhs), _p(x), u64(rhs.size), C.c_double(konv_rad), u64(max_iters), C.byref(it))
        return x, it.value, st

    def cg_iters_q1(self, bands, rhs, iters):
        """ConjugateGradients<NONE>(Q1, rhs, x, unsigned long iters)
        (conjugate_gradients.hh:526-569); x starts at zero."""
        x = np.zeros_like(rhs)
        used = u64()
        st = getattr(self.lib, f"or_cg_iters_q1_{_sfx(rhs.dtype)}")(
            _pp(bands), _p(rhs), _p(x), u64(rhs.size), u64(iters), C.byref(used))
        return x, used.value, st

    def cg_iters_ell(self, E, rhs, x0, max_iters):
        """ConjugateGradients<NONE>(ELL, rhs, x, unsigned long max_iters)
        (conjugate_gradients.hh:571-586, kernel :203-248) from x0."""
        x = np.ascontiguousarray(x0, rhs.dtype).copy()
        used = u64()
        st = getattr(self.lib, f"or_cg_iters_ell_{_sfx(rhs.dtype)}")(
            _ells([E]), _p(rhs), _p(x), u64(max_iters), C.byref(used))
        return x, used.value, st

    def cg_var_ell_pm(self, E, Pm, b, x0, max_iters=10000, eps=1e-8):
        """CGSolver<VAR> with a SparseMatrixELL preconditioner Pm (cg.hh:141-219)."""
        x = np.ascontiguousarray(x0, b.dtype).copy()
        used = u64()
        st = getattr(self.lib, f"or_cg_var_ell_pm_{_sfx(b.dtype)}")(
            _ells([E]), _ells([Pm]), _p(b), _p(x), u64(max_iters), eps, C.byref(used))
        return x, used.value, st

    def mg_mixed_q1(self, A64, b, x0, a32_levels, dinv32_levels, *, min_level, max_level,
                    n_max_iter=2, tolerance=1e-8, n_pre_smooth=2, adapt=1.0, mask=(2,) * 8,
                    max_outer=0):
        """a32_levels / dinv32_levels: dict level -> list of 9 bands / array."""
        x = x0.copy()
        dummy = np.zeros(1, np.float32)
        a_ptrs = (vp * (9 * (max_level + 1)))()
        d_ptrs = (vp * (max_level + 1))()
        for L in range(max_level + 1):
            for k in range(9):
                a_ptrs[9 * L + k] = (a32_levels[L][k] if L in a32_levels else dummy).ctypes.data
            d_ptrs[L] = (dinv32_levels[L] if L in dinv32_levels else dummy).ctypes.data
        p = MGParams(min_level, max_level, n_max_iter, 1, 0, tolerance, n_pre_smooth, adapt,
                     (u64 * 8)(*mask), max_outer)
        outer, inner = u64(), u64()
        fd, idf = C.c_double(), C.c_double()
        st = self.lib.or_mg_mixed_q1(_pp(A64), _p(b), _p(x), a_ptrs, d_ptrs, C.byref(p),
                                     C.byref(outer), C.byref(inner), C.byref(fd), C.byref(idf))
        return dict(x=x, outer=outer.value, inner=inner.value, final_defect=fd.value,
                    initial_defect=idf.value, status=st)

    # -- operator-list MG (mg.hh, ri.hh, ir.hh) --------------------------------
    def mgop_solve(self, data, b, x0, *, cycle="V", coarse_var=False, min_level=1, max_iters=100,
                   max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8):
        """MGSolver over MGCycleCreation<cycle, CGSolver<NONE|VAR>, RISmoother, PROLMAT>.
        data: dict with lists A, P, prol, res (see tests/honei_fixtures.load_mg_data)."""
        dt = np.dtype(b.dtype)
        p = MGOpParams(CYCLES[cycle], int(coarse_var), len(data["A"]), min_level, max_iters, max_iters_coarse,
                       n_pre, n_post, eps)
        x = np.ascontiguousarray(x0, dt).copy()
        used, used_c = u64(), u64()
        fn = C.c_double() if dt == np.float64 else C.c_float()
        st = getattr(self.lib, f"or_mgop_solve_{_sfx(dt)}")(
            C.byref(p), _ells(data["A"]), _pp(data["P"]), _ells(data["prol"]), _ells(data["res"]),
            _p(b), _p(x), C.byref(used), C.byref(used_c), C.byref(fn))
        assert st == 0, st
        return dict(x=x, used=used.value, used_coarse=used_c.value, final_norm=fn.value)

    def mgop_solve_pm(self, data, b, x0, *, cycle="V", coarse_var=False, min_level=1, max_iters=100,
                      max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8):
        """mgop_solve with SparseMatrixELL preconditioners: data["Pm"] (one ELL
        dict per level) in RISmoother and in CGSolver<VAR>."""
        dt = np.dtype(b.dtype)
        p = MGOpParams(CYCLES[cycle], int(coarse_var), len(data["A"]), min_level, max_iters, max_iters_coarse,
                       n_pre, n_post, eps)
        x = np.ascontiguousarray(x0, dt).copy()
        used, used_c = u64(), u64()
        fn = C.c_double() if dt == np.float64 else C.c_float()
        st = getattr(self.lib, f"or_mgop_solve_pm_{_sfx(dt)}")(
            C.byref(p), _ells(data["A"]), _ells(data["Pm"]), _ells(data["prol"]), _ells(data["res"]),
            _p(b), _p(x), C.byref(used), C.byref(used_c), C.byref(fn))
        assert st == 0, st
        return dict(x=x, used=used.value, used_coarse=used_c.value, final_norm=fn.value)

    def ir_mgop(self, A64, b, x0, data32, x_inner0, *, min_level=1, max_iters_inner=8,
                max_iters_coarse=100, n_pre=4, n_post=4, eps_inner=1e-8, eps=1e-8, max_iters=100):
        """IRSolver<Outer, Inner, Norm<l2, true>> with the IR test's inner set-up
        (V cycle, CGSolver<VAR> coarse, RISmoother, PROLMAT)."""
        p = MGOpParams(0, 1, len(data32["A"]), min_level, max_iters_inner, max_iters_coarse, n_pre, n_post,
                       float(np.float32(eps_inner)))
        x = np.ascontiguousarray(x0, np.float64).copy()
        xi = np.ascontiguousarray(x_inner0, np.float32).copy()
        used, iu, cu = u64(), u64(), u64()
        st = self.lib.or_ir_mgop(C.byref(p), _ells([A64]), _p(b), _p(x), _ells(data32["A"]), _pp(data32["P"]),
                                 _ells(data32["prol"]), _ells(data32["res"]), _p(xi), C.c_double(eps),
                                 u64(max_iters), C.byref(used), C.byref(iu), C.byref(cu))
        assert st == 0, st
        return dict(x=x, x_inner=xi, used=used.value, inner_used=iu.value, coarse_used=cu.value)

    # -- RNG -----------------------------------------------------------------
    def rng_uniform(self, seed, stream, index, lo, hi):
        return self.lib.or_rng_uniform(u64(seed), u64(stream), u64(index), lo, hi)


class Reference:
    """HONEI's own tags::CPU templates, compiled from /root/reference by
    oracle/ref/Makefile.  Available only where that build ran."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        L = self.lib
        for s, T in (("f64", C.c_double), ("f32", C.c_float)):
            getattr(L, f"ref_dot_{s}").restype = T
            getattr(L, f"ref_norm_{s}").restype = T
            for fn in ("ref_cg_none_q1", "ref_cg_none_ell", "ref_cg_var_ell"):
                getattr(L, f"{fn}_{s}").restype = u64
            getattr(L, f"ref_cg_none_q1_{s}").argtypes = [vp, vp, vp, u64, u64, T]
            getattr(L, f"ref_cg_none_ell_{s}").argtypes = [u64, u64, u64, u64, vp, vp, vp, vp, u64, T]
            getattr(L, f"ref_cg_var_ell_{s}").argtypes = [u64, u64, u64, u64, vp, vp, vp, vp, vp, u64, T]
            getattr(L, f"ref_pcg_jac_ell_{s}").restype = u64
            getattr(L, f"ref_pcg_jac_ell_{s}").argtypes = [u64, u64, u64, u64, vp, vp, vp, vp, vp, u64, T]
            getattr(L, f"ref_cg_konv_q1_{s}").argtypes = [vp, vp, vp, u64, C.c_double]
            getattr(L, f"ref_cg_iters_q1_{s}").argtypes = [vp, vp, vp, u64, u64]
            getattr(L, f"ref_cg_iters_ell_{s}").argtypes = [u64, u64, u64, u64, vp, vp, vp, vp, u64]
            getattr(L, f"ref_cg_var_ell_pm_{s}").restype = u64
            getattr(L, f"ref_cg_var_ell_pm_{s}").argtypes = [u64, u64, u64, u64, vp, vp, u64, u64, u64, vp, vp, vp,
                                                             vp, u64, T]
        L.ref_mg_mixed_q1.argtypes = [vp, vp, vp, vp, u64, u64, u64, C.c_double, u64, C.c_double,
                                      vp, vp, vp]
        L.ref_mgop_solve_f64.argtypes = [C.c_char_p, C.c_int, C.c_int, u64, u64, u64, u64, u64, u64,
                                         C.c_double, C.c_double, vp, vp, vp]
        L.ref_mgop_solve_pm_f64.argtypes = [C.c_char_p, C.c_int, C.c_int, u64, u64, u64, u64, u64, u64,
                                            C.c_double, C.c_double, C.c_char_p, vp, vp, vp]
        L.ref_ir_mgop.argtypes = [C.c_char_p, u64, u64, u64, u64, u64, u64, C.c_double, C.c_double,
                                  C.c_double, u64, vp, vp, vp, vp]

    def ell_from_csr(self, rows, cols, rowptr, col, val, threads=1):
        rowptr = np.ascontiguousarray(rowptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint64)
        val = np.ascontiguousarray(val)
        stride, ncpr = u64(), u64()
        f = getattr(self.lib, f"ref_ell_from_csr_{_sfx(val.dtype)}")
        f(u64(rows), u64(cols), _p(rowptr), _p(col), _p(val), u64(threads), C.byref(stride),
          C.byref(ncpr), None, None, None)
        size = stride.value * ncpr.value
        Aj = np.zeros(size, np.uint64)
        Ax = np.zeros(size, val.dtype)
        Arl = np.zeros(rows, np.uint64)
        f(u64(rows), u64(cols), _p(rowptr), _p(col), _p(val), u64(threads), C.byref(stride),
          C.byref(ncpr), _p(Aj), _p(Ax), _p(Arl))
        return dict(rows=rows, stride=stride.value, ncpr=ncpr.value, threads=threads,
                    Aj=Aj, Ax=Ax, Arl=Arl)

    def ell_product(self, E, x, cols=None):
        rows = E["rows"]
        cols = x.size if cols is None else cols
        y = np.zeros(rows, x.dtype)
        Arl = np.zeros(rows, np.uint64)
        getattr(self.lib, f"ref_ell_product_{_sfx(x.dtype)}")(
            _p(y), u64(rows), u64(cols), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]),
            _p(E["Aj"]), _p(E["Ax"]), _p(x), _p(Arl))
        return y, Arl

    def ell_defect(self, E, rhs, x):
        r = np.zeros(E["rows"], x.dtype)
        getattr(self.lib, f"ref_ell_defect_{_sfx(x.dtype)}")(
            _p(r), _p(rhs), u64(E["rows"]), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]),
            _p(E["Aj"]), _p(E["Ax"]), _p(x))
        return r

    def q1_product(self, bands, x):
        y = np.zeros_like(x)
        getattr(self.lib, f"ref_q1_product_{_sfx(x.dtype)}")(_p(y), _pp(bands), _p(x), u64(x.size))
        return y

    def q1_defect(self, bands, rhs, x):
        r = np.zeros_like(x)
        getattr(self.lib, f"ref_q1_defect_{_sfx(x.dtype)}")(_p(r), _p(rhs), _pp(bands), _p(x), u64(x.size))
        return r

    def dot(self, x, y):
        return getattr(self.lib, f"ref_dot_{_sfx(x.dtype)}")(_p(x), _p(y), u64(x.size))

    def norm(self, x, root=False):
        return getattr(self.lib, f"ref_norm_{_sfx(x.dtype)}")(_p(x), u64(x.size), C.c_int(int(root)))

    def restriction(self, fine, mask=(2,) * 8):
        m = np.asarray(mask, np.uint64)
        side_f = int(round(np.sqrt(fine.size)))
        side_c = (side_f - 1) // 2 + 1
        coarse = np.zeros(side_c * side_c, fine.dtype)
        getattr(self.lib, f"ref_restriction_{_sfx(fine.dtype)}")(
            _p(coarse), u64(coarse.size), _p(fine), u64(fine.size), _p(m))
        return coarse

    def prolongation(self, coarse, n_fine, mask=(2,) * 8):
        m = np.asarray(mask, np.uint64)
        fine = np.zeros(n_fine, coarse.dtype)
        getattr(self.lib, f"ref_prolongation_{_sfx(coarse.dtype)}")(
            _p(fine), u64(n_fine), _p(coarse), u64(coarse.size), _p(m))
        return fine

    def jacobi_q1(self, x, bands, rhs, dinv, iters):
        x = x.copy()
        getattr(self.lib, f"ref_jacobi_q1_{_sfx(x.dtype)}")(
            _p(x), _pp(bands), _p(rhs), _p(dinv), u64(x.size), u64(iters))
        return x

    def cg_none_q1(self, bands, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = getattr(self.lib, f"ref_cg_none_q1_{_sfx(b.dtype)}")(
            _pp(bands), _p(b), _p(x), u64(b.size), u64(max_iters), eps)
        return x, used

    def cg_none_ell(self, E, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = getattr(self.lib, f"ref_cg_none_ell_{_sfx(b.dtype)}")(
            u64(E["rows"]), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]), _p(E["Aj"]),
            _p(E["Ax"]), _p(b), _p(x), u64(max_iters), eps)
        return x, used

    def cg_var_ell(self, E, P, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = getattr(self.lib, f"ref_cg_var_ell_{_sfx(b.dtype)}")(
            u64(E["rows"]), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]), _p(E["Aj"]),
            _p(E["Ax"]), _p(P), _p(b), _p(x), u64(max_iters), eps)
        return x, used

    def pcg_jac_ell(self, E, dd_inv, b, x0, max_iters=10000, eps=1e-8):
        x = x0.copy()
        used = getattr(self.lib, f"ref_pcg_jac_ell_{_sfx(b.dtype)}")(
            u64(E["rows"]), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]), _p(E["Aj"]),
            _p(E["Ax"]), _p(dd_inv), _p(b), _p(x), u64(max_iters), eps)
        return x, used

    def _csr_result(self, f, args, rows, cols, bs, dtype):
        nb, used, axs = u64(), u64(), u64()
        f(*args, None, None, None, C.byref(nb), C.byref(used), C.byref(axs))
        Ar = np.zeros(rows + 1, np.uint64)
        Aj = np.zeros(max(nb.value, 1), np.uint64)
        Ax = np.zeros(max(axs.value, 1), dtype)
        f(*args, _p(Ar), _p(Aj), _p(Ax), C.byref(nb), C.byref(used), C.byref(axs))
        n = nb.value
        return dict(rows=rows, cols=cols, bs=bs, Ar=Ar, Aj=Aj[:n].copy(), Ax=Ax[:axs.value].copy(), nblocks=n,
                    used=used.value)

    def csr_from_rows(self, rows, cols, rowptr, col, val, bs=1):
        rowptr = np.ascontiguousarray(rowptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint64)
        val = np.ascontiguousarray(val)
        f = getattr(self.lib, f"ref_csr_from_rows_{_sfx(val.dtype)}")
        return self._csr_result(f, (u64(rows), u64(cols), _p(rowptr), _p(col), _p(val), u64(bs)),
                                rows, cols, bs, val.dtype)

    def csr_from_ell(self, E, bs=1, cols=None):
        cols = cols if cols is not None else E.get("cols", E["rows"])
        f = getattr(self.lib, f"ref_csr_from_ell_{_sfx(E['Ax'].dtype)}")
        return self._csr_result(f, (u64(E["rows"]), u64(cols), u64(E["stride"]), u64(E["ncpr"]),
                                    u64(E.get("threads", 1)), _p(E["Aj"]), _p(E["Ax"]), u64(bs)),
                                E["rows"], cols, bs, E["Ax"].dtype)

    def csr_product(self, M, x, row_start=0, row_end=None, y=None):
        y = np.zeros(M["rows"], x.dtype) if y is None else y.copy()
        row_end = M["rows"] if row_end is None else row_end
        getattr(self.lib, f"ref_csr_product_{_sfx(x.dtype)}")(
            _p(y), u64(M["rows"]), u64(M["cols"]), u64(M["bs"]), u64(M["nblocks"]), _p(M["Ar"]), _p(M["Aj"]),
            _p(M["Ax"]), _p(x), u64(row_start), u64(row_end))
        return y

    def csr_defect(self, M, rhs, x, row_start=0, row_end=None, r=None):
        r = np.zeros(M["rows"], x.dtype) if r is None else r.copy()
        row_end = M["rows"] if row_end is None else row_end
        getattr(self.lib, f"ref_csr_defect_{_sfx(x.dtype)}")(
            _p(r), _p(rhs), u64(M["rows"]), u64(M["cols"]), u64(M["bs"]), u64(M["nblocks"]), _p(M["Ar"]),
            _p(M["Aj"]), _p(M["Ax"]), _p(x), u64(row_start), u64(row_end))
        return r

    def cg_konv_q1(self, bands, rhs, konv_rad):
        x = np.zeros_like(rhs)
        getattr(self.lib, f"ref_cg_konv_q1_{_sfx(rhs.dtype)}")(
            _pp(bands), _p(rhs), _p(x), u64(rhs.size), C.c_double(konv_rad))
        return x

    def cg_iters_q1(self, bands, rhs, iters):
        x = np.zeros_like(rhs)
        getattr(self.lib, f"ref_cg_iters_q1_{_sfx(rhs.dtype)}")(_pp(bands), _p(rhs), _p(x), u64(rhs.size),
                                                                u64(iters))
        return x

    def cg_iters_ell(self, E, rhs, x0, max_iters):
        x = np.ascontiguousarray(x0, rhs.dtype).copy()
        getattr(self.lib, f"ref_cg_iters_ell_{_sfx(rhs.dtype)}")(
            u64(E["rows"]), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]), _p(E["Aj"]), _p(E["Ax"]),
            _p(rhs), _p(x), u64(max_iters))
        return x

    def cg_var_ell_pm(self, E, Pm, b, x0, max_iters=10000, eps=1e-8):
        x = np.ascontiguousarray(x0, b.dtype).copy()
        used = getattr(self.lib, f"ref_cg_var_ell_pm_{_sfx(b.dtype)}")(
            u64(E["rows"]), u64(E["stride"]), u64(E["ncpr"]), u64(E["threads"]), _p(E["Aj"]), _p(E["Ax"]),
            u64(Pm["stride"]), u64(Pm["ncpr"]), u64(Pm["threads"]), _p(Pm["Aj"]), _p(Pm["Ax"]), _p(b), _p(x),
            u64(max_iters), eps)
        return x, used

    def mg_mixed_q1(self, A64, b, x0, a32_levels, *, min_level, max_level, n_max_iter=2,
                    tolerance=1e-8, n_pre_smooth=2, adapt=1.0, mask=(2,) * 8):
        x = x0.copy()
        dummy = np.zeros(9, np.float32)
        a_ptrs = (vp * (9 * (max_level + 1)))()
        for L in range(max_level + 1):
            for k in range(9):
                a_ptrs[9 * L + k] = (a32_levels[L][k] if L in a32_levels else dummy).ctypes.data
        m = np.asarray(mask, np.uint64)
        outer, inner = u64(), u64()
        self.lib.ref_mg_mixed_q1(_pp(A64), _p(b), _p(x), a_ptrs, u64(min_level), u64(max_level),
                                 u64(n_max_iter), C.c_double(tolerance), u64(n_pre_smooth),
                                 C.c_double(adapt), _p(m), C.byref(outer), C.byref(inner))
        return dict(x=x, outer=outer.value, inner=inner.value)


def _ref_mgop_solve(self, base, n, *, cycle="V", coarse_var=False, levels=6, min_level=1, max_iters=100,
                    max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8, damping=0.7):
    """The reference's MGUtil::load_data(base, levels, damping, "jac") + MGSolver."""
    base = base if base.endswith("/") else base + "/"
    x = np.zeros(n, np.float64)
    used, used_c = u64(), u64()
    st = self.lib.ref_mgop_solve_f64(base.encode(), CYCLES[cycle], int(coarse_var), levels, min_level, max_iters,
                                     max_iters_coarse, n_pre, n_post, eps, damping, _p(x), C.byref(used),
                                     C.byref(used_c))
    assert st == 0
    return dict(x=x, used=used.value, used_coarse=used_c.value)


def _ref_ir_mgop(self, base, n, *, levels=4, min_level=1, max_iters_inner=8, max_iters_coarse=100, n_pre=4,
                 n_post=4, eps_inner=1e-8, damping=0.7, eps=1e-8, max_iters=100):
    """The IRSolver test's set-up run by the reference (ir_TEST.cc:30-80)."""
    base = base if base.endswith("/") else base + "/"
    x = np.zeros(n, np.float64)
    used, iu, cu = u64(), u64(), u64()
    self.lib.ref_ir_mgop(base.encode(), levels, min_level, max_iters_inner, max_iters_coarse, n_pre, n_post,
                         eps_inner, damping, eps, max_iters, _p(x), C.byref(used), C.byref(iu), C.byref(cu))
    return dict(x=x, used=used.value, inner_used=iu.value, coarse_used=cu.value)


def _ref_mgop_solve_pm(self, base, n, suffix, *, cycle="V", coarse_var=False, levels=6, min_level=1,
                       max_iters=100, max_iters_coarse=100, n_pre=4, n_post=4, eps=1e-8, damping=1.0):
    """MGUtil<..., PreconContType_ = SparseMatrixELL>::load_data(base, levels, damping, suffix) + MGSolver."""
    base = base if base.endswith("/") else base + "/"
    x = np.zeros(n, np.float64)
    used, used_c = u64(), u64()
    st = self.lib.ref_mgop_solve_pm_f64(base.encode(), CYCLES[cycle], int(coarse_var), levels, min_level, max_iters,
                                        max_iters_coarse, n_pre, n_post, eps, damping, suffix.encode(), _p(x),
                                        C.byref(used), C.byref(used_c))
    assert st == 0
    return dict(x=x, used=used.value, used_coarse=used_c.value)


Reference.mgop_solve = _ref_mgop_solve
Reference.mgop_solve_pm = _ref_mgop_solve_pm
Reference.ir_mgop = _ref_ir_mgop


def have_reference() -> bool:
    return os.path.exists(REF_SO)


# -- Multigrid<..., FIXED | MIXED> on Q1 bands or ELL matrices ---------------
class RefEll(C.Structure):
    """ref_ell of oracle/ref/refdrv.cc."""
    _fields_ = [("rows", u64), ("cols", u64), ("stride", u64), ("ncpr", u64), ("Aj", vp), ("Ax", vp)]


def _level_ells(struct, mats, max_level):
    """Array indexed by level (entries below the lowest given level are empty)."""
    arr = (struct * (max_level + 1))()
    for L, E in mats.items():
        if struct is OrEll:
            arr[L] = OrEll(E["rows"], E.get("cols", E["rows"]), E["stride"], E["threads"], E["Aj"].ctypes.data,
                           E["Ax"].ctypes.data, E["Arl"].ctypes.data)
        else:
            arr[L] = RefEll(E["rows"], E.get("cols", E["rows"]), E["stride"], E["ncpr"], E["Aj"].ctypes.data,
                            E["Ax"].ctypes.data)
    return arr


def _level_ptrs(arrs, max_level, per=1):
    """vp array [L * per + k] from a dict level -> array (per == 1) or list of arrays."""
    dummy = np.zeros(16, np.float64)
    out = (vp * ((max_level + 1) * per))()
    for L in range(max_level + 1):
        for k in range(per):
            a = arrs.get(L) if per == 1 else (arrs[L][k] if L in arrs else None)
            out[L * per + k] = (a if a is not None else dummy).ctypes.data
    return out


def _mg_params(min_level, max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check,
               max_outer=0):
    return MGParams(min_level, max_level, n_max_iter, int(initial_zero), int(convergence_check), tolerance, n_pre,
                    adapt, (u64 * 8)(*mask), max_outer)


def _oracle_mg_mixed(self, *, kind, b, x0, A64, a32, dinv32, prol32=None, res32=None, min_level, max_level,
                     n_max_iter=2, tolerance=1e-8, n_pre=2, adapt=1.0, mask=(2,) * 8, initial_zero=True,
                     convergence_check=False):
    """kind "q1": A64 = 9 FP64 bands, a32 = {L: 9 bands}; kind "ell": A64 = ELL, a32 = {L: ELL}.
    prol32 / res32 = {L: ELL} for PROLMAT transfers, else structured."""
    p = _mg_params(min_level, max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check)
    x = np.ascontiguousarray(x0, np.float64).copy()
    o, i, fd, idf = u64(), u64(), C.c_double(), C.c_double()
    L = self.lib
    L.or_mg_mixed.argtypes = [C.c_int] + [vp] * 9 + [C.POINTER(MGParams), vp, vp, vp, vp]
    if kind == "q1":
        args = (0, _pp(A64), None, _p(b), _p(x), _level_ptrs(a32, max_level, 9), None)
    else:
        args = (1, None, _ells([A64]), _p(b), _p(x), None, _level_ells(OrEll, a32, max_level))
    pr = _level_ells(OrEll, prol32, max_level) if prol32 else None
    rs = _level_ells(OrEll, res32, max_level) if res32 else None
    st = L.or_mg_mixed(*args, _level_ptrs(dinv32, max_level), pr, rs, C.byref(p), C.byref(o), C.byref(i),
                       C.byref(fd), C.byref(idf))
    return dict(x=x, outer=o.value, inner=i.value, final_defect=fd.value, initial_defect=idf.value, status=st)


def _oracle_mg_fixed(self, *, kind, b, x0, a, dinv, prol=None, res=None, min_level, max_level, n_max_iter=16,
                     tolerance=1e-8, n_pre=2, adapt=1.0, mask=(2,) * 8, initial_zero=True, convergence_check=False):
    dt = np.dtype(b.dtype)
    s = _sfx(dt)
    p = _mg_params(min_level, max_level, n_max_iter, tolerance, n_pre, adapt, mask, initial_zero, convergence_check)
    x = np.ascontiguousarray(x0, dt).copy()
    it = u64()
    fd = C.c_double() if dt == np.float64 else C.c_float()
    idf = C.c_double() if dt == np.float64 else C.c_float()
    f = getattr(self.lib, f"or_mg_fixed_{s}")
    f.argtypes = [C.c_int] + [vp] * 7 + [C.POINTER(MGParams), vp, vp, vp]
    if kind == "q1":
        aq, ae = _level_ptrs(a, max_level, 9), None
    else:
        aq, ae = None, _level_ells(OrEll, a, max_level)
    pr = _level_ells(OrEll, prol, max_level) if prol else None
    rs = _level_ells(OrEll, res, max_level) if res else None
    st = f(0 if kind == "q1" else 1, aq, ae, _level_ptrs(dinv, max_level), pr, rs, _p(b), _p(x), C.byref(p),
           C.byref(it), C.byref(fd), C.byref(idf))
    return dict(x=x, iters=it.value, final_defect=fd.value, initial_defect=idf.value, status=st)


Oracle.mg_mixed = _oracle_mg_mixed
Oracle.mg_fixed = _oracle_mg_fixed


def _ref_mg_mixed(self, *, kind, b, x0, A64, a32, dinv32, prol32=None, res32=None, min_level, max_level,
                  n_max_iter=2, tolerance=1e-8, n_pre=2, adapt=1.0, mask=(2,) * 8, initial_zero=True,
                  convergence_check=False):
    x = np.ascontiguousarray(x0, np.float64).copy()
    o, i = u64(), u64()
    L = self.lib
    L.ref_mg_mixed.argtypes = [C.c_int, C.c_int] + [vp] * 9 + [u64, u64, u64, C.c_double, u64, C.c_double, vp,
                                                             C.c_int, C.c_int, vp, vp]
    m = np.asarray(mask, np.uint64)
    if kind == "q1":
        a64q, a64e, a32q, a32e = _pp(A64), None, _level_ptrs(a32, max_level, 9), None
    else:
        a64q, a64e = None, _level_ells(RefEll, {0: A64}, 0)
        a32q, a32e = None, _level_ells(RefEll, a32, max_level)
    pr = _level_ells(RefEll, prol32, max_level) if prol32 else None
    rs = _level_ells(RefEll, res32, max_level) if res32 else None
    L.ref_mg_mixed(0 if kind == "q1" else 1, int(bool(prol32)), a64q, a64e, _p(b), _p(x), a32q, a32e, pr, rs,
                   _level_ptrs(dinv32, max_level), min_level, max_level, n_max_iter, tolerance, n_pre, adapt, _p(m),
                   int(initial_zero), int(convergence_check), C.byref(o), C.byref(i))
    return dict(x=x, outer=o.value, inner=i.value)


def _ref_mg_fixed(self, *, kind, b, x0, a, dinv, prol=None, res=None, min_level, max_level, n_max_iter=16,
                  tolerance=1e-8, n_pre=2, adapt=1.0, mask=(2,) * 8, initial_zero=True, convergence_check=False):
    dt = np.dtype(b.dtype)
    x = np.ascontiguousarray(x0, dt).copy()
    f = getattr(self.lib, f"ref_mg_fixed_{_sfx(dt)}")
    f.argtypes = [C.c_int, C.c_int] + [vp] * 7 + [u64, u64, u64, C.c_double, u64, C.c_double, vp, C.c_int, C.c_int]
    m = np.asarray(mask, np.uint64)
    if kind == "q1":
        aq, ae = _level_ptrs(a, max_level, 9), None
    else:
        aq, ae = None, _level_ells(RefEll, a, max_level)
    pr = _level_ells(RefEll, prol, max_level) if prol else None
    rs = _level_ells(RefEll, res, max_level) if res else None
    f(0 if kind == "q1" else 1, int(bool(prol)), aq, ae, pr, rs, _level_ptrs(dinv, max_level), _p(b), _p(x),
      min_level, max_level, n_max_iter, tolerance, n_pre, adapt, _p(m), int(initial_zero), int(convergence_check))
    return dict(x=x)


Reference.mg_mixed = _ref_mg_mixed
Reference.mg_fixed = _ref_mg_fixed
