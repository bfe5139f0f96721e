"""Device-resident containers over libhsolve, for Python callers (tests,
benchmarks, the serving loop).  Every operation runs on the GPU through the
C-ABI; nothing here computes on the host."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import EllHost, MGParams, ServeParams, ServeResult, ServeStats, check, lib, u64, vp

_DT = {np.dtype(np.float32): 0, np.dtype(np.float64): 1}


def _sfx(dtype):
    return "double" if np.dtype(dtype) == np.float64 else "float"


class DeviceBuffer:
    """A hipMalloc'ed byte range, freed with the object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = vp()
        check(lib().hsolve_malloc(C.byref(p), u64(max(self.nbytes, 1))), "malloc")
        self.ptr = p.value

    def free(self):
        if getattr(self, "ptr", None):
            lib().hsolve_free(vp(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Span:
    """Part of another buffer (keeps the owner alive, frees nothing)."""

    def __init__(self, owner, ptr):
        self.owner = owner
        self.ptr = ptr


class DeviceVector:
    """DenseVector<DT> kept on the device."""

    def __init__(self, size: int, dtype=np.float64, data=None):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.buf = DeviceBuffer(self.size * self.dtype.itemsize)
        if data is not None:
            self.upload(data)

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        return cls(arr.size, arr.dtype, arr)

    @classmethod
    def zeros(cls, size, dtype=np.float64):
        v = cls(size, dtype)
        v.fill(0)
        return v

    def view(self, offset, size):
        """Elements [offset, offset + size) as a vector sharing this one's memory."""
        offset, size = int(offset), int(size)
        assert 0 <= offset and offset + size <= self.size
        v = DeviceVector.__new__(DeviceVector)
        v.dtype, v.size = self.dtype, size
        v.buf = _Span(self.buf, self.buf.ptr + offset * self.dtype.itemsize)
        return v

    @property
    def ptr(self):
        return self.buf.ptr

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, self.dtype)
        assert arr.size == self.size
        check(lib().hsolve_upload(arr.ctypes.data_as(vp), vp(self.ptr), u64(arr.nbytes), vp(stream)), "upload")

    def get(self, stream=None):
        out = np.empty(self.size, self.dtype)
        check(lib().hsolve_download(vp(self.ptr), out.ctypes.data_as(vp), u64(out.nbytes), vp(stream)), "download")
        return out

    def fill(self, value, stream=None):
        if self.dtype == np.float64:
            check(lib().hsolve_fill_double(vp(self.ptr), float(value), u64(self.size), vp(stream)), "fill")
        else:
            check(lib().hsolve_fill_float(vp(self.ptr), float(value), u64(self.size), vp(stream)), "fill")

    def copy_from(self, other, stream=None):
        assert other.size == self.size and other.dtype == self.dtype
        check(lib().hsolve_copy(vp(other.ptr), vp(self.ptr), u64(self.size * self.dtype.itemsize), vp(stream)),
              "copy")


class Q1Matrix:
    """BandedMatrixQx<Q1Type, DT>: nine device bands LL LD LU DL DD DU UL UD UU."""

    def __init__(self, bands):
        bands = [np.ascontiguousarray(b) for b in bands]
        assert len(bands) == 9
        self.dtype = bands[0].dtype
        self.size = bands[0].size
        self.root = int(round(np.sqrt(self.size)))
        self.bands = [DeviceVector.from_numpy(b) for b in bands]
        self.ptrs = (vp * 9)(*[b.ptr for b in self.bands])

    def band_ptrs(self):
        return [vp(b.ptr) for b in self.bands]


def fill_poisson_q1(A: Q1Matrix, stream=None):
    """FillMatrix<Tag, POISSON, DIRICHLET_0>::value(A) on A's device bands."""
    fn = lib().hsolve_fill_poisson_q1_double if A.dtype == np.float64 else lib().hsolve_fill_poisson_q1_float
    check(fn(A.ptrs, u64(A.size), vp(stream)), "fill_poisson_q1")


class EllMatrix:
    """SparseMatrixELL<DT> uploaded into a device handle (uint32 columns)."""

    def __init__(self, rows, columns, stride, num_cols_per_row, threads, Aj, Ax, Arl=None):
        Aj = np.ascontiguousarray(Aj, np.uint64)
        Ax = np.ascontiguousarray(Ax)
        self.dtype = Ax.dtype
        self.rows, self.columns = int(rows), int(columns)
        h = vp()
        arl_p = None
        if Arl is not None:
            self._arl = np.ascontiguousarray(Arl, np.uint64)
            arl_p = self._arl.ctypes.data_as(vp)
        check(lib().hsolve_ell_create(_DT[self.dtype], u64(rows), u64(columns), u64(stride),
                                      u64(num_cols_per_row), u64(threads), Aj.ctypes.data_as(vp),
                                      Ax.ctypes.data_as(vp), arl_p, C.byref(h)), "ell_create")
        self.handle = h.value
        vals = [u64() for _ in range(7)]
        check(lib().hsolve_ell_info(vp(self.handle), *[C.byref(v) for v in vals]), "ell_info")
        (self.rows, self.columns, self.stride, self.ncpr, self.threads, self.nnz,
         self.device_bytes) = (v.value for v in vals)

    @classmethod
    def from_dict(cls, E, columns=None):
        return cls(E["rows"], columns if columns is not None else E.get("cols", E["rows"]), E["stride"],
                   E["ncpr"], E.get("threads", 1), E["Aj"], E["Ax"], E.get("Arl"))

    def set_variant(self, variant: int):
        check(lib().hsolve_ell_set_variant(vp(self.handle), int(variant)), "ell_set_variant")

    @property
    def q1_side(self):
        """Grid side when the matrix was recognised as a 9-point structured matrix, else 0."""
        v = u64()
        check(lib().hsolve_ell_q1_side(vp(self.handle), C.byref(v)), "ell_q1_side")
        return v.value

    def spmv(self, x, y, rhs=None, stream=None):
        check(lib().hsolve_ell_spmv(vp(self.handle), vp(x.ptr), vp(y.ptr), vp(rhs.ptr) if rhs else None,
                                    vp(stream)), "ell_spmv")

    def spmv_rows(self, x, y, row_start, row_end, rhs=None, stream=None):
        """Rows [row_start, row_end) of y = A x (or rhs - A x); the rest of y untouched."""
        check(lib().hsolve_ell_spmv_rows(vp(self.handle), vp(x.ptr), vp(y.ptr), vp(rhs.ptr) if rhs else None,
                                         u64(row_start), u64(row_end), vp(stream)), "ell_spmv_rows")

    def spmm(self, nrhs, X, Y, rhs=None, stream=None):
        """Y = A X (or RHS - A X) on column-blocked multivectors of nrhs columns."""
        check(lib().hsolve_ell_spmm(vp(self.handle), u64(nrhs), vp(X.ptr), vp(Y.ptr), vp(rhs.ptr) if rhs else None,
                                    vp(stream)), "ell_spmm")

    def jacobi(self, rhs, dinv, x_old, x_new, stream=None):
        check(lib().hsolve_ell_jacobi(vp(self.handle), vp(rhs.ptr), vp(dinv.ptr), vp(x_old.ptr), vp(x_new.ptr),
                                      vp(stream)), "ell_jacobi")

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hsolve_ell_destroy(vp(self.handle))
                self.handle = None
        except Exception:
            pass


class CsrMatrix:
    """SparseMatrixCSR<DT> (blocked CSR: Ar block offsets, Aj block start
    columns, Ax `bs` values per block) uploaded into a device handle."""

    def __init__(self, rows, columns, blocksize, Ar, Aj, Ax, used_elements=0):
        Ar = np.ascontiguousarray(Ar, np.uint64)
        Aj = np.ascontiguousarray(Aj, np.uint64)
        Ax = np.ascontiguousarray(Ax)
        assert Ar.size == rows + 1 and Aj.size >= int(Ar[rows]) and Ax.size >= int(Ar[rows]) * blocksize
        self.dtype = Ax.dtype
        h = vp()
        check(lib().hsolve_csr_create(_DT[self.dtype], u64(rows), u64(columns), u64(blocksize),
                                      Ar.ctypes.data_as(vp), Aj.ctypes.data_as(vp), Ax.ctypes.data_as(vp),
                                      u64(used_elements), C.byref(h)), "csr_create")
        self.handle = h.value
        vals = [u64() for _ in range(6)]
        check(lib().hsolve_csr_info(vp(self.handle), *[C.byref(v) for v in vals]), "csr_info")
        (self.rows, self.columns, self.blocksize, self.nblocks, self.used,
         self.device_bytes) = (v.value for v in vals)

    @classmethod
    def from_dict(cls, M):
        return cls(M["rows"], M["cols"], M["bs"], M["Ar"], M["Aj"], M["Ax"], M.get("used", 0))

    def set_variant(self, variant: int):
        """0 automatic, 1 always the CSR kernel."""
        check(lib().hsolve_csr_set_variant(vp(self.handle), int(variant)), "csr_set_variant")

    @property
    def ell_form(self):
        """True when full-range products run on the handle's ELL form."""
        f = C.c_int()
        check(lib().hsolve_csr_form(vp(self.handle), C.byref(f)), "csr_form")
        return bool(f.value)

    def spmv(self, x, y, rhs=None, row_start=0, row_end=0, stream=None):
        """y = A x (or rhs - A x) on rows [row_start, row_end), row_end 0 = all rows."""
        check(lib().hsolve_csr_spmv(vp(self.handle), vp(x.ptr), vp(y.ptr), vp(rhs.ptr) if rhs else None,
                                    u64(row_start), u64(row_end), vp(stream)), "csr_spmv")

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hsolve_csr_destroy(vp(self.handle))
                self.handle = None
        except Exception:
            pass


class Event:
    """hipEvent_t for timing work enqueued on a stream."""

    def __init__(self):
        e = vp()
        check(lib().hsolve_event_create(C.byref(e)), "event_create")
        self.ev = e.value

    def record(self, stream=None):
        check(lib().hsolve_event_record(vp(self.ev), vp(stream)), "event_record")

    def elapsed_ms(self, stop):
        ms = C.c_float()
        check(lib().hsolve_event_elapsed_ms(vp(self.ev), vp(stop.ev), C.byref(ms)), "event_elapsed")
        return ms.value

    def __del__(self):
        try:
            if getattr(self, "ev", None):
                lib().hsolve_event_destroy(vp(self.ev))
                self.ev = None
        except Exception:
            pass


def synchronize(stream=None):
    check(lib().hsolve_synchronize(vp(stream)), "synchronize")


def copy_kernel(src: DeviceBuffer, dst: DeviceBuffer, nbytes: int, stream=None):
    check(lib().hsolve_copy_kernel(vp(src.ptr), vp(dst.ptr), u64(nbytes), vp(stream)), "copy_kernel")


# ----------------------------------------------------------------- ops
def q1_product(A: Q1Matrix, x: DeviceVector, y: DeviceVector, stream=None):
    f = getattr(lib(), f"hsolve_product_bmdv_q1_{_sfx(A.dtype)}")
    check(f(*A.band_ptrs(), vp(x.ptr), vp(y.ptr), u64(A.size), u64(A.root), vp(stream)), "product_q1")


def q1_defect(A: Q1Matrix, rhs: DeviceVector, x: DeviceVector, y: DeviceVector, stream=None):
    f = getattr(lib(), f"hsolve_defect_q1_{_sfx(A.dtype)}")
    check(f(vp(rhs.ptr), *A.band_ptrs(), vp(x.ptr), vp(y.ptr), u64(A.size), u64(A.root), vp(stream)), "defect_q1")


def q1_jacobi(A: Q1Matrix, rhs, dinv, x_old, x_new, stream=None):
    f = getattr(lib(), f"hsolve_jacobi_q1_{_sfx(A.dtype)}")
    check(f(vp(rhs.ptr), A.ptrs, vp(dinv.ptr), vp(x_old.ptr), vp(x_new.ptr), u64(A.size), vp(stream)), "jacobi_q1")


def dot(a: DeviceVector, b: DeviceVector, stream=None):
    if a.dtype == np.float64:
        out = C.c_double()
        check(lib().hsolve_dot_product_two_double_host(vp(a.ptr), vp(b.ptr), u64(a.size), C.byref(out),
                                                       vp(stream)), "dot")
    else:
        out = C.c_float()
        check(lib().hsolve_dot_product_two_float_host(vp(a.ptr), vp(b.ptr), u64(a.size), C.byref(out),
                                                      vp(stream)), "dot")
    return out.value


def norm(a: DeviceVector, root=False, stream=None):
    if a.dtype == np.float64:
        out = C.c_double()
        check(lib().hsolve_norm_l2_one_double_host(vp(a.ptr), u64(a.size), int(root), C.byref(out), vp(stream)),
              "norm")
    else:
        out = C.c_float()
        check(lib().hsolve_norm_l2_one_float_host(vp(a.ptr), u64(a.size), int(root), C.byref(out), vp(stream)),
              "norm")
    return out.value


def restriction(coarse: DeviceVector, fine: DeviceVector, mask=(2,) * 8, stream=None):
    m = np.asarray(mask, np.uint64)
    f = getattr(lib(), f"hsolve_restriction_{_sfx(fine.dtype)}")
    check(f(vp(coarse.ptr), u64(coarse.size), vp(fine.ptr), u64(fine.size), m.ctypes.data_as(vp), vp(stream)),
          "restriction")


def prolongation(fine: DeviceVector, coarse: DeviceVector, mask=(2,) * 8, stream=None):
    m = np.asarray(mask, np.uint64)
    f = getattr(lib(), f"hsolve_prolongation_{_sfx(fine.dtype)}")
    check(f(vp(fine.ptr), u64(fine.size), vp(coarse.ptr), u64(coarse.size), m.ctypes.data_as(vp), vp(stream)),
          "prolongation")


def prolongation_axpy(fine_x: DeviceVector, coarse: DeviceVector, alpha, mask=(2,) * 8, stream=None):
    m = np.asarray(mask, np.uint64)
    f = getattr(lib(), f"hsolve_prolongation_axpy_{_sfx(fine_x.dtype)}")
    check(f(vp(fine_x.ptr), u64(fine_x.size), vp(coarse.ptr), u64(coarse.size), alpha, m.ctypes.data_as(vp),
            vp(stream)), "prolongation_axpy")


def cg_konv_q1(A: Q1Matrix, rhs: DeviceVector, x: DeviceVector, konv_rad, max_iters=0, stream=None):
    it = DeviceVector(1, np.uint64)
    f = getattr(lib(), f"hsolve_cg_konv_q1_{_sfx(A.dtype)}")
    check(f(A.ptrs, vp(rhs.ptr), vp(x.ptr), u64(A.size), float(konv_rad), u64(max_iters), vp(it.ptr), vp(stream)),
          "cg_konv_q1")
    return int(it.get()[0])


def cg_iters_q1(A: Q1Matrix, rhs: DeviceVector, x: DeviceVector, iters, stream=None):
    """ConjugateGradients<NONE>(Q1, rhs, x, unsigned long iters); returns the steps taken."""
    used = DeviceVector(1, np.uint64)
    f = getattr(lib(), f"hsolve_cg_iters_q1_{_sfx(A.dtype)}")
    check(f(A.ptrs, vp(rhs.ptr), vp(x.ptr), u64(A.size), u64(iters), vp(used.ptr), vp(stream)), "cg_iters_q1")
    return int(used.get()[0])


# ----------------------------------------------------------------- solvers
def cg_q1(A: Q1Matrix, b: DeviceVector, x: DeviceVector, max_iters=10000, eps=1e-8, stream=None, raise_on_cap=True):
    used, defect = u64(), C.c_double()
    rc = lib().hsolve_cg_q1(_DT[A.dtype], A.ptrs, u64(A.size), vp(b.ptr), vp(x.ptr), u64(max_iters), float(eps),
                            C.byref(used), C.byref(defect), vp(stream))
    if rc != 5 or raise_on_cap:
        check(rc, "cg_q1")
    return used.value, defect.value


def cg_ell(A: EllMatrix, b: DeviceVector, x: DeviceVector, max_iters=10000, eps=1e-8, stream=None,
           raise_on_cap=True):
    used, defect = u64(), C.c_double()
    rc = lib().hsolve_cg_ell(vp(A.handle), vp(b.ptr), vp(x.ptr), u64(max_iters), float(eps), C.byref(used),
                             C.byref(defect), vp(stream))
    if rc != 5 or raise_on_cap:
        check(rc, "cg_ell")
    return used.value, defect.value


def pcg_ell(A: EllMatrix, P: DeviceVector, b: DeviceVector, x: DeviceVector, max_iters=10000, eps=1e-8,
            stream=None, raise_on_cap=True):
    used, defect = u64(), C.c_double()
    rc = lib().hsolve_pcg_ell(vp(A.handle), vp(P.ptr), vp(b.ptr), vp(x.ptr), u64(max_iters), float(eps),
                              C.byref(used), C.byref(defect), vp(stream))
    if rc != 5 or raise_on_cap:
        check(rc, "pcg_ell")
    return used.value, defect.value


def pcg_ell_pmat(A: EllMatrix, P: EllMatrix, b: DeviceVector, x: DeviceVector, max_iters=10000, eps=1e-8,
                 stream=None, raise_on_cap=True):
    """CGSolver<VAR> with a SparseMatrixELL preconditioner P (z = P r)."""
    used, defect = u64(), C.c_double()
    rc = lib().hsolve_pcg_ell_pmat(vp(A.handle), vp(P.handle), vp(b.ptr), vp(x.ptr), u64(max_iters), float(eps),
                                   C.byref(used), C.byref(defect), vp(stream))
    if rc != 5 or raise_on_cap:
        check(rc, "pcg_ell_pmat")
    return used.value, defect.value


def cg_iters_ell(A: EllMatrix, rhs: DeviceVector, x: DeviceVector, max_iters, stream=None):
    """ConjugateGradients<NONE>(ELL, rhs, x, unsigned long max_iters); returns
    (iterations, final relative defect)."""
    used, defect = u64(), C.c_double()
    check(lib().hsolve_cg_iters_ell(vp(A.handle), vp(rhs.ptr), vp(x.ptr), u64(max_iters), C.byref(used),
                                    C.byref(defect), vp(stream)), "cg_iters_ell")
    return used.value, defect.value


def pcg_jac_ell(A: EllMatrix, dd_inverted: DeviceVector, b: DeviceVector, x: DeviceVector, max_iters=10000,
                eps=1e-8, stream=None, raise_on_cap=True):
    """ConjugateGradients<JAC> on ELL; returns (iterations, final relative defect)."""
    used, defect = u64(), C.c_double()
    rc = lib().hsolve_pcg_jac_ell(vp(A.handle), vp(dd_inverted.ptr), vp(b.ptr), vp(x.ptr), u64(max_iters),
                                  float(eps), C.byref(used), C.byref(defect), vp(stream))
    if rc != 5 or raise_on_cap:
        check(rc, "pcg_jac_ell")
    return used.value, defect.value


def cg_ell_multi(A: EllMatrix, P, B: DeviceVector, X: DeviceVector, nrhs, max_iters=10000, eps=1e-8, stream=None,
                 raise_on_cap=True):
    """CGSolver<VAR> (P a DeviceVector) or <NONE> (P None) on every column of
    the column-blocked multivector B; returns (iterations, relative defects)."""
    used = np.zeros(nrhs, np.uint64)
    defect = np.zeros(nrhs, np.float64)
    rc = lib().hsolve_cg_ell_multi(vp(A.handle), u64(nrhs), vp(P.ptr) if P is not None else None, vp(B.ptr),
                                   vp(X.ptr), u64(max_iters), float(eps), used.ctypes.data_as(vp),
                                   defect.ctypes.data_as(vp), vp(stream))
    if rc != 5 or raise_on_cap:
        check(rc, "cg_ell_multi")
    return used, defect


def rhs_uniform_multi(b: DeviceVector, side, count, row_stride, col_stride, mask=None, seed=None, stream_id=1,
                      stream=None):
    """Seeded U[-1, 1] right-hand sides, column j from seed + j, zero on the Dirichlet border nodes."""
    from . import problems
    m = None if mask is None else (u64 * 8)(*mask)
    check(lib().hsolve_rhs_uniform_multi(vp(b.ptr), u64(side), u64(count), u64(row_stride), u64(col_stride), m,
                                         u64(problems.SEED if seed is None else seed), u64(stream_id), vp(stream)),
          "rhs_uniform_multi")


def mg_params(min_level, max_level, n_max_iter=2, tolerance=1e-8, n_pre_smooth=2, adapt=1.0, mask=(2,) * 8,
              initial_zero=True, convergence_check=False, max_outer=0):
    return MGParams(min_level, max_level, n_max_iter, int(initial_zero), int(convergence_check), tolerance,
                    n_pre_smooth, adapt, (u64 * 8)(*mask), max_outer)


class MGSolver:
    """Multigrid<MI355X, MI355X, NONE, JAC, CYCLE::V, MIXED> hierarchy, device resident."""

    def __init__(self, params: MGParams, bands64=None, bands32_levels=None, dinv32_levels=None):
        self.params = params
        h = vp()
        if bands64 is None:
            check(lib().hsolve_mg_create_poisson(C.byref(params), C.byref(h)), "mg_create_poisson")
        else:
            maxl = params.max_level
            self._keep = [np.ascontiguousarray(b, np.float64) for b in bands64]
            b64 = (vp * 9)(*[b.ctypes.data for b in self._keep])
            b32 = (vp * (9 * (maxl + 1)))()
            d32 = (vp * (maxl + 1))()
            for L in range(params.min_level, maxl + 1):
                for k in range(9):
                    a = np.ascontiguousarray(bands32_levels[L][k], np.float32)
                    self._keep.append(a)
                    b32[9 * L + k] = a.ctypes.data
                if dinv32_levels is not None and L in dinv32_levels:
                    a = np.ascontiguousarray(dinv32_levels[L], np.float32)
                    self._keep.append(a)
                    d32[L] = a.ctypes.data
            check(lib().hsolve_mg_create(C.byref(params), b64, b32, d32, C.byref(h)), "mg_create")
            self._keep = None
        self.handle = h.value
        nb = u64()
        check(lib().hsolve_mg_device_bytes(vp(self.handle), C.byref(nb)), "mg_device_bytes")
        self.device_bytes = nb.value

    def set_value_codes(self, enable: bool):
        """Read the value-coded bands where the hierarchy has them (default) or the plain bands."""
        check(lib().hsolve_mg_set_value_codes(vp(self.handle), int(bool(enable))), "mg_set_value_codes")

    @classmethod
    def from_ell(cls, params: MGParams, A64, a32, dinv32=None, prol32=None, res32=None, slots=1):
        """Multigrid<..., MIXED> on ELL matrices (hsolve_mg_create_ell).  A64: FP64 ELL dict
        (rows, cols, stride, ncpr, Aj, Ax, Arl); a32 / dinv32 / prol32 / res32: dicts level -> ..."""
        self = cls.__new__(cls)
        self.params = params
        keep = []
        maxl = params.max_level

        def host(E, dtype):
            Aj = np.ascontiguousarray(E["Aj"], np.uint64)
            Ax = np.ascontiguousarray(E["Ax"], dtype)
            Arl = np.ascontiguousarray(E["Arl"], np.uint64)
            keep.extend([Aj, Ax, Arl])
            return EllHost(E["rows"], E.get("cols", E["rows"]), E["stride"], E["ncpr"], Aj.ctypes.data,
                           Ax.ctypes.data, Arl.ctypes.data)

        def levels(d):
            arr = (EllHost * (maxl + 1))()
            for L, E in d.items():
                arr[L] = host(E, np.float32)
            return arr

        a64 = host(A64, np.float64)
        dv = None
        if dinv32 is not None:
            dv = (vp * (maxl + 1))()
            for L, d in dinv32.items():
                d = np.ascontiguousarray(d, np.float32)
                keep.append(d)
                dv[L] = d.ctypes.data
        h = vp()
        check(lib().hsolve_mg_create_ell(C.byref(params), C.byref(a64), levels(a32), dv,
                                         levels(prol32) if prol32 else None, levels(res32) if res32 else None,
                                         u64(slots), C.byref(h)), "mg_create_ell")
        self.handle = h.value
        nb = u64()
        check(lib().hsolve_mg_device_bytes(vp(self.handle), C.byref(nb)), "mg_device_bytes")
        self.device_bytes = nb.value
        return self

    def solve_batch(self, bs, xs, stream=None, raise_on_cap=True):
        """Solves bs[i] into xs[i] in the same launches; returns one dict per solve."""
        k = len(bs)
        b = (vp * k)(*[v.ptr for v in bs])
        x = (vp * k)(*[v.ptr for v in xs])
        o, i = (u64 * k)(), (u64 * k)()
        fd, idf = (C.c_double * k)(), (C.c_double * k)()
        rc = lib().hsolve_mg_solve_batch(vp(self.handle), u64(k), b, x, o, i, fd, idf, vp(stream))
        if rc != 5 or raise_on_cap:
            check(rc, "mg_solve_batch")
        return [dict(outer=o[j], inner=i[j], final_defect=fd[j], initial_defect=idf[j], status=rc) for j in range(k)]

    def solve(self, b: DeviceVector, x: DeviceVector, stream=None, raise_on_cap=True):
        o, i, fd, idf = u64(), u64(), C.c_double(), C.c_double()
        rc = lib().hsolve_mg_solve(vp(self.handle), vp(b.ptr), vp(x.ptr), C.byref(o), C.byref(i), C.byref(fd),
                                   C.byref(idf), vp(stream))
        if rc != 5 or raise_on_cap:
            check(rc, "mg_solve")
        return dict(outer=o.value, inner=i.value, final_defect=fd.value, initial_defect=idf.value, status=rc)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hsolve_mg_destroy(vp(self.handle))
                self.handle = None
        except Exception:
            pass


def serve_params(min_level=2, min_served_level=6, max_served_level=11, n_max_iter=2, tolerance=1e-8,
                 n_pre_smooth=2, adapt=1.0, mask=(2,) * 8, initial_zero=True, convergence_check=False,
                 max_outer=0, memory_budget=16 << 30, max_slots=64, rows_per_tick=1 << 24, use_graphs=True):
    return ServeParams(min_level, min_served_level, max_served_level, n_max_iter, int(initial_zero),
                       int(convergence_check), tolerance, n_pre_smooth, adapt, (u64 * 8)(*mask), max_outer,
                       memory_budget, max_slots, rows_per_tick, int(use_graphs))


def _result_dict(r):
    return dict(id=r.id, tag=r.tag, level=r.level, status=r.status, outer=r.outer_iters, inner=r.inner_iters,
                ticks=r.ticks, initial_defect=r.initial_defect, final_defect=r.final_defect,
                t_submit=r.t_submit, t_admit=r.t_admit, t_done=r.t_done)


class Server:
    """Request server for Poisson Q1 MG-IR solves of mixed sizes (hsolve_serve_*)."""

    def __init__(self, params: ServeParams):
        self.params = params
        h = vp()
        check(lib().hsolve_serve_create(C.byref(params), C.byref(h)), "serve_create")
        self.handle = h.value

    def request_bytes(self, level):
        b = u64()
        check(lib().hsolve_serve_request_bytes(vp(self.handle), u64(level), C.byref(b)), "serve_request_bytes")
        return b.value

    def submit(self, level, b: DeviceVector = None, seed=0, x: DeviceVector = None, tag=0, t_submit=-1.0):
        i = u64()
        check(lib().hsolve_serve_submit(vp(self.handle), u64(level), vp(b.ptr if b is not None else None),
                                        u64(seed), vp(x.ptr if x is not None else None), u64(tag),
                                        C.c_double(t_submit), C.byref(i)), "serve_submit")
        return i.value

    def step(self):
        busy = u64()
        check(lib().hsolve_serve_step(vp(self.handle), C.byref(busy)), "serve_step")
        return busy.value

    def poll(self, max_results=256):
        out = (ServeResult * max_results)()
        n = u64()
        check(lib().hsolve_serve_poll(vp(self.handle), out, u64(max_results), C.byref(n)), "serve_poll")
        return [_result_dict(out[k]) for k in range(n.value)]

    def stats(self):
        s = ServeStats()
        check(lib().hsolve_serve_get_stats(vp(self.handle), C.byref(s)), "serve_get_stats")
        return {name: getattr(s, name) for name, _ in ServeStats._fields_}

    def drain(self, expected):
        """Steps until `expected` results have been polled; returns them by id."""
        got = {}
        while len(got) < expected:
            self.step()
            for r in self.poll():
                got[r["id"]] = r
        return got

    def run(self, arrivals, levels, seeds, outputs=None):
        """Replays an arrival trace in real time (hsolve_serve_run); returns one dict per request."""
        n = len(arrivals)
        arr = np.ascontiguousarray(arrivals, np.float64)
        lv = np.ascontiguousarray(levels, np.uint64)
        sd = np.ascontiguousarray(seeds, np.uint64)
        xs = None
        if outputs is not None:
            xs = (vp * n)(*[o.ptr if o is not None else None for o in outputs])
        res = (ServeResult * max(n, 1))()
        check(lib().hsolve_serve_run(vp(self.handle), u64(n), arr.ctypes.data_as(vp), lv.ctypes.data_as(vp),
                                     sd.ctypes.data_as(vp), xs, res), "serve_run")
        return [_result_dict(res[k]) for k in range(n)]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().hsolve_serve_destroy(vp(self.handle))
                self.handle = None
        except Exception:
            pass
