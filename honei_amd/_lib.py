"""ctypes binding of libhsolve.so (the gfx950 backend's C-ABI, include/hsolve.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhsolve.so")

u64 = C.c_uint64
i32 = C.c_int32
vp = C.c_void_p
dbl = C.c_double
flt = C.c_float

_lib = None


class HsolveError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hsolve error {code}: {msg}")
        self.code = code


class MGParams(C.Structure):
    """hsolve_mg_params (MGInfo configuration fields)."""
    _fields_ = [
        ("min_level", u64), ("max_level", u64), ("n_max_iter", u64),
        ("initial_zero", i32), ("convergence_check", i32),
        ("tolerance", dbl), ("n_pre_smooth", u64), ("adapt_correction_factor", dbl),
        ("macro_border_mask", u64 * 8), ("max_outer", u64),
    ]


# name -> argtypes (restype is int for every entry point)
_SIGS = {
    "hsolve_device_count": [C.POINTER(C.c_int)],
    "hsolve_get_device": [C.POINTER(C.c_int)],
    "hsolve_set_device": [C.c_int],
    "hsolve_malloc": [C.POINTER(vp), u64],
    "hsolve_malloc_host": [C.POINTER(vp), u64],
    "hsolve_free": [vp],
    "hsolve_free_host": [vp],
    "hsolve_upload": [vp, vp, u64, vp],
    "hsolve_download": [vp, vp, u64, vp],
    "hsolve_copy": [vp, vp, u64, vp],
    "hsolve_fill_zero": [vp, u64, vp],
    "hsolve_fill_float": [vp, flt, u64, vp],
    "hsolve_fill_double": [vp, dbl, u64, vp],
    "hsolve_synchronize": [vp],
    "hsolve_stream_create": [C.POINTER(vp)],
    "hsolve_stream_destroy": [vp],
    "hsolve_event_create": [C.POINTER(vp)],
    "hsolve_event_destroy": [vp],
    "hsolve_event_record": [vp, vp],
    "hsolve_event_elapsed_ms": [vp, vp, C.POINTER(flt)],
    "hsolve_copy_kernel": [vp, vp, u64, vp],
    "hsolve_product_smell_dv_float": [vp, vp, vp, vp, vp, u64, u64, u64, u64, u64, vp],
    "hsolve_product_smell_dv_double": [vp, vp, vp, vp, vp, u64, u64, u64, u64, u64, vp],
    "hsolve_defect_smell_dv_float": [vp, vp, vp, vp, vp, vp, u64, u64, u64, u64, u64, vp],
    "hsolve_defect_smell_dv_double": [vp, vp, vp, vp, vp, vp, u64, u64, u64, u64, u64, vp],
    "hsolve_product_bmdv_q1_float": [vp] * 11 + [u64, u64, vp],
    "hsolve_product_bmdv_q1_double": [vp] * 11 + [u64, u64, vp],
    "hsolve_defect_q1_float": [vp] * 12 + [u64, u64, vp],
    "hsolve_defect_q1_double": [vp] * 12 + [u64, u64, vp],
    "hsolve_dot_product_two_float": [vp, vp, u64, vp, vp],
    "hsolve_dot_product_two_double": [vp, vp, u64, vp, vp],
    "hsolve_norm_l2_one_float": [vp, u64, C.c_int, vp, vp],
    "hsolve_norm_l2_one_double": [vp, u64, C.c_int, vp, vp],
    "hsolve_dot_product_two_float_host": [vp, vp, u64, C.POINTER(flt), vp],
    "hsolve_dot_product_two_double_host": [vp, vp, u64, C.POINTER(dbl), vp],
    "hsolve_norm_l2_one_float_host": [vp, u64, C.c_int, C.POINTER(flt), vp],
    "hsolve_norm_l2_one_double_host": [vp, u64, C.c_int, C.POINTER(dbl), vp],
    "hsolve_scaled_sum_three_float_s": [vp, vp, vp, flt, u64, vp],
    "hsolve_scaled_sum_three_double_s": [vp, vp, vp, dbl, u64, vp],
    "hsolve_scaled_sum_three_float": [vp, vp, vp, u64, vp],
    "hsolve_scaled_sum_three_double": [vp, vp, vp, u64, vp],
    "hsolve_sum_two_float": [vp, vp, u64, vp],
    "hsolve_sum_two_double": [vp, vp, u64, vp],
    "hsolve_scale_one_float": [vp, flt, u64, vp],
    "hsolve_scale_one_double": [vp, dbl, u64, vp],
    "hsolve_difference_two_float": [vp, vp, u64, vp],
    "hsolve_difference_two_double": [vp, vp, u64, vp],
    "hsolve_difference_three_float": [vp, vp, vp, u64, vp],
    "hsolve_difference_three_double": [vp, vp, vp, u64, vp],
    "hsolve_element_product_three_float": [vp, vp, vp, u64, vp],
    "hsolve_element_product_three_double": [vp, vp, vp, u64, vp],
    "hsolve_element_inverse_one_float": [vp, u64, vp],
    "hsolve_element_inverse_one_double": [vp, u64, vp],
    "hsolve_prolongation_float": [vp, u64, vp, u64, vp, vp],
    "hsolve_prolongation_double": [vp, u64, vp, u64, vp, vp],
    "hsolve_restriction_float": [vp, u64, vp, u64, vp, vp],
    "hsolve_restriction_double": [vp, u64, vp, u64, vp, vp],
    "hsolve_apply_dirichlet_float": [vp, u64, vp, vp],
    "hsolve_apply_dirichlet_double": [vp, u64, vp, vp],
    "hsolve_convert_float_double": [vp, vp, u64, vp],
    "hsolve_convert_double_float": [vp, vp, u64, vp],
    "hsolve_fill_poisson_q1_float": [vp, u64, vp],
    "hsolve_fill_poisson_q1_double": [vp, u64, vp],
    "hsolve_jacobi_q1_float": [vp, vp, vp, vp, vp, u64, vp],
    "hsolve_jacobi_q1_double": [vp, vp, vp, vp, vp, u64, vp],
    "hsolve_prolongation_axpy_float": [vp, u64, vp, u64, flt, vp, vp],
    "hsolve_prolongation_axpy_double": [vp, u64, vp, u64, dbl, vp, vp],
    "hsolve_cg_konv_q1_float": [vp, vp, vp, u64, dbl, u64, vp, vp],
    "hsolve_cg_konv_q1_double": [vp, vp, vp, u64, dbl, u64, vp, vp],
    "hsolve_cg_iters_q1_float": [vp, vp, vp, u64, u64, vp, vp],
    "hsolve_cg_iters_q1_double": [vp, vp, vp, u64, u64, vp, vp],
    "hsolve_ell_create": [C.c_int, u64, u64, u64, u64, u64, vp, vp, vp, C.POINTER(vp)],
    "hsolve_ell_destroy": [vp],
    "hsolve_ell_info": [vp] + [C.POINTER(u64)] * 7,
    "hsolve_ell_set_variant": [vp, C.c_int],
    "hsolve_ell_q1_side": [vp, C.POINTER(u64)],
    "hsolve_ell_spmv": [vp, vp, vp, vp, vp],
    "hsolve_ell_spmv_rows": [vp, vp, vp, vp, u64, u64, vp],
    "hsolve_ell_jacobi": [vp, vp, vp, vp, vp, vp],
    "hsolve_cg_konv_ell": [vp, vp, vp, dbl, u64, vp, vp],
    "hsolve_cg_q1": [C.c_int, vp, u64, vp, vp, u64, dbl, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_cg_ell": [vp, vp, vp, u64, dbl, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_pcg_ell": [vp, vp, vp, vp, u64, dbl, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_cg_ell_noalloc": [vp, vp, vp, u64, dbl, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_pcg_ell_pmat": [vp, vp, vp, vp, u64, dbl, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_cg_iters_ell": [vp, vp, vp, u64, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_pcg_jac_ell": [vp, vp, vp, vp, u64, dbl, C.POINTER(u64), C.POINTER(dbl), vp],
    "hsolve_ell_spmm": [vp, u64, vp, vp, vp, vp],
    "hsolve_product_csr_dv_float": [vp, vp, vp, vp, vp, u64, u64, u64, u64, vp],
    "hsolve_product_csr_dv_double": [vp, vp, vp, vp, vp, u64, u64, u64, u64, vp],
    "hsolve_defect_csr_dv_float": [vp, vp, vp, vp, vp, vp, u64, u64, u64, vp],
    "hsolve_defect_csr_dv_double": [vp, vp, vp, vp, vp, vp, u64, u64, u64, vp],
    "hsolve_csr_create": [C.c_int, u64, u64, u64, vp, vp, vp, u64, C.POINTER(vp)],
    "hsolve_csr_destroy": [vp],
    "hsolve_csr_info": [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64),
                        C.POINTER(u64)],
    "hsolve_csr_spmv": [vp, vp, vp, vp, u64, u64, vp],
    "hsolve_csr_set_variant": [vp, C.c_int],
    "hsolve_csr_form": [vp, C.POINTER(C.c_int)],
    "hsolve_cg_ell_multi": [vp, u64, vp, vp, vp, u64, dbl, vp, vp, vp],
    "hsolve_rhs_uniform_multi": [vp, u64, u64, u64, u64, vp, u64, u64, vp],
    "hsolve_mg_create": [C.POINTER(MGParams), vp, vp, vp, C.POINTER(vp)],
    "hsolve_mg_create_poisson": [C.POINTER(MGParams), C.POINTER(vp)],
    "hsolve_mg_destroy": [vp],
    "hsolve_mg_device_bytes": [vp, C.POINTER(u64)],
    "hsolve_mg_set_value_codes": [vp, C.c_int],
    "hsolve_mg_solve": [vp, vp, vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(dbl), C.POINTER(dbl), vp],
    "hsolve_mg_create_ell": [C.POINTER(MGParams), vp, vp, vp, vp, vp, u64, C.POINTER(vp)],
    "hsolve_mg_slots": [vp, C.POINTER(u64)],
    "hsolve_mg_solve_batch": [vp, u64, vp, vp, vp, vp, vp, vp, vp],
    "hsolve_serve_create": [C.c_void_p, C.POINTER(vp)],
    "hsolve_serve_destroy": [vp],
    "hsolve_serve_request_bytes": [vp, u64, C.POINTER(u64)],
    "hsolve_serve_submit": [vp, u64, vp, u64, vp, u64, dbl, C.POINTER(u64)],
    "hsolve_serve_step": [vp, C.POINTER(u64)],
    "hsolve_serve_poll": [vp, vp, u64, C.POINTER(u64)],
    "hsolve_serve_now": [vp, C.POINTER(dbl)],
    "hsolve_serve_get_stats": [vp, vp],
    "hsolve_serve_run": [vp, u64, vp, vp, vp, vp, vp],
}


class EllHost(C.Structure):
    """hsolve_ell_host: an ELL matrix in HONEI's host layout (one lane per row)."""
    _fields_ = [("rows", u64), ("cols", u64), ("stride", u64), ("num_cols_per_row", u64),
                ("Aj", vp), ("Ax", vp), ("Arl", vp)]


class ServeParams(C.Structure):
    """hsolve_serve_params."""
    _fields_ = [
        ("min_level", u64), ("min_served_level", u64), ("max_served_level", u64),
        ("n_max_iter", u64), ("initial_zero", i32), ("convergence_check", i32),
        ("tolerance", dbl), ("n_pre_smooth", u64), ("adapt_correction_factor", dbl),
        ("macro_border_mask", u64 * 8), ("max_outer", u64),
        ("memory_budget", u64), ("max_slots", u64), ("rows_per_tick", u64), ("use_graphs", i32),
    ]


class ServeResult(C.Structure):
    """hsolve_serve_result."""
    _fields_ = [
        ("id", u64), ("tag", u64), ("level", u64), ("status", i32),
        ("outer_iters", u64), ("inner_iters", u64), ("ticks", u64),
        ("initial_defect", dbl), ("final_defect", dbl),
        ("t_submit", dbl), ("t_admit", dbl), ("t_done", dbl),
    ]


class ServeStats(C.Structure):
    """hsolve_serve_stats."""
    _fields_ = [
        ("bytes_in_use", u64), ("bytes_held", u64), ("bytes_shared", u64), ("budget", u64),
        ("submitted", u64), ("admitted", u64), ("completed", u64), ("queued", u64),
        ("running", u64), ("peak_running", u64), ("ticks", u64), ("graph_captures", u64),
    ]


def exported_names():
    return sorted(_SIGS) + ["hsolve_last_error", "hsolve_version"]


def lib():
    """Load libhsolve.so once and attach signatures; raises if not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.hsolve_last_error.restype = C.c_char_p
    L.hsolve_last_error.argtypes = []
    L.hsolve_version.restype = C.c_char_p
    L.hsolve_version.argtypes = []
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(code, what=""):
    if code != 0:
        msg = lib().hsolve_last_error().decode(errors="replace")
        raise HsolveError(code, f"{what}: {msg}" if what else msg)
    return code
