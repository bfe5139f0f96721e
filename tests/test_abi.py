"""The C-ABI library: it loads without a GPU and exports every entry point
include/hsolve.h declares (no compute call is made here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hsolve.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\s*\*)\s*(hsolve_\w+)\s*\(", text)))


def test_header_declares_the_reference_launcher_set():
    names = set(declared_symbols())
    # one HIP entry point per CUDA launcher of operations.hh on the solver path
    for base in ("product_smell_dv", "defect_smell_dv", "product_bmdv_q1", "defect_q1", "dot_product_two",
                 "norm_l2_one", "scaled_sum_three", "sum_two", "scale_one", "difference_two",
                 "difference_three", "element_product_three", "element_inverse_one", "prolongation",
                 "restriction", "product_csr_dv", "defect_csr_dv"):
        for t in ("float", "double"):
            assert f"hsolve_{base}_{t}" in names, base
    for n in ("hsolve_convert_float_double", "hsolve_convert_double_float", "hsolve_malloc", "hsolve_free",
              "hsolve_upload", "hsolve_download", "hsolve_copy", "hsolve_fill_zero", "hsolve_device_count"):
        assert n in names


def test_library_loads_and_exports_all():
    from honei_amd import LIB_PATH, exported_names
    assert os.path.exists(LIB_PATH), "libhsolve.so not built"
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), f"{name} declared in hsolve.h but not exported"
    # the Python binding covers exactly the declared set
    assert set(exported_names()) == set(declared_symbols())
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (hsolve_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_version_string():
    from honei_amd import lib
    assert b"gfx950" in lib().hsolve_version()


def test_no_cpu_fallback_when_missing(tmp_path, monkeypatch):
    """Without the built library the binding raises instead of computing on the host."""
    import honei_amd._lib as L
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        L.lib()


def test_problem_generators_deterministic(oracle):
    from honei_amd import problems as P
    idx = np.arange(64, dtype=np.uint64)
    bits = P.rng_bits(P.SEED, 5, idx)
    assert all(int(bits[i]) == oracle.lib.or_rng_bits(P.SEED, 5, i) for i in range(64))
    E1 = P.matrix_b_ell(4096)
    E2 = P.matrix_b_ell(4096)
    assert all(np.array_equal(E1[k], E2[k]) for k in ("Aj", "Ax", "Arl"))
    n = E1["rows"]
    Aj = E1["Aj"].reshape(E1["ncpr"], E1["stride"])[:, :n].astype(np.int64)
    Ax = E1["Ax"].reshape(E1["ncpr"], E1["stride"])[:, :n]
    L = E1["Arl"].astype(np.int64)
    assert L.min() >= 5 and L.max() <= 13
    for r in (0, 1, 1000, 2047, 2048, 4095):
        cols = Aj[:L[r], r]
        assert np.all(np.diff(cols) > 0) and np.all(np.abs(cols - r) <= 2048) and r in cols
        d = Ax[:L[r], r][cols == r][0]
        assert d == pytest.approx(1.0 + np.abs(Ax[:L[r], r][cols != r]).sum())


def test_matrix_a_shape():
    from honei_amd import problems as P
    side = 65
    E = P.q1_to_ell(P.matrix_a_bands(side))
    n = side * side
    assert P.ell_nnz(E) == (side - 2) ** 2 * 9 + (4 * side - 4)
    # 4097^2: the figure the byte model uses
    assert (4097 - 2) ** 2 * 9 + (4 * 4097 - 4) == 150_937_609
    assert P.ell_spmv_bytes(16_785_409, 150_937_609, 8) == 2_146_959_488
    Aj = E["Aj"].reshape(E["ncpr"], E["stride"])[:, :n]
    Ax = E["Ax"].reshape(E["ncpr"], E["stride"])[:, :n]
    i = 2 * side + 5
    assert np.array_equal(Aj[:, i], i + P.q1_offsets(n))
    assert Ax[4, i] > np.abs(np.delete(Ax[:, i], 4)).sum()       # strictly dominant
