"""Runs the C++ template-layer test (tests/cpp/mi355x_TEST.cc): HONEI's
container and operation templates instantiated on tags::GPU::MI355X, as a
HONEI caller would use them, each checked against the oracle."""
import os
import subprocess

import pytest

from honei_fixtures import TESTDATA

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "mi355x_TEST")


def test_template_layer_against_oracle(hip):
    assert os.path.exists(BIN), "tests/cpp/mi355x_TEST not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, TESTDATA], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
    for name in ("ell_product_defect<double>", "q1_and_blas1<float>", "cg", "multigrid_mixed", "lock_semantics",
                 "multigrid_fixed", "multigrid_ell", "mgop_cycles", "ir_mgop", "cg_iters_q1<double>",
                 "cg_iters_q1<float>", "cg_iters_ell<double>", "cg_iters_ell<float>", "cg_var_pmat", "ell_threads_io",
                 "mgop_spai"):
        assert f"PASS {name}" in r.stdout
