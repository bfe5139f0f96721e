"""The vectorised ELL -> CSR conversion used for the CSR workloads equals the
oracle's SparseMatrixCSR(SparseMatrixELL) restatement (itself pinned against
the compiled reference)."""
import numpy as np

from honei_amd import problems as P


def test_ell_to_csr_matches_oracle(oracle):
    for E in (P.q1_to_ell(P.matrix_a_bands(33)), P.matrix_b_ell(3000), P.diffusion_ell(4),
              P.q1_to_ell(P.matrix_a_bands(17, dtype=np.float32))):
        A = P.ell_to_csr(E)
        B = oracle.csr_from_ell(E, 1)
        for k in ("Ar", "Aj", "Ax", "nblocks"):
            assert np.array_equal(A[k], B[k]), k
        x = np.linspace(-1, 1, E["rows"]).astype(E["Ax"].dtype)
        assert np.array_equal(oracle.csr_product(A, x), oracle.csr_product(B, x))
