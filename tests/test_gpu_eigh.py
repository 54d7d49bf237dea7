"""Symmetric eigensolver on the GPU (qfe_eigh_gpu, rocSOLVER dsyevd) against
numpy.linalg.eigh (LAPACK)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 7, 300, 1536])
def test_eigh_gpu_matches_lapack(n):
    """Eigenvalues to 1e-12 of the spectral radius; eigenvectors orthonormal
    and reconstructing A to the same level (1536 is the SCF's size at
    (H2O)64/cc-pVDZ)."""
    from qdk_chemistry_amd.eri import eigh_gpu
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = 0.5 * (A + A.T)
    eigh_gpu(A)  # first call creates the solver handle
    t0 = time.perf_counter()
    w, V = eigh_gpu(A)
    t1 = time.perf_counter()
    wr, _ = np.linalg.eigh(A)
    t2 = time.perf_counter()
    scale = max(1.0, np.max(np.abs(wr)))
    print(f"n={n}: GPU {1e3 * (t1 - t0):.1f} ms (host copies included), LAPACK {1e3 * (t2 - t1):.1f} ms, "
          f"max|dw| {np.max(np.abs(w - wr)):.2e}", flush=True)
    assert np.max(np.abs(w - wr)) < 1e-12 * scale * max(1, n) ** 0.5
    assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-11 * max(1, n) ** 0.5
    assert np.max(np.abs(V @ np.diag(w) @ V.T - A)) < 1e-11 * scale * max(1, n) ** 0.5


def test_eigh_gpu_rejects_bad_input():
    from qdk_chemistry_amd._lib import QfeError
    from qdk_chemistry_amd.eri import eigh_gpu
    with pytest.raises(ValueError):
        eigh_gpu(np.zeros((3, 4)))
    with pytest.raises(QfeError):
        eigh_gpu(np.zeros((0, 0)))
