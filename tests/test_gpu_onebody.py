"""One-electron integrals on the GPU (qfe_one_body_gpu / qfe_one_body_grad_gpu)
against the reference's Psi4 fixture and against the host path."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_INT1E = [(0, "overlap", 1e-9), (1, "kinetic", 1e-9), (2, "nuclear", 1e-8)]


def _teh2hi():
    """TeH2-HI in def2-SVP, cartesian, unsorted shells (int1e_test.cpp:38-46)."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "int1e_teh2hi.npz"))
    mol = Structure(f["numbers"], f["coords_angstrom"] / 0.52917721092)
    return f, mol, load_basis("def2-svp", mol.numbers, mol.coords, pure=False, sort=False)


@pytest.mark.parametrize("op,name,tol", _INT1E)
def test_one_body_gpu_vs_reference_int1e_fixture(op, name, tol):
    """S, T, V of the GPU kernels against the reference's values (its
    tolerances, int1e_test.cpp:96-99) and against the host path."""
    from qdk_chemistry_amd.eri import one_body, one_body_gpu
    f, mol, b = _teh2hi()
    M = one_body_gpu(b, mol, op)
    err_ref = np.max(np.abs(M - f[name + "_integral"]))
    err_host = np.max(np.abs(M - one_body(b, mol, op)))
    print(f"{name}: GPU vs fixture {err_ref:.2e}, vs host {err_host:.2e}", flush=True)
    assert err_ref < tol
    assert err_host < 1e-12 * max(1.0, np.max(np.abs(M)))


@pytest.mark.parametrize("op,name,tol", _INT1E)
def test_one_body_grad_gpu_vs_reference_int1e_fixture(op, name, tol):
    """sum_mn D_mn dO_mn/dR on the GPU against the reference (int1e_test.cpp:
    101-105) and the host path."""
    from qdk_chemistry_amd.eri import one_body_grad, one_body_grad_gpu
    f, mol, b = _teh2hi()
    g = one_body_grad_gpu(b, mol, op, f["dm"])
    err_ref = np.max(np.abs(g - f[name + "_gradient"]))
    err_host = np.max(np.abs(g - one_body_grad(b, mol, op, f["dm"])))
    print(f"{name} gradient: GPU vs fixture {err_ref:.2e}, vs host {err_host:.2e}", flush=True)
    assert err_ref < tol
    assert err_host < 1e-10


def test_one_body_gpu_matches_host_pure_fshells():
    """Pure d and f shells (ethane, cc-pVTZ): S, T, V and the three gradient
    contractions with a random symmetric weight agree with the host path."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import one_body, one_body_gpu, one_body_grad, one_body_grad_gpu
    from qdk_chemistry_amd.molecule import Structure
    R = np.array([[0.0, 0.0, 1.455], [0.0, 0.0, -1.455], [1.94, 0.0, 2.20], [-0.97, 1.68, 2.20],
                  [-0.97, -1.68, 2.20], [-1.94, 0.0, -2.20], [0.97, 1.68, -2.20], [0.97, -1.68, -2.20]])
    mol = Structure(np.array([6, 6, 1, 1, 1, 1, 1, 1]), R)
    b = load_basis("cc-pvtz", mol.numbers, mol.coords)
    rng = np.random.default_rng(5)
    W = rng.uniform(-1, 1, (b.nbf, b.nbf))
    W = 0.5 * (W + W.T)
    for op in range(3):
        M, Mh = one_body_gpu(b, mol, op), one_body(b, mol, op)
        g, gh = one_body_grad_gpu(b, mol, op, W), one_body_grad(b, mol, op, W)
        print(f"op {op}: max|dM| {np.max(np.abs(M - Mh)):.2e}, max|dg| {np.max(np.abs(g - gh)):.2e}", flush=True)
        assert np.max(np.abs(M - Mh)) < 1e-11
        assert np.max(np.abs(g - gh)) < 1e-9
        assert np.max(np.abs(g.sum(axis=0))) < 1e-9  # translation invariance
