"""f shells (l = 3): J/K against the oracle and the reference's def2-TZVP
water energy.  Skipped when the library was built without l = 3 classes."""
import numpy as np
import pytest

from conftest import random_density

pytestmark = pytest.mark.gpu

CRAWFORD = [[0.0, -0.0757918436, 0.0], [0.866811829, 0.6014357793, 0.0], [-0.866811829, 0.6014357793, 0.0]]


@pytest.fixture(autouse=True)
def _needs_f():
    from qdk_chemistry_amd.eri import compiled_max_l
    if compiled_max_l() < 3:
        pytest.skip("library built without f-shell classes (QFE_MAX_L < 3)")


def _ethane():
    from qdk_chemistry_amd.molecule import Structure
    return Structure.from_angstrom(["C", "C", "H", "H", "H", "H", "H", "H"], [
        [0.0, 0.0, 0.7650], [0.0, 0.0, -0.7650],
        [1.0190, 0.0, 1.1600], [-0.5095, 0.8825, 1.1600], [-0.5095, -0.8825, 1.1600],
        [-1.0190, 0.0, -1.1600], [0.5095, -0.8825, -1.1600], [0.5095, 0.8825, -1.1600]])


@pytest.mark.parametrize("system,basis", [("water", "def2-tzvp"), ("ethane", "cc-pvtz")])
def test_fshell_jk_vs_oracle(system, basis):
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import Structure
    mol = Structure.from_angstrom(["O", "H", "H"], CRAWFORD) if system == "water" else _ethane()
    b = load_basis(basis, mol.numbers, mol.coords)
    assert max(s.l for s in b.shells) == 3
    P = random_density(b.nbf, 17, 0.2)
    J, K = create(b, mol).build_JK(P)
    Jr, Kr = OracleERI(b, mol).build_JK(P)
    print(system, basis, "nbf", b.nbf, "max|dJ| %.2e max|dK| %.2e" % (np.max(np.abs(J - Jr)), np.max(np.abs(K - Kr))),
          flush=True)
    assert np.max(np.abs(J - Jr)) < 1e-10
    assert np.max(np.abs(K - Kr)) < 1e-10


def test_water_def2tzvp_reference_energy():
    # cpp/tests/test_scf.cpp:136-150
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    mol = Structure.from_angstrom(["O", "H", "H"], CRAWFORD)
    b = load_basis("def2-tzvp", mol.numbers, mol.coords)
    r = RHF(mol, b, SCFConfig(convergence_threshold=1e-8)).run()
    print("E =", r["energy"], flush=True)
    assert abs(r["energy"] - (-76.0205776517675)) < 1e-8


def test_fshell_gradient_is_unsupported():
    from qdk_chemistry_amd import _lib
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import Structure
    mol = Structure.from_angstrom(["O", "H", "H"], CRAWFORD)
    b = load_basis("def2-tzvp", mol.numbers, mol.coords)
    with pytest.raises(_lib.QfeError):
        create(b, mol).get_gradients(random_density(b.nbf, 3))
