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


@pytest.mark.parametrize("system,basis", [("water", "def2-tzvp"), ("ethane", "cc-pvtz")])
def test_fshell_eri_gradient_vs_oracle(system, basis):
    """The adjoint kernel on classes with f shells (runtime-l kernel) against
    the oracle's restatement of ERI::get_gradients (libint2_direct.cpp:
    646-890), RHF total density and a spin pair."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import Structure
    mol = Structure.from_angstrom(["O", "H", "H"], CRAWFORD) if system == "water" else _ethane()
    b = load_basis(basis, mol.numbers, mol.coords)
    eng, o = create(b, mol), OracleERI(b, mol)
    P = random_density(b.nbf, 19, 0.2)
    dJ, dK = eng.get_gradients(P)
    rJ, rK = o.get_gradients(P, mode=1)
    print(system, basis, "max|ddJ| %.2e max|ddK| %.2e" % (np.max(np.abs(dJ - rJ)), np.max(np.abs(dK - rK))),
          flush=True)
    assert np.max(np.abs(dJ - rJ)) < 1e-10 and np.max(np.abs(dK - rK)) < 1e-10
    P2 = np.stack([P, random_density(b.nbf, 20, 0.1)])
    dJ, dK = eng.get_gradients(P2)
    rJ, rK = o.get_gradients(P2, mode=1)
    assert np.max(np.abs(dJ - rJ)) < 1e-10 and np.max(np.abs(dK - rK)) < 1e-10


def test_water_def2tzvp_rhf_gradient_vs_central_differences():
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    mol = Structure.from_angstrom(["O", "H", "H"], CRAWFORD)
    cfg = SCFConfig(convergence_threshold=1e-10)
    scf = RHF(mol, load_basis("def2-tzvp", mol.numbers, mol.coords), cfg)
    scf.run()
    g = scf.gradient()
    h = 1e-4
    worst = 0.0
    for a, x in ((0, 1), (1, 0), (1, 1), (2, 0)):
        e = []
        for sgn in (1.0, -1.0):
            R = mol.coords.copy()
            R[a, x] += sgn * h
            m = Structure(mol.numbers, R)
            e.append(RHF(m, load_basis("def2-tzvp", m.numbers, m.coords), cfg).run()["energy"])
        worst = max(worst, abs((e[0] - e[1]) / (2 * h) - g[a, x]))
    print("def2-TZVP water: max |g - fd| = %.2e" % worst, flush=True)
    assert worst <= 1e-6
