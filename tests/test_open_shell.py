"""Charged and open-shell SCF drivers (RHF with a charge, UHF, ROHF) against
the reference's energies (cpp/tests/test_scf.cpp: Lithium, Oxygen,
OH_ROHF_DIIS, Oxygen_atom_ROHF_GDM; tolerance 1e-8).  Here on the CPU: the
drivers' host logic over the oracle's two-density J/K and the host
one-electron integrals; tests/test_gpu_open_shell.py runs the same cases
through the GPU engine."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FX = json.load(open(os.path.join(GOLD, "reference_fixtures.json")))["open_shell"]


def case_structure(c):
    from qdk_chemistry_amd.basis import ELEMENTS
    from qdk_chemistry_amd.molecule import Structure
    els, xyz = c["structure"]
    return Structure(np.array([ELEMENTS[e] for e in els]), np.array(xyz, dtype=np.float64) / FX["bohr_to_angstrom"],
                     charge=c["charge"])


def run_case(c, cfg, make_eri, guess):
    """The driver named by the case; O2 singlet starts from the ROHF triplet
    density, which breaks the spatial symmetry that the core and atomic
    guesses keep (they settle in a state 0.5 Eh higher)."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.scf import RHF, ROHF, UHF
    mol = case_structure(c)
    b = load_basis(c["basis"], mol.numbers, mol.coords)
    eri = make_eri(b, mol)
    if c["kind"] == "RHF":
        P0 = None
        if c["name"] == "O2 singlet":
            r3 = ROHF(mol, b, 3, cfg, eri=eri).run(guess=guess)
            P0 = r3["Pa"] + r3["Pb"]
        return RHF(mol, b, cfg, eri=eri).run(P0=P0, guess=guess)
    if c["kind"] == "UHF":
        return UHF(mol, b, c["multiplicity"], cfg, eri=eri).run(guess="rohf:" + guess)
    return ROHF(mol, b, c["multiplicity"], cfg, eri=eri).run(guess=guess)


@pytest.mark.parametrize("c", FX["cases"], ids=[c["name"] for c in FX["cases"]])
def test_open_shell_reference_energy_host_logic(c):
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.scf import SCFConfig
    cfg = SCFConfig(convergence_threshold=1e-8, one_body_on_gpu=False, eigen_on_gpu=False, device_resident=False,
                    max_iterations=200)
    r = run_case(c, cfg, lambda b, m: OracleERI(b, m), "core")
    assert abs(r["energy"] - c["energy"]) < FX["tolerance"], (c["name"], r["energy"])


def test_multiplicity_checks():
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure
    from qdk_chemistry_amd.scf import UHF, SCFConfig
    mol = Structure(np.array([8]), np.zeros((1, 3)))
    b = load_basis("sto-3g", mol.numbers, mol.coords)
    cfg = SCFConfig(one_body_on_gpu=False, eigen_on_gpu=False)
    for m in (0, 2, 10):
        with pytest.raises(ValueError):
            UHF(mol, b, m, cfg, eri=object())
