"""Quarter and four-index MO transforms on the GPU (qfe_quarter_trans,
qfe_mo_eri): against the oracle on the same inputs, and the reference's RMP2
correlation energy of singlet O2 / cc-pVDZ (cpp/tests/test_mp2.cpp) through
the GPU SCF and the GPU transform."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _case(name, nbasis="cc-pvdz"):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure, water_monomer
    if name == "water":
        mol = water_monomer()
    else:
        rec = json.load(open(os.path.join(GOLD, "reference_fixtures.json")))["o2_rmp2"]
        mol = Structure(np.array(rec["numbers"]), np.array(rec["coords_bohr"], dtype=np.float64))
    return mol, load_basis(nbasis, mol.numbers, mol.coords)


@pytest.mark.parametrize("name,nt", [("water", 7), ("water", 24), ("o2", 11)])
def test_quarter_trans_vs_oracle(name, nt):
    from oracle.oracle import OracleERI, quarter_trans
    from qdk_chemistry_amd.eri import create
    mol, b = _case(name)
    C = np.random.Generator(np.random.PCG64(nt)).uniform(-1.0, 1.0, (b.nbf, nt))
    print(f"quarter_trans {name} nbf {b.nbf} nt {nt}", flush=True)
    got = create(b, mol).quarter_trans(C)
    ref = quarter_trans(OracleERI(b, mol).eri_full(), C)
    assert got.shape == (b.nbf, b.nbf, b.nbf, nt)
    err = np.max(np.abs(got - ref))
    print(f"  max|diff| = {err:.2e}", flush=True)
    assert err < 1e-11


def test_mo_eri_vs_oracle_distinct_coefficients():
    """Four different C: each slot contracts its own AO index."""
    from oracle.oracle import OracleERI, mo_eri
    from qdk_chemistry_amd.eri import MOERI, create
    mol, b = _case("o2")
    rng = np.random.Generator(np.random.PCG64(5))
    Cs = [rng.uniform(-1.0, 1.0, (b.nbf, 9)) for _ in range(4)]
    got = MOERI(create(b, mol)).compute(b.nbf, 9, *Cs)
    ref = mo_eri(OracleERI(b, mol).eri_full(), *Cs)
    err = np.max(np.abs(got - ref))
    print(f"mo_eri O2 nbf {b.nbf} nt 9: max|diff| = {err:.2e}", flush=True)
    assert err < 1e-10


def test_rmp2_o2_vs_reference():
    """GPU RHF, GPU MO transform of the (ov|ov) block, RMP2 correlation
    energy within the reference's tolerance of its value."""
    from oracle.oracle import rmp2_correlation
    from qdk_chemistry_amd.eri import MOERI, create
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    rec = json.load(open(os.path.join(GOLD, "reference_fixtures.json")))["o2_rmp2"]
    mol, b = _case("o2")
    eri = create(b, mol)
    r = RHF(mol, b, SCFConfig(convergence_threshold=1e-10), eri=eri).run()
    no = mol.nelectrons // 2
    C, e = r["C"], r["eps"]
    nv = C.shape[1] - no
    # the (o v | o v) block through the square transform with nt = nmo
    full = MOERI(eri).compute(b.nbf, C.shape[1], C, C, C, C)
    ovov = full[:no, no:, :no, no:]
    corr = rmp2_correlation(ovov, e[:no], e[no:])
    print(f"E_HF {r['energy']:.10f}; RMP2 corr {corr:.12f} vs {rec['correlation_energy']:.12f}", flush=True)
    assert nv > 0
    assert abs(corr - rec["correlation_energy"]) < rec["tolerance"]


def test_mo_transform_errors():
    from qdk_chemistry_amd._lib import QfeError
    from qdk_chemistry_amd.eri import MOERI, create
    mol, b = _case("water")
    eri = create(b, mol)
    C = np.zeros((b.nbf, b.nbf + 1))
    with pytest.raises(QfeError, match="num_molecular_orbitals > num_atomic_orbitals"):
        MOERI(eri).compute(b.nbf, b.nbf + 1, C, C, C, C)
    mol2, tz = _case("water", "def2-tzvp")  # O carries f shells
    C = np.zeros((tz.nbf, 3))
    with pytest.raises(QfeError, match="shells above l = 2"):
        create(tz, mol2).quarter_trans(C)
