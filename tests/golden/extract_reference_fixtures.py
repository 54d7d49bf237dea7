"""Copy the reference's RHF gradient fixtures into reference_fixtures.json,
and its one-electron integral fixture into int1e_teh2hi.npz.

Sources (data only): scf/tests/data/h2o.json and water10.json (RHF/HF/6-31g*/
pure energy + gradient, atom-major), geometries from scf/tests/test_common.h
(make_h2o, make_water10, make_TeH2_HI; angstrom with the local bohr factor
0.52917721092); scf/tests/data/h2o_gauxc.json (RHF with density-fitted J);
cpp/tests/test_scf.cpp (charged and open-shell RHF/UHF/ROHF energies of
Li, O2, OH and the O atom, structures from cpp/tests/ut_common.hpp);
cpp/tests/test_mp2.cpp (RMP2 correlation energy of singlet O2 in
cc-pVDZ at the test's default bond length, a PySCF value, with the
tolerance testing::mp2_tolerance of cpp/tests/ut_common.hpp);
scf/tests/data/int1e.json (Psi4 S, T, V of TeH2-HI in
def2-SVP, cartesian, unsorted shells, with the density `dm` and the three
gradients sum dm dO/dR, natoms x 3).  Run in a checkout that holds the
reference; the outputs are committed.
"""
import json
import os
import re

ROOT = "/root/reference/cpp/src/qdk/chemistry/algorithms/microsoft/scf/tests"
TESTS = "/root/reference/cpp/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def coords_from_header(fn_name):
    src = open(os.path.join(ROOT, "test_common.h")).read()
    body = src[src.index(f"inline std::shared_ptr<Molecule> {fn_name}()"):]
    body = body[: body.index("return mol;")]
    nums = re.search(r"atomic_nums = \{([^}]*)\}", body).group(1)
    Z = [int(x) for x in nums.replace("\n", " ").split(",") if x.strip()]
    block = body[body.index("mol->coords = {"):]
    triples = re.findall(r"\{\s*([-\d.]+),\s*([-\d.]+),\s*([-\d.]+)\s*\}", block)
    return Z, [[float(a), float(b), float(c)] for a, b, c in triples[: len(Z)]]


def main():
    out = {"bohr_to_angstrom": 0.52917721092}
    for name, fn, data in (("h2o", "make_h2o", "h2o.json"), ("water10", "make_water10", "water10.json")):
        Z, R = coords_from_header(fn)
        rec = json.load(open(os.path.join(ROOT, "data", data)))["RHF/HF/6-31g*/pure"]
        grad = rec["gradient"]
        if isinstance(grad, str):
            grad = json.loads(grad)
        out[name] = dict(numbers=Z, coords_angstrom=R, basis="6-31g*", pure=True,
                         energy=float(rec["energy"]), gradient=[float(g) for g in grad])
    # RHF with density-fitted J (def2-universal-jfit) and exact K, from the
    # data file the reference's test reads for it (scf_test.cpp:22-25:
    # "h2o_gauxc"; molecule make_h2o)
    h2o = json.load(open(os.path.join(ROOT, "data", "h2o_gauxc.json")))
    dfj = h2o["RHF-DFJ/def2-svp/def2-universal-jfit/pure"]
    assert dfj["mol"] == "h2o" and dfj["dfj"] and dfj["aux_basis"] == "def2-universal-jfit"
    out["h2o_dfj"] = dict(basis="def2-svp", aux_basis="def2-universal-jfit", pure=True,
                          energy=float(dfj["energy"]))
    # RMP2 of singlet O2 (create_o2_hf_setup(1): O at 0 and at bond_length
    # bohr on x, cc-pvdz; test RMP2Energies_CCPVDZ)
    mp2 = open(os.path.join(TESTS, "test_mp2.cpp")).read()
    bond = float(re.search(r"double bond_length = ([\d.]+)", mp2).group(1))
    corr = float(re.search(r"pyscf_rmp2_corr_cc_pvdz = (-?[\d.]+)", mp2).group(1))
    tol = float(re.search(r"mp2_tolerance = ([\d.e+-]+);",
                          open(os.path.join(TESTS, "ut_common.hpp")).read()).group(1))
    out["o2_rmp2"] = dict(numbers=[8, 8], coords_bohr=[[0.0, 0.0, 0.0], [bond, 0.0, 0.0]], basis="cc-pvdz",
                          pure=True, correlation_energy=corr, tolerance=tol)
    # open-shell and charged SCF energies of cpp/tests/test_scf.cpp with the
    # structures of cpp/tests/ut_common.hpp (angstrom; CODATA-2022 bohr)
    scf = open(os.path.join(TESTS, "test_scf.cpp")).read()
    ut = open(os.path.join(TESTS, "ut_common.hpp")).read()

    def test_body(name):
        b = scf[scf.index(f"TEST_F(ScfTest, {name})"):]
        return b[: b.index("\n}\n")]

    def structure(fn):
        b = ut[ut.index(f"inline std::shared_ptr<Structure> {fn}()"):]
        b = b[: b.index("\n}\n")]
        xyz = [[float(v) for v in t] for t in re.findall(r"\{\s*([-\d.]+),\s*([-\d.]+),\s*([-\d.]+)\s*\}", b)]
        els = re.findall(r"Element::(\w+)", b)
        return els, xyz

    def near(body, var):
        return float(re.search(rf"EXPECT_NEAR\({var},\s*(-?[\d.]+)", body).group(1))

    open_shell = []
    li, oxy, oh, o2 = (structure(f) for f in ("create_li_structure", "create_oxygen_structure",
                                              "create_oh_structure", "create_o2_structure"))
    b = test_body("Lithium")
    open_shell.append(dict(name="Li doublet", test="Lithium", structure=li, charge=0, multiplicity=2,
                           basis="def2-svp", kind="UHF", energy=near(b, "E_default")))
    open_shell.append(dict(name="Li+ singlet", test="Lithium", structure=li, charge=1, multiplicity=1,
                           basis="def2-svp", kind="RHF", energy=near(b, "E_li_plus_1")))
    b = test_body("Oxygen")
    open_shell.append(dict(name="O2 singlet", test="Oxygen", structure=o2, charge=0, multiplicity=1,
                           basis="sto-3g", kind="RHF", energy=near(b, "E_singlet")))
    open_shell.append(dict(name="O2 triplet", test="Oxygen", structure=o2, charge=0, multiplicity=3,
                           basis="sto-3g", kind="UHF", energy=near(b, "E_triplet")))
    b = test_body("OH_ROHF_DIIS")
    open_shell.append(dict(name="OH doublet", test="OH_ROHF_DIIS", structure=oh, charge=0, multiplicity=2,
                           basis="sto-3g", kind="ROHF", energy=near(b, "E_doublet")))
    b = test_body("Oxygen_atom_ROHF_GDM")
    open_shell.append(dict(name="O triplet", test="Oxygen_atom_ROHF_GDM", structure=oxy, charge=0,
                           multiplicity=3, basis="cc-pvdz", kind="ROHF", energy=near(b, "E_triplet")))
    tol = float(re.search(r"scf_energy_tolerance = ([\d.e+-]+);", ut).group(1))
    out["open_shell"] = dict(bohr_to_angstrom=0.529177210544, tolerance=tol, cases=open_shell)
    json.dump(out, open(os.path.join(HERE, "reference_fixtures.json"), "w"), indent=1)
    import numpy as np
    rec = json.load(open(os.path.join(ROOT, "data", "int1e.json")))
    assert rec["mol"] == "TeH2-HI" and rec["basis"] == "def2-svp"
    Z, R = coords_from_header("make_TeH2_HI")
    n = int(round(len(rec["overlap_integral"]) ** 0.5))
    arrs = dict(numbers=np.array(Z), coords_angstrom=np.array(R))
    for k in ("overlap_integral", "kinetic_integral", "nuclear_integral", "dm"):
        arrs[k] = np.array(rec[k], dtype=np.float64).reshape(n, n)
    for k in ("overlap_gradient", "kinetic_gradient", "nuclear_gradient"):
        arrs[k] = np.array(rec[k], dtype=np.float64).reshape(len(Z), 3)
    np.savez_compressed(os.path.join(HERE, "int1e_teh2hi.npz"), **arrs)


if __name__ == "__main__":
    main()
