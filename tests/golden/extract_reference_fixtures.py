"""Copy the reference's RHF gradient fixtures into reference_fixtures.json,
and its one-electron integral fixture into int1e_teh2hi.npz.

Sources (data only): scf/tests/data/h2o.json and water10.json (RHF/HF/6-31g*/
pure energy + gradient, atom-major), geometries from scf/tests/test_common.h
(make_h2o, make_water10, make_TeH2_HI; angstrom with the local bohr factor
0.52917721092); scf/tests/data/h2o_gauxc.json (RHF with density-fitted J);
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
