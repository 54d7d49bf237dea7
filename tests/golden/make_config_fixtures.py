"""Oracle values for the benchmark configurations, written once in the build
container (the oracle is slow for the larger systems) into
config_fixtures.json, which the GPU tests read.

    python tests/golden/make_config_fixtures.py [c1] [c2] [c5] [--threads N]

c1  RHF/cc-pVDZ energy of H2O at (0, +-0.757, 0.587) A (oracle SCF).
c2  RHF/cc-pVDZ energy of the seeded (H2O)16 cluster (seed 20261202).
c5  J and K of C20H42/cc-pVTZ (seed 20261220) at the seeded decaying density
    decaying_density(basis, 20261220) of tests/conftest.py: Tr(P J),
    Tr(P K) over the whole matrices and 1000 sampled elements.

Every value comes from oracle/oracle.c (the CPU restatement of the
reference's LIBINT2_DIRECT loop), at eri_threshold = shell_pair_threshold =
1e-12.  Existing entries are kept unless recomputed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (REPO, os.path.join(REPO, "qdk-chemistry_amd"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)
OUT = os.path.join(HERE, "config_fixtures.json")


def c1(threads):
    from oracle.oracle import OracleERI, rhf
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import water_monomer
    mol = water_monomer()
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    r = rhf(mol, b, eri=OracleERI(b, mol, nthreads=threads), conv_energy=1e-12, conv_grad=1e-9)
    return {"system": "H2O C1 geometry, cc-pVDZ", "nbf": b.nbf, "energy": r["energy"], "iterations": r["iterations"]}


def c2(threads):
    from oracle.oracle import OracleERI, rhf
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(16, 20261202)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    r = rhf(mol, b, eri=OracleERI(b, mol, nthreads=threads), conv_energy=1e-11, conv_grad=1e-8)
    return {"system": "(H2O)16 seed 20261202, cc-pVDZ", "nbf": b.nbf, "energy": r["energy"],
            "iterations": r["iterations"]}


def c5(threads):
    from conftest import decaying_density
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import linear_alkane
    mol = linear_alkane(20, 20261220)
    b = load_basis("cc-pvtz", mol.numbers, mol.coords)
    P = decaying_density(b, 20261220)
    o = OracleERI(b, mol, nthreads=threads)
    t = time.time()
    J, K = o.build_JK(P)
    dt = time.time() - t
    rng = np.random.Generator(np.random.PCG64(5))
    idx = rng.integers(0, b.nbf, (1000, 2))
    return {"system": "C20H42 seed 20261220, cc-pVTZ", "nbf": b.nbf,
            "density": "decaying_density(basis, 20261220) of tests/conftest.py",
            "trace_PJ": float(np.sum(P * J)), "trace_PK": float(np.sum(P * K)),
            "quartets": o.last_stats["quartets"], "oracle_seconds": dt, "threads": threads,
            "samples": [[int(i), int(j), float(J[i, j]), float(K[i, j])] for i, j in idx]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", nargs="*", default=["c1", "c2", "c5"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    fns = {"c1": c1, "c2": c2, "c5": c5}
    for w in a.which:
        t = time.time()
        data[w] = fns[w](a.threads)
        print(f"{w}: done in {time.time() - t:.0f} s", flush=True)
        with open(OUT + ".tmp", "w") as f:
            json.dump(data, f, indent=1)
        os.replace(OUT + ".tmp", OUT)


if __name__ == "__main__":
    main()
