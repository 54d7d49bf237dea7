"""Regenerate the mpmath golden vectors used by the CPU tests.

boys.json      F_m(T) = int_0^1 t^2m exp(-T t^2) dt at 40 digits (mpmath
               lower incomplete gamma: gamma(m+1/2, T) / (2 T^(m+1/2))).
rys.json       n-point Rys nodes/weights (Gauss rule of the Boys measure)
               from the 40-digit moments by the Chebyshev algorithm and a
               Jacobi-matrix eigen-decomposition (Golub-Welsch).
structures     seeded benchmark geometries (xyz, angstrom).
"""
import json
import os
import sys

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "qdk-chemistry_amd"))
sys.path.insert(0, os.path.join(HERE, "..", "..", "qdk-chemistry_amd", "csrc"))


def boys():
    mp.mp.dps = 40
    Ts = [0.0, 1e-8, 1e-3, 0.137, 0.5, 1.0, 2.718, 7.3, 12.0, 19.99, 25.0, 33.3, 39.9, 40.1, 47.0, 80.0, 250.0]
    out = []
    for T in Ts:
        for m in range(0, 17):
            if T == 0.0:
                v = mp.mpf(1) / (2 * m + 1)
            else:
                v = mp.gammainc(m + mp.mpf(0.5), 0, T) / (2 * mp.mpf(T) ** (m + mp.mpf(0.5)))
            out.append([m, T, mp.nstr(v, 30)])
    json.dump(out, open(os.path.join(HERE, "boys.json"), "w"), indent=0)


def rys():
    import gen_rys_table as g
    mp.mp.dps = 40
    out = []
    for n in range(1, 8):
        for T in (0.0, 0.3, 1.7, 5.5, 13.25, 27.9, 39.5, 45.1, 60.7, 77.0, 120.0):
            u, w = g.rys_rule(T, n)
            out.append(dict(n=n, T=T, u=[mp.nstr(x, 30) for x in u], w=[mp.nstr(x, 30) for x in w]))
    json.dump(out, open(os.path.join(HERE, "rys.json"), "w"), indent=0)


def structures():
    from qdk_chemistry_amd.molecule import linear_alkane, water_cluster
    for name, mol in (("h2o16_seed20261202", water_cluster(16, 20261202)),
                      ("h2o64_seed20261264", water_cluster(64, 20261264)),
                      ("c20h42_seed20261220", linear_alkane(20, 20261220))):
        with open(os.path.join(HERE, name + ".xyz"), "w") as f:
            f.write(mol.to_xyz(name))


if __name__ == "__main__":
    boys()
    rys()
    structures()
