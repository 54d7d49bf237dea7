"""Basis sets and AO conventions of qdk-chemistry's SCF plugin path.

The order of shells and the normalisation of contraction coefficients fix
which AO index every Fock-matrix element refers to, so they follow the
reference exactly:

* BSE JSON loading, one shell per general-contraction column
  (cpp/src/qdk/chemistry/data/basis_set.cpp:174-212).
* Primitives sorted by descending exponent inside each shell, then all shells
  of the molecule stable-sorted by angular momentum and by descending leading
  exponent (data/basis_set.cpp:120-162, applied at :626-632).
* PSI4-mode normalisation: each primitive normalised as an x^l Gaussian, then
  the contraction renormalised to unit self-overlap
  (algorithms/microsoft/utils.cpp:26-79, core/basis_set.cpp:365-403).
* Pure (solid-harmonic) functions for l >= 2 only, m = -l..l; cartesian
  components in CCA order xx, xy, xz, yy, yz, zz (libint2_util.cpp:15-28).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

_DATA = os.path.join(os.path.dirname(__file__), "data", "basis")

ELEMENTS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9}
SYMBOLS = {v: k for k, v in ELEMENTS.items()}


def _basis_file(name: str) -> str:
    """Basis name -> JSON file (lower case, '*' spelled '_st_' as in
    data/basis_set.cpp normalize_basis_set_name)."""
    norm = name.lower().replace("*", "_st_")
    path = os.path.join(_DATA, norm + ".json")
    if not os.path.exists(path):
        raise ValueError(f"basis set {name} is not supported")
    return path


@dataclass
class Shell:
    atom: int
    l: int
    exps: np.ndarray
    coefs: np.ndarray
    center: np.ndarray

    @property
    def nprim(self) -> int:
        return len(self.exps)


def _double_factorial_table(n: int) -> list:
    # same recurrence as the reference (df[i] = df[i-2] * (i-1)); df[2l] = (2l-1)!!
    df = [1.0] * max(n, 3)
    for i in range(3, n):
        df[i] = df[i - 2] * (i - 1)
    return df


def normalize_psi4(l: int, exps: np.ndarray, coefs: np.ndarray) -> np.ndarray:
    """PSI4-mode normalisation (utils.cpp:26-79)."""
    df = _double_factorial_table(20)
    sqrt_pi3 = math.sqrt(math.pi ** 3)
    c = np.array(coefs, dtype=np.float64).copy()
    for i, a in enumerate(exps):
        if a <= 0:
            raise ValueError("Shell exponents must be positive")
        e2 = 2.0 * a
        e2_am32 = e2 ** (l + 1) * math.sqrt(e2)
        c[i] *= math.sqrt(2.0 ** l * e2_am32 / (sqrt_pi3 * df[2 * l]))
    norm = 0.0
    for i in range(len(exps)):
        for j in range(i + 1):
            g = exps[i] + exps[j]
            norm += (1 if i == j else 2) * df[2 * l] * sqrt_pi3 * c[i] * c[j] / (
                2.0 ** l * g ** (l + 1) * math.sqrt(g))
    return c / math.sqrt(norm)


@dataclass
class BasisSet:
    name: str
    shells: list
    pure: bool = True
    natoms: int = 0
    shell_offsets: np.ndarray = field(default=None)
    # core electrons an effective core potential replaces, per atom (the
    # nuclear charges of the one- and two-electron Hamiltonian are reduced
    # by them: core/basis_set.cpp:85-88)
    ecp_electrons: np.ndarray = field(default=None)

    def __post_init__(self):
        offs = [0]
        for sh in self.shells:
            offs.append(offs[-1] + self.nfun(sh.l))
        self.shell_offsets = np.array(offs, dtype=np.int64)
        if self.ecp_electrons is None:
            self.ecp_electrons = np.zeros(self.natoms, dtype=np.int64)

    def effective_charges(self, numbers) -> np.ndarray:
        """Nuclear charges seen by the electrons: Z minus the ECP core electrons."""
        return np.asarray(numbers, dtype=np.float64) - self.ecp_electrons

    def nfun(self, l: int) -> int:
        return 2 * l + 1 if (self.pure and l >= 2) else (l + 1) * (l + 2) // 2

    @property
    def nbf(self) -> int:
        return int(self.shell_offsets[-1])

    @property
    def nsh(self) -> int:
        return len(self.shells)

    @property
    def max_l(self) -> int:
        return max(sh.l for sh in self.shells) if self.shells else 0

    def flat(self):
        """Flat arrays (l, atom, centers[nsh,3], prim_off[nsh+1], exps, coefs)
        in the sorted AO order -- the layout the C-ABI takes."""
        l = np.array([s.l for s in self.shells], dtype=np.int32)
        atom = np.array([s.atom for s in self.shells], dtype=np.int32)
        cen = np.ascontiguousarray(np.array([s.center for s in self.shells], dtype=np.float64).reshape(-1, 3))
        po = np.zeros(len(self.shells) + 1, dtype=np.int32)
        for i, s in enumerate(self.shells):
            po[i + 1] = po[i] + s.nprim
        exps = np.ascontiguousarray(np.concatenate([s.exps for s in self.shells]) if self.shells else np.zeros(0))
        coefs = np.ascontiguousarray(np.concatenate([s.coefs for s in self.shells]) if self.shells else np.zeros(0))
        return l, atom, cen, po, exps.astype(np.float64), coefs.astype(np.float64)


def load_basis(name: str, numbers, coords_bohr, pure: bool = True, sort: bool = True) -> BasisSet:
    """Build the basis for atoms with nuclear charges `numbers` at
    `coords_bohr` (natoms x 3) following data::BasisSet::from_basis_name
    (data/basis_set.cpp:594-638) and the PSI4-mode conversion of the plugin
    (ms/scf.cpp:273 -> utils.cpp:287-310)."""
    with open(_basis_file(name)) as f:
        data = json.load(f)
    coords = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
    shells = []
    ecp = np.zeros(len(numbers), dtype=np.int64)
    for ia, z in enumerate(numbers):
        key = str(int(z))
        if key not in data["elements"]:
            raise ValueError(f"Element (Z={key}) is not supported in basis set {name}")
        ecp[ia] = int(data["elements"][key].get("ecp_electrons", 0))
        for entry in data["elements"][key]["electron_shells"]:
            ams = entry["angular_momentum"]
            exps = np.array([float(x) for x in entry["exponents"]])
            for j, col in enumerate(entry["coefficients"]):
                l = int(ams[j] if len(ams) > 1 else ams[0])
                coefs = np.array([float(x) for x in col])
                order = np.argsort(-exps, kind="stable")  # descending exponent
                shells.append(Shell(ia, l, exps[order].copy(), coefs[order].copy(), coords[ia].copy()))
    if sort:
        # stable sort by l, then by descending leading exponent (data/basis_set.cpp:147-161)
        shells = sorted(shells, key=lambda s: (s.l, -s.exps[0]))
    for s in shells:
        s.coefs = normalize_psi4(s.l, s.exps, s.coefs)
    return BasisSet(name=name.lower(), shells=shells, pure=pure, natoms=len(numbers), ecp_electrons=ecp)


# ---------------------------------------------------------------------------
# cartesian <-> solid-harmonic conventions
# ---------------------------------------------------------------------------

def cart_list(l: int):
    """CCA order: lx descending, then ly descending."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def _dfact(n: int) -> float:
    r = 1.0
    k = n
    while k > 1:
        r *= k
        k -= 2
    return r


def solid_harmonic_matrix(l: int) -> np.ndarray:
    """Rows m = -l..l: real solid harmonics in terms of x^l-normalised
    cartesian monomials (Helgaker, Jorgensen & Olsen eq. 6.4.47-6.4.50),
    each row rescaled to unit self-overlap."""
    comps = cart_list(l)
    idx = {c: i for i, c in enumerate(comps)}
    M = np.zeros((2 * l + 1, len(comps)))
    for m in range(-l, l + 1):
        am = abs(m)
        vvm = 1 if m < 0 else 0
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                for vv in range(vvm, am + 1, 2):
                    sgn = -1.0 if ((t + (vv - vvm) // 2) & 1) else 1.0
                    c = sgn * 0.25 ** t * math.comb(l, t) * math.comb(l - t, am + t) * math.comb(t, u) * math.comb(am, vv)
                    px, py, pz = 2 * t + am - 2 * u - vv, 2 * u + vv, l - 2 * t - am
                    if px < 0 or pz < 0:
                        continue
                    M[m + l, idx[(px, py, pz)]] += c
    # normalise under x^l-normalised cartesians
    S = np.zeros((len(comps), len(comps)))
    for i, a in enumerate(comps):
        for j, b in enumerate(comps):
            s = [a[k] + b[k] for k in range(3)]
            if any(x & 1 for x in s):
                continue
            S[i, j] = _dfact(s[0] - 1) * _dfact(s[1] - 1) * _dfact(s[2] - 1) / _dfact(2 * l - 1)
    for r in range(M.shape[0]):
        M[r] /= math.sqrt(M[r] @ S @ M[r])
    return M
