"""Molecular structures used by the Fock-build workloads.

Units: bohr internally; xyz input in angstrom converted with the CODATA-2022
factor the reference's Structure uses (cpp/include/qdk/chemistry/constants.hpp:64-66).
The seeded generators produce the benchmark systems: water clusters on a
3.0 angstrom cubic lattice with +-0.3 angstrom jitter and uniformly random
orientations, and an all-trans alkane with dihedral jitter.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .basis import ELEMENTS, SYMBOLS

BOHR_TO_ANGSTROM = 0.529177210544  # CODATA 2022
ANGSTROM_TO_BOHR = 1.0 / BOHR_TO_ANGSTROM


@dataclass
class Structure:
    numbers: np.ndarray       # nuclear charges (int)
    coords: np.ndarray        # natoms x 3, bohr
    charge: int = 0           # molecular charge

    @property
    def natoms(self) -> int:
        return len(self.numbers)

    @property
    def nelectrons(self) -> int:
        return int(np.sum(self.numbers)) - int(self.charge)

    def nuclear_repulsion(self) -> float:
        e = 0.0
        Z, R = self.numbers, self.coords
        for i in range(len(Z)):
            d = R[i + 1:] - R[i]
            r = np.sqrt(np.einsum("ij,ij->i", d, d))
            e += float(np.sum(Z[i] * Z[i + 1:] / r))
        return e

    def nuclear_repulsion_gradient(self) -> np.ndarray:
        """d E_nuc / d R, atom-major (natoms x 3)."""
        Z, R = self.numbers.astype(np.float64), self.coords
        g = np.zeros_like(R)
        for i in range(len(Z)):
            d = R[i] - R
            r = np.sqrt(np.einsum("ij,ij->i", d, d))
            r[i] = np.inf
            g[i] = -np.sum((Z[i] * Z / r ** 3)[:, None] * d, axis=0)
        return g

    def to_xyz(self, comment: str = "") -> str:
        lines = [str(self.natoms), comment]
        for z, c in zip(self.numbers, self.coords * BOHR_TO_ANGSTROM):
            lines.append(f"{SYMBOLS[int(z)]:2s} {c[0]:18.12f} {c[1]:18.12f} {c[2]:18.12f}")
        return "\n".join(lines) + "\n"

    @staticmethod
    def from_xyz(text: str) -> "Structure":
        lines = text.strip().splitlines()
        n = int(lines[0])
        Z, R = [], []
        for ln in lines[2:2 + n]:
            p = ln.split()
            Z.append(ELEMENTS[p[0]])
            R.append([float(x) for x in p[1:4]])
        return Structure(np.array(Z, dtype=np.int64), np.array(R) * ANGSTROM_TO_BOHR)

    @staticmethod
    def from_angstrom(symbols, coords_ang) -> "Structure":
        return Structure(np.array([ELEMENTS[s] for s in symbols], dtype=np.int64),
                         np.asarray(coords_ang, dtype=np.float64) * ANGSTROM_TO_BOHR)


def water_monomer() -> Structure:
    """O at the origin, H at (0, +-0.757, 0.587) angstrom."""
    return Structure.from_angstrom(["O", "H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.757, 0.587], [0.0, -0.757, 0.587]])


def _quat_rotation(u1, u2, u3) -> np.ndarray:
    """Uniform random rotation from three uniforms (Shoemake 1992)."""
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    q = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
         b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(nmol: int, seed: int, spacing: float = 3.0, jitter: float = 0.3) -> Structure:
    """Seeded (H2O)n on a cubic lattice (angstrom units for the rules).

    Lattice sites fill x fastest: 16 -> 4x2x2, 64 -> 4x4x4.  Each oxygen gets
    U(-jitter, jitter) per axis; each molecule a uniform random rotation about
    its oxygen.  A molecule's jitter and rotation are redrawn together while
    any intermolecular H..H < 1.5 A or O..O < 2.5 A (an O..O clash cannot be
    cured by rotation alone: 3.0 - 2*0.3 < 2.5).
    """
    dims = {1: (1, 1, 1), 2: (2, 1, 1), 4: (2, 2, 1), 8: (2, 2, 2), 16: (4, 2, 2), 32: (4, 4, 2), 64: (4, 4, 4)}[nmol]
    rng = np.random.Generator(np.random.PCG64(seed))
    mono = np.array([[0.0, 0.0, 0.0], [0.0, 0.757, 0.587], [0.0, -0.757, 0.587]])
    placed = []  # list of (3x3) angstrom coordinates
    for k in range(nmol):
        ix = k % dims[0]
        iy = (k // dims[0]) % dims[1]
        iz = k // (dims[0] * dims[1])
        site = np.array([ix, iy, iz], dtype=np.float64) * spacing
        for _ in range(1000):
            o = site + rng.uniform(-jitter, jitter, 3)
            R = _quat_rotation(*rng.uniform(0.0, 1.0, 3))
            mol = o + mono @ R.T
            ok = True
            for other in placed:
                if np.linalg.norm(mol[0] - other[0]) < 2.5:
                    ok = False
                    break
                hh = np.linalg.norm(mol[1:, None, :] - other[None, 1:, :], axis=-1)
                if np.min(hh) < 1.5:
                    ok = False
                    break
            if ok:
                break
        else:
            raise RuntimeError("water_cluster: rejection sampling did not converge")
        placed.append(mol)
    xyz = np.concatenate(placed, axis=0)
    syms = ["O", "H", "H"] * nmol
    return Structure.from_angstrom(syms, xyz)


def linear_alkane(ncarbon: int, seed: int, dihedral_jitter_deg: float = 5.0) -> Structure:
    """All-trans CnH(2n+2): C-C 1.54 A, C-H 1.09 A, tetrahedral angles, each
    C-C-C-C dihedral 180 + U(-jitter, jitter) degrees (seeded PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tet = math.acos(-1.0 / 3.0)
    cc, ch = 1.54, 1.09

    def place(a, b, c, bond, angle, dihed):
        """Atom d bonded to c with |cd| = bond, angle bcd, dihedral abcd."""
        bc = c - b
        bc /= np.linalg.norm(bc)
        n = np.cross(b - a, bc)
        n /= np.linalg.norm(n)
        m = np.cross(n, bc)
        d2 = np.array([-bond * math.cos(angle), bond * math.sin(angle) * math.cos(dihed),
                       bond * math.sin(angle) * math.sin(dihed)])
        return c + d2[0] * bc + d2[1] * m + d2[2] * n

    C = [np.array([0.0, 0.0, 0.0]), np.array([cc, 0.0, 0.0])]
    # third carbon in the xy plane
    C.append(C[1] + cc * np.array([-math.cos(tet), math.sin(tet), 0.0]))
    for k in range(3, ncarbon):
        dih = math.radians(180.0 + rng.uniform(-dihedral_jitter_deg, dihedral_jitter_deg))
        C.append(place(C[k - 3], C[k - 2], C[k - 1], cc, tet, dih))
    C = C[:ncarbon]
    H = []
    for k in range(ncarbon):
        nb = [C[j] for j in (k - 1, k + 1) if 0 <= j < ncarbon]
        if len(nb) == 2:
            # two H in the plane bisecting the C-C-C angle
            u = (nb[0] - C[k]) / np.linalg.norm(nb[0] - C[k])
            v = (nb[1] - C[k]) / np.linalg.norm(nb[1] - C[k])
            bis = -(u + v)
            bis /= np.linalg.norm(bis)
            perp = np.cross(u, v)
            perp /= np.linalg.norm(perp)
            half = tet / 2.0
            for s in (1.0, -1.0):
                H.append(C[k] + ch * (math.cos(half) * bis + s * math.sin(half) * perp))
        else:
            # terminal CH3: staggered around the C-C axis
            other = nb[0]
            ref = C[k + 2] if k == 0 else C[k - 2]
            for j in range(3):
                H.append(place(ref, other, C[k], ch, tet, math.pi + j * 2 * math.pi / 3))
    syms = ["C"] * ncarbon + ["H"] * len(H)
    return Structure.from_angstrom(syms, np.array(C + H))
