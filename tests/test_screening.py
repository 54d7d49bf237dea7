"""Pins the oracle's screening, Schwarz factors, significant pairs and exact
J/K rows, so that the screened quartet census can serve as the checker of
the GPU task list (libint2_direct.cpp:83-151,357-411; schwarz.cpp:26-78).

The census is recounted here by a plain Python restatement of the
reference's canonical loop, and the row oracle is compared with the full
screened build: a dropped or mis-oriented norm, a wrong loop bound, or a
primitive-pair rule with a different cut fails one of them.
"""
import math

import numpy as np
import pytest

from conftest import decaying_density, random_density


def _system(nwat, basis="cc-pvdz", seed=11):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import water_cluster, water_monomer
    mol = water_monomer() if nwat == 1 else water_cluster(nwat, seed)
    return mol, load_basis(basis, mol.numbers, mol.coords)


def _norms(b, P):
    off = b.shell_offsets
    n = b.nsh
    out = np.zeros((n, n))
    for s1 in range(n):
        for s2 in range(n):
            out[s1, s2] = np.max(np.abs(P[off[s1]:off[s1 + 1], off[s2]:off[s2 + 1]]))
    return out


def _sig_prims(b, s1, s2):
    A, B = b.shells[s1], b.shells[s2]
    ab2 = float(np.sum((A.center - B.center) ** 2))
    n = 0
    for a, ca in zip(A.exps, A.coefs):
        for bb, cb in zip(B.exps, B.coefs):
            if not abs(ca * cb * math.exp(-a * bb / (a + bb) * ab2)) < np.finfo(float).eps / 1e10:
                n += 1
    return n


def _census_loop(b, Q, sig, P, thr):
    """The canonical loop of libint2_direct.cpp:357-402 in Python: s1 >= s2
    significant, s3 <= s1, s4 <= (s1 == s3 ? s2 : s3) significant; skip when
    max6 * Q12 * Q34 < thr.  Counts per (l1, l2, l3, l4)."""
    Pn = _norms(b, P)
    nsh = b.nsh
    pairs = {s1: [s2 for s2 in range(s1 + 1) if sig[s1, s2]] for s1 in range(nsh)}
    nq = np.zeros((5, 5, 5, 5))
    npq = np.zeros((5, 5, 5, 5))
    nsp = np.zeros((5, 5, 5, 5))
    spp = {}
    for s1 in range(nsh):
        for s2 in pairs[s1]:
            for s3 in range(s1 + 1):
                s4max = s2 if s1 == s3 else s3
                for s4 in pairs[s3]:
                    if s4 > s4max:
                        break
                    m = max(Pn[s1, s2], Pn[s1, s3], Pn[s2, s3], Pn[s1, s4], Pn[s2, s4], Pn[s3, s4])
                    if m * Q[s1, s2] * Q[s3, s4] < thr:
                        continue
                    ls = tuple(b.shells[s].l for s in (s1, s2, s3, s4))
                    nq[ls] += 1
                    npq[ls] += np.prod([b.shells[s].nprim for s in (s1, s2, s3, s4)])
                    for key in ((s1, s2), (s3, s4)):
                        if key not in spp:
                            spp[key] = _sig_prims(b, *key)
                    nsp[ls] += spp[(s1, s2)] * spp[(s3, s4)]
    return nq, npq, nsp


def test_schwarz_is_sqrt_of_frobenius_norm(water):
    """Q_ab = sqrt(||(ab|ab)||_F) (schwarz.cpp:60-66), every shell pair."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    o = OracleERI(b, water)
    Q = o.schwarz()
    for s1 in range(b.nsh):
        for s2 in range(s1 + 1):
            blk = o.quartet(s1, s2, s1, s2)
            ref = math.sqrt(math.sqrt(float(np.sum(blk * blk))))
            assert abs(Q[s1, s2] - ref) <= 1e-14 * ref
            assert Q[s2, s1] == Q[s1, s2]


def test_significant_pairs_rule():
    """Kept pairs: same centre, or ||S_ab||_F >= shell_pair_threshold
    (libint2_direct.cpp:120-134); checked on a cluster where the overlap test
    removes pairs, against overlaps recomputed from the oracle's S."""
    from oracle.oracle import OracleERI
    mol, b = _system(8, seed=21)
    o = OracleERI(b, mol, shell_pair_threshold=1e-6)
    got = {tuple(x) for x in o.pairs()}
    S = o.onee(0)
    off = b.shell_offsets
    ref = set()
    for s1 in range(b.nsh):
        for s2 in range(s1 + 1):
            same = np.array_equal(b.shells[s1].center, b.shells[s2].center)
            fro = np.sqrt(np.sum(S[off[s1]:off[s1 + 1], off[s2]:off[s2 + 1]] ** 2))
            if same or fro >= 1e-6:
                ref.add((s1, s2))
    assert len(ref) < b.nsh * (b.nsh + 1) // 2  # the rule removes pairs here
    assert got == ref


@pytest.mark.parametrize("nwat,kind,scale", [(1, "random", 1e-10), (2, "random", 1e-8), (2, "decaying", 1e-3)])
def test_census_equals_the_restated_loop(nwat, kind, scale):
    """Per-class surviving quartets and primitive quartets of the oracle's
    census equal a Python recount of the reference's loop.  Densities are
    scaled so that the density test removes quartets the Schwarz bound
    alone keeps."""
    from oracle.oracle import OracleERI
    mol, b = _system(nwat)
    o = OracleERI(b, mol)
    P = random_density(b.nbf, 7, scale) if kind == "random" else decaying_density(b, 7, scale)
    nq, npq, nsp = o.census(P, significant=True)
    Q = o.schwarz()
    sig = np.zeros((b.nsh, b.nsh), dtype=bool)
    for s1, s2 in o.pairs():
        sig[s1, s2] = sig[s2, s1] = True
    rq, rp, rs = _census_loop(b, Q, sig, P, 1e-12)
    assert 0 < rq.sum()
    # the density test removes quartets here (else it would be untested)
    rq_all, _, _ = _census_loop(b, Q, sig, np.ones_like(P), 1e-12)
    assert rq.sum() < rq_all.sum()
    assert np.array_equal(nq, rq)
    assert np.array_equal(npq, rp)
    assert np.array_equal(nsp, rs)


def test_census_of_a_spin_pair_uses_the_first_density():
    """The reference screens both spins with the norms of the first density
    block (compute_shellblock_norm(obs_, P, nbf), libint2_direct.cpp:271)."""
    from oracle.oracle import OracleERI
    mol, b = _system(2)
    o = OracleERI(b, mol)
    Pa = decaying_density(b, 3, 1e-3)
    Pb = decaying_density(b, 4, 1.0)
    two = np.stack([Pa, Pb])
    assert np.array_equal(o.census(two)[0], o.census(Pa)[0])
    assert not np.array_equal(o.census(Pb)[0], o.census(Pa)[0])


def test_jk_rows_equal_the_full_screened_build():
    """The row oracle gives the rows of the full canonical build at the same
    threshold, so the screened J/K of large systems can be checked row by
    row."""
    from oracle.oracle import OracleERI
    mol, b = _system(4)
    o = OracleERI(b, mol)
    P = decaying_density(b, 5, 0.3)
    J, K = o.build_JK(P)
    rows = [0, 5, 17, b.nsh - 1, 30]
    Jr, Kr = o.jk_rows(P, rows)
    off = b.shell_offsets
    idx = np.concatenate([np.arange(off[s], off[s + 1]) for s in rows])
    assert np.max(np.abs(Jr - J[idx])) < 1e-13
    assert np.max(np.abs(Kr - K[idx])) < 1e-13
    # the screening is active at this density: a zero threshold changes K
    Jz, Kz = OracleERI(b, mol, eri_threshold=0.0).jk_rows(P, rows)
    assert np.max(np.abs(Kz - Kr)) > 0.0


def _e2_spin(o, Pa, Pb):
    """Two-electron energy of a spin pair at fixed densities:
    1/2 Tr(Pt J(Pt)) - 1/2 sum_s Tr(Ps K(Ps))."""
    Pt = Pa + Pb
    J, _ = o.build_JK(Pt)
    _, Ka = o.build_JK(Pa)
    _, Kb = o.build_JK(Pb)
    return 0.5 * np.sum(Pt * J) - 0.5 * (np.sum(Pa * Ka) + np.sum(Pb * Kb))


def test_oracle_spin_pair_gradient_vs_central_differences(water):
    """dJ + dK of a spin pair (libint2_direct.cpp:683-692,866-886) equal the
    central differences of the two-electron energy at fixed densities, and
    equal spins reduce to the RHF gradient of their sum."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure
    b0 = load_basis("cc-pvdz", water.numbers, water.coords)
    Pa = decaying_density(b0, 8, 0.2)
    Pb = decaying_density(b0, 9, 0.1)
    o0 = OracleERI(b0, water, eri_threshold=0.0)
    dJ, dK = o0.get_gradients(np.stack([Pa, Pb]), mode=1)
    na = water.natoms
    h = 1e-4
    for atom, x in ((0, 2), (1, 0), (2, 1)):
        e = []
        for sgn in (1, -1):
            R = water.coords.copy()
            R[atom, x] += sgn * h
            m = Structure(water.numbers, R)
            e.append(_e2_spin(OracleERI(load_basis("cc-pvdz", m.numbers, m.coords), m, eri_threshold=0.0), Pa, Pb))
        fd = (e[0] - e[1]) / (2 * h)
        an = dJ[atom + x * na] + dK[atom + x * na]
        assert abs(fd - an) < 1e-7, (atom, x, fd, an)
    P = Pa + Pa
    rJ, rK = o0.get_gradients(P, mode=1)
    sJ, sK = o0.get_gradients(np.stack([Pa, Pa]), mode=1)
    assert np.max(np.abs(sJ - rJ)) < 1e-12 and np.max(np.abs(sK - rK)) < 1e-12
