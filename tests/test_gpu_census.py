"""The GPU task list is the reference's screened quartet set: per ERI class,
the shell quartets that pass max6 * Q_ab * Q_cd >= eri_threshold
(libint2_direct.cpp:357-402) and their primitive quartets over the
significant primitive pairs (:29,87,146) equal the oracle's census exactly,
on H2O and (H2O)16 in cc-pVDZ at the atomic-guess and the converged
densities.  Quartets/s is therefore counted over the reference's set."""
import numpy as np
import pytest

from conftest import decaying_density

pytestmark = pytest.mark.gpu


def _pair_class(la, lb):
    hi, lo = max(la, lb), min(la, lb)
    return hi * (hi + 1) // 2 + lo


def canonical_census(nq5):
    """Oracle census [l1, l2, l3, l4] (reference orientation) -> dict keyed
    by the engine's class (la >= lb | lc >= ld), bra pair class >= ket's."""
    out = {}
    for idx in zip(*np.nonzero(nq5)):
        l1, l2, l3, l4 = (int(x) for x in idx)
        p, q = (max(l1, l2), min(l1, l2)), (max(l3, l4), min(l3, l4))
        if _pair_class(*p) < _pair_class(*q):
            p, q = q, p
        key = p + q
        out[key] = out.get(key, 0.0) + float(nq5[idx])
    return out


def gpu_census(eng):
    q, pq = {}, {}
    for prec in (0, 1):
        for cls, nq, npq, _ in eng.class_stats(prec):
            if nq > 0:
                q[cls] = q.get(cls, 0.0) + nq
                pq[cls] = pq.get(cls, 0.0) + npq
    return q, pq


def _check(eng, o, P, label):
    J, K = eng.build_JK(P)
    gq, gp = gpu_census(eng)
    nq, _, nsig = o.census(P, significant=True)
    rq, rp = canonical_census(nq), canonical_census(nsig)
    print(f"{label}: GPU {sum(gq.values()):.0f} quartets, oracle {sum(rq.values()):.0f}", flush=True)
    assert gq == rq, label
    assert gp == rp, label
    assert eng.stats()["quartets"] == int(sum(rq.values()))
    return J, K


@pytest.mark.parametrize("name", ["h2o", "h2o16"])
def test_census_equals_oracle(name):
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster, water_monomer
    from qdk_chemistry_amd.scf import RHF, SCFConfig, atom_guess
    mol = water_monomer() if name == "h2o" else water_cluster(16, 20261202)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    eng = create(b, mol, ERIConfig())
    o = OracleERI(b, mol)
    P0 = atom_guess(mol, "cc-pvdz")
    _check(eng, o, P0, f"{name} atomic guess")
    P = RHF(mol, b, SCFConfig(), eri=eng).run()["P"]
    _check(eng, o, P, f"{name} converged")
    # a density whose small blocks put the density test to work
    _check(eng, o, decaying_density(b, 5, 1e-4), f"{name} decaying")


def test_census_mixed_precision_split_is_a_partition():
    """With mixed precision the FP64 and FP32 parts together are the same set."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    eng = create(b, mol, ERIConfig(is_mix_precision=True))
    P = decaying_density(b, 6, 0.5)
    _check(eng, OracleERI(b, mol), P, "(H2O)4 mixed")
    assert eng.stats()["quartets_fp32"] > 0


def test_census_of_a_spin_pair_screens_with_the_first_density(monkeypatch):
    """ndm = 2: both spins are screened with the first block's norms
    (libint2_direct.cpp:271), so the census is that of P_alpha.  The engine
    precision also follows the first block, eps / max|P_alpha| (:279-287):
    with P_alpha at 1e-4 that drops primitive quartets below 2e-12, which
    moves the J/K of the O(1) P_beta by ~1e-10 -- as in the reference.  The
    element-wise comparison with the oracle (which keeps every primitive
    quartet) therefore runs with primitive screening off."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(2, 5)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    o = OracleERI(b, mol)
    Pa = decaying_density(b, 3, 1e-4)
    Pb = decaying_density(b, 4, 1.0)
    Jr, Kr = o.build_JK(np.stack([Pa, Pb]))
    J, K = create(b, mol).build_JK(np.stack([Pa, Pb]))
    assert np.max(np.abs(J - Jr)) < 1e-9 and np.max(np.abs(K - Kr)) < 1e-9
    monkeypatch.setenv("QFE_PRIM_SCREEN", "0")
    eng = create(b, mol)
    J, K = eng.build_JK(np.stack([Pa, Pb]))
    gq, _ = gpu_census(eng)
    assert gq == canonical_census(o.census(Pa)[0])
    assert np.max(np.abs(J - Jr)) < 1e-10 and np.max(np.abs(K - Kr)) < 1e-10


def test_executed_primitive_count(monkeypatch):
    """The profiled count of evaluated primitive quartets (the FLOP model's
    N_prim): with primitive screening off it is every primitive quartet of
    the significant primitive pairs (the oracle census); with it on (the
    default) it is smaller, in both precisions."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = decaying_density(b, 17, 1.0)
    _, _, nsig = OracleERI(b, mol).census(P, significant=True)
    rp = canonical_census(nsig)
    ex_on = {}
    for mix in (False, True):
        eng = create(b, mol, ERIConfig(profile=True, is_mix_precision=mix))
        eng.build_JK(P)
        for prec in (0, 1):
            for cls, nq, npq, _ in eng.class_stats(prec, executed=True):
                if nq > 0:
                    ex_on[(mix, cls)] = ex_on.get((mix, cls), 0.0) + npq
    monkeypatch.setenv("QFE_PRIM_SCREEN", "0")
    for mix in (False, True):
        eng = create(b, mol, ERIConfig(profile=True, is_mix_precision=mix))
        eng.build_JK(P)
        ex = {}
        for prec in (0, 1):
            for cls, nq, npq, _ in eng.class_stats(prec, executed=True):
                if nq > 0:
                    ex[cls] = ex.get(cls, 0.0) + npq
        assert ex == rp
        on = sum(v for (m, _), v in ex_on.items() if m == mix)
        print(f"mixed={mix}: executed {on:.0f} of {sum(rp.values()):.0f} primitive quartets", flush=True)
        assert 0 < on < sum(rp.values())
