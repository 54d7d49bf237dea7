"""GPU J/K parity against the oracle (libint2_direct.cpp:256-626 semantics)."""
import numpy as np
import pytest

from conftest import random_density

pytestmark = pytest.mark.gpu


def _engine(mol, basis_name, **kw):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    b = load_basis(basis_name, mol.numbers, mol.coords)
    return b, create(b, mol, ERIConfig(**kw))


def _oracle(b, mol):
    from oracle.oracle import OracleERI
    return OracleERI(b, mol)


@pytest.mark.parametrize("seed", [1, 2])
def test_water_ccpvdz_random_density(water, seed):
    b, eng = _engine(water, "cc-pvdz")
    P = random_density(b.nbf, seed)
    J, K = eng.build_JK(P)
    Jr, Kr = _oracle(b, water).build_JK(P)
    print("max|dJ| %.3e max|dK| %.3e" % (np.max(np.abs(J - Jr)), np.max(np.abs(K - Kr))), flush=True)
    assert np.max(np.abs(J - Jr)) <= 1e-10
    assert np.max(np.abs(K - Kr)) <= 1e-10


def test_water_cluster4_random_density():
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    b, eng = _engine(mol, "cc-pvdz")
    P = random_density(b.nbf, 5, 0.5)
    J, K = eng.build_JK(P)
    Jr, Kr = _oracle(b, mol).build_JK(P)
    print("max|dJ| %.3e max|dK| %.3e" % (np.max(np.abs(J - Jr)), np.max(np.abs(K - Kr))), flush=True)
    assert np.max(np.abs(J - Jr)) <= 1e-10
    assert np.max(np.abs(K - Kr)) <= 1e-10
    st = eng.stats()
    assert st["quartets"] > 0


def test_mixed_precision_jk_close_to_oracle():
    """FP32 only where Q_ab Q_cd < 1e-6: J/K stay within 1e-8 of the oracle."""
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    b, eng = _engine(mol, "cc-pvdz", is_mix_precision=True, mix_precision_threshold=1e-6)
    P = random_density(b.nbf, 8, 0.5)
    J, K = eng.build_JK(P)
    st = eng.stats()
    assert st["quartets_fp32"] > 0
    Jr, Kr = _oracle(b, mol).build_JK(P)
    print("mixed: fp32 share %.3f max|dJ| %.3e max|dK| %.3e" % (st["quartets_fp32"] / st["quartets"],
          np.max(np.abs(J - Jr)), np.max(np.abs(K - Kr))), flush=True)
    assert np.max(np.abs(J - Jr)) <= 1e-8
    assert np.max(np.abs(K - Kr)) <= 1e-8
    # switching back to FP64 restores the 1e-10 parity
    eng.set_precision(False)
    J, K = eng.build_JK(P)
    assert eng.stats()["quartets_fp32"] == 0
    assert np.max(np.abs(J - Jr)) <= 1e-10 and np.max(np.abs(K - Kr)) <= 1e-10


@pytest.mark.parametrize("mix", [False, True])
def test_segment_digestion_variants_agree(monkeypatch, mix):
    """The segment kernel with exchange and density rows in LDS (default) and
    with exchange rows only (QFE_SEGMENT_DIGEST=0, density read from L2) give
    the same J and K, and both match the oracle."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = random_density(b.nbf, 5, 0.3)
    cfg = ERIConfig(is_mix_precision=mix)
    J1, K1 = create(b, mol, cfg).build_JK(P)
    monkeypatch.setenv("QFE_SEGMENT_DIGEST", "0")
    J0, K0 = create(b, mol, cfg).build_JK(P)
    print("K+P rows vs K rows: max|dJ| %.2e max|dK| %.2e" % (np.max(np.abs(J1 - J0)), np.max(np.abs(K1 - K0))),
          flush=True)
    assert np.max(np.abs(J1 - J0)) < 1e-11
    assert np.max(np.abs(K1 - K0)) < 1e-11
    if not mix:
        Jr, Kr = OracleERI(b, mol).build_JK(P)
        assert np.max(np.abs(J1 - Jr)) < 1e-10
        assert np.max(np.abs(K1 - Kr)) < 1e-10


@pytest.mark.parametrize("mix", [False, True])
def test_jcd_slots_and_atomic_paths_agree(monkeypatch, mix):
    """J_cd summed per ket pair from the share slots (default) and added per
    task with global atomics (QFE_JCD_SLOTS=0) give the same J; both match
    the oracle."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 13)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = random_density(b.nbf, 8, 0.3)
    cfg = ERIConfig(is_mix_precision=mix)
    J1, K1 = create(b, mol, cfg).build_JK(P)
    monkeypatch.setenv("QFE_JCD_SLOTS", "0")
    J0, K0 = create(b, mol, cfg).build_JK(P)
    print("slots vs atomic J_cd: max|dJ| %.2e max|dK| %.2e" % (np.max(np.abs(J1 - J0)), np.max(np.abs(K1 - K0))),
          flush=True)
    assert np.max(np.abs(J1 - J0)) < 1e-11
    assert np.max(np.abs(K1 - K0)) < 1e-11
    if not mix:
        Jr, Kr = OracleERI(b, mol).build_JK(P)
        assert np.max(np.abs(J1 - Jr)) < 1e-10
        assert np.max(np.abs(K1 - Kr)) < 1e-10


def test_jcd_slots_zeroed_between_builds(monkeypatch):
    """Share slots of quartets that pass the Schwarz bound but fail the
    density test keep the previous build's values; the survivor masks leave
    them out: after a dense density, a block-diagonal one (no density between
    molecules, so many such quartets) gives the same J as the atomic path and
    the oracle."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(8, 29)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    dense = random_density(b.nbf, 12, 0.3)
    blk = np.zeros_like(dense)
    atom_bf = np.concatenate([[sh.atom] * b.nfun(sh.l) for sh in b.shells])  # functions follow the shells
    mol_bf = atom_bf // 3  # water_cluster lists each molecule's three atoms together
    for m in range(mol.natoms // 3):
        i = np.where(mol_bf == m)[0]
        blk[np.ix_(i, i)] = dense[np.ix_(i, i)]
    eng = create(b, mol)
    eng.build_JK(dense)
    J1, K1 = eng.build_JK(blk)
    st = eng.stats()
    monkeypatch.setenv("QFE_JCD_SLOTS", "0")
    J0, K0 = create(b, mol).build_JK(blk)
    Jr, Kr = OracleERI(b, mol).build_JK(blk)
    print("block-diagonal after dense: quartets %d, max|J - J_atomic| %.2e, max|J - J_oracle| %.2e"
          % (st["quartets"], np.max(np.abs(J1 - J0)), np.max(np.abs(J1 - Jr))), flush=True)
    assert np.max(np.abs(J1 - J0)) < 1e-11
    assert np.max(np.abs(J1 - Jr)) < 1e-10
    assert np.max(np.abs(K1 - Kr)) < 1e-10


@pytest.mark.parametrize("mix", [False, True])
def test_default_build_is_bitwise_repeatable(mix):
    """use_atomics = False (the reference's default, core/scf.h:81-82):
    repeated builds of one density agree bit for bit, as the reference's
    ordered reduction of thread-local buffers does (libint2_direct.cpp:
    576-592); the fixed-point sums stay within 1e-10 of the oracle."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 13)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = random_density(b.nbf, 9, 0.3)
    eng = create(b, mol, ERIConfig(is_mix_precision=mix))
    J1, K1 = eng.build_JK(P)
    for _ in range(3):
        J2, K2 = eng.build_JK(P)
        assert np.array_equal(J1, J2) and np.array_equal(K1, K2)
    J3, K3 = create(b, mol, ERIConfig(is_mix_precision=mix)).build_JK(P)
    assert np.array_equal(J1, J3) and np.array_equal(K1, K3)
    Jr, Kr = OracleERI(b, mol).build_JK(P)
    tol = 1e-8 if mix else 1e-10
    assert np.max(np.abs(J1 - Jr)) < tol and np.max(np.abs(K1 - Kr)) < tol


def test_atomic_build_matches_oracle_and_deterministic_build():
    """use_atomics = True: FP64 atomics, same values to rounding."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 13)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = random_density(b.nbf, 10, 0.3)
    Ja, Ka = create(b, mol, ERIConfig(use_atomics=True)).build_JK(P)
    Jd, Kd = create(b, mol, ERIConfig(use_atomics=False)).build_JK(P)
    Jr, Kr = OracleERI(b, mol).build_JK(P)
    print("atomics vs fixed point: max|dJ| %.2e max|dK| %.2e" % (np.max(np.abs(Ja - Jd)), np.max(np.abs(Ka - Kd))),
          flush=True)
    assert np.max(np.abs(Ja - Jd)) < 1e-12 and np.max(np.abs(Ka - Kd)) < 1e-12
    assert np.max(np.abs(Ja - Jr)) < 1e-10 and np.max(np.abs(Ka - Kr)) < 1e-10


@pytest.mark.parametrize("atomics,mix", [(False, False), (True, False), (False, True)])
def test_column_tiled_exchange_rows(monkeypatch, atomics, mix):
    """Exchange rows split into shell-aligned column tiles (the path that
    lifts the LDS limit on nbf), forced at small nbf with QFE_KCOL_TILE: the
    same J and K as whole rows, and the oracle's.  Under mixed precision the
    tiled kernel evaluates the FP32 tasks in FP64, so it is held to the
    mixed tolerance against the oracle only."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = random_density(b.nbf, 14, 0.3)
    cfg = ERIConfig(use_atomics=atomics, is_mix_precision=mix)
    J1, K1 = create(b, mol, cfg).build_JK(P)
    monkeypatch.setenv("QFE_KCOL_TILE", "20")  # nbf 96: five or more tiles
    J0, K0 = create(b, mol, cfg).build_JK(P)
    Jr, Kr = OracleERI(b, mol).build_JK(P)
    print("tiled vs whole rows: max|dJ| %.2e max|dK| %.2e" % (np.max(np.abs(J1 - J0)), np.max(np.abs(K1 - K0))),
          flush=True)
    tol = 1e-8 if mix else 1e-10
    if not mix:
        assert np.max(np.abs(J1 - J0)) < 1e-12 and np.max(np.abs(K1 - K0)) < 1e-12
    assert np.max(np.abs(J0 - Jr)) < tol and np.max(np.abs(K0 - Kr)) < tol


def test_primitive_screening_changes_nothing_visible(monkeypatch):
    """Primitive quartets below the engine precision (eps / max|P|, the
    reference's Libint2 setting, libint2_direct.cpp:279-287) are skipped by
    default; QFE_PRIM_SCREEN=0 keeps them all.  Both agree with each other
    and with the oracle (which keeps every primitive quartet)."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import water_cluster
    from conftest import decaying_density
    mol = water_cluster(8, 29)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = decaying_density(b, 15, 1.0)
    J1, K1 = create(b, mol).build_JK(P)
    monkeypatch.setenv("QFE_PRIM_SCREEN", "0")
    J0, K0 = create(b, mol).build_JK(P)
    Jr, Kr = OracleERI(b, mol).build_JK(P)
    print("screened vs all primitive quartets: max|dJ| %.2e max|dK| %.2e"
          % (np.max(np.abs(J1 - J0)), np.max(np.abs(K1 - K0))), flush=True)
    assert np.max(np.abs(J1 - J0)) < 1e-12 and np.max(np.abs(K1 - K0)) < 1e-12
    assert np.max(np.abs(J1 - Jr)) < 1e-10 and np.max(np.abs(K1 - Kr)) < 1e-10


@pytest.mark.parametrize("mix", [False, True])
def test_concurrent_class_streams_match_one_stream(monkeypatch, mix):
    """The class kernels spread over the engine's streams (the default) give
    bit for bit the J and K of one stream (QFE_STREAMS=1): the deterministic
    sums do not depend on the order in which the classes run."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 13)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    P = random_density(b.nbf, 19, 0.3)
    J1, K1 = create(b, mol, ERIConfig(is_mix_precision=mix)).build_JK(P)
    monkeypatch.setenv("QFE_STREAMS", "1")
    J0, K0 = create(b, mol, ERIConfig(is_mix_precision=mix)).build_JK(P)
    assert np.array_equal(J1, J0) and np.array_equal(K1, K0)
