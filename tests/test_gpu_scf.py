"""RHF through the GPU Fock build: the reference's fixture energies, the
oracle's SCF energy on a cluster, and the error paths of the boundary."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CRAWFORD = [[0.0, -0.0757918436, 0.0], [0.866811829, 0.6014357793, 0.0], [-0.866811829, 0.6014357793, 0.0]]


def _rhf(mol, basis_name, conv=1e-8):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    b = load_basis(basis_name, mol.numbers, mol.coords)
    return RHF(mol, b, SCFConfig(convergence_threshold=conv)).run()


def test_water_ccpvdz_reference_energy(water):
    # python/tests/test_sample_workflow_sci.py:165,201-204 (printed to 1e-8)
    r = _rhf(water, "cc-pvdz")
    print("E =", r["energy"], flush=True)
    assert abs(r["energy"] - (-76.02286895)) < 1e-8


def test_water_def2svp_reference_energy():
    from qdk_chemistry_amd.molecule import Structure
    # cpp/tests/test_scf.cpp:136-142
    r = _rhf(Structure.from_angstrom(["O", "H", "H"], CRAWFORD), "def2-svp")
    assert abs(r["energy"] - (-75.9229032345009)) < 1e-8


def test_cluster_energy_matches_oracle():
    """(H2O)4: GPU SCF vs the oracle's SCF, same geometry and basis."""
    from oracle.oracle import rhf as oracle_rhf
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 11)
    r = _rhf(mol, "cc-pvdz", conv=1e-9)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    ro = oracle_rhf(mol, b, conv_energy=1e-12, conv_grad=1e-9)
    print("GPU %.12f oracle %.12f" % (r["energy"], ro["energy"]), flush=True)
    assert abs(r["energy"] - ro["energy"]) < 1e-9


def test_fock_matrix_matches_oracle_at_converged_density(water):
    """F = H + J - K/2 element-wise at the converged density (<= 1e-10)."""
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    r = _rhf(water, "cc-pvdz")
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    J, K = create(b, water).build_JK(r["P"])
    Jr, Kr = OracleERI(b, water).build_JK(r["P"])
    assert np.max(np.abs((J - 0.5 * K) - (Jr - 0.5 * Kr))) < 1e-10


def test_nan_density_is_an_error(water):
    from qdk_chemistry_amd import _lib
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    eng = create(b, water)
    P = np.zeros((b.nbf, b.nbf))
    P[2, 5] = P[5, 2] = np.nan
    with pytest.raises(_lib.QfeError) as ei:
        eng.build_JK(P)
    assert ei.value.code == _lib.QFE_ERR_RUNTIME and "NaN" in str(ei.value)
    # the engine stays usable after the error
    J, K = eng.build_JK(np.eye(b.nbf))
    assert np.all(np.isfinite(J))


def test_range_separated_exchange_unsupported(water):
    from qdk_chemistry_amd import _lib
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    eng = create(b, water)
    with pytest.raises(_lib.QfeError) as ei:
        eng.build_JK(np.eye(b.nbf), 1.0, 0.0, 0.3)
    assert ei.value.code == _lib.QFE_ERR_UNSUPPORTED


def test_device_entry_matches_host_entry(water):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from conftest import random_density
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    eng = create(b, water)
    P = random_density(b.nbf, 12)
    J, K = eng.build_JK(P)
    dP, dJ, dK = eng.device_buffers()
    eng.upload(dP, P)
    eng.build_JK_device(dP, dJ, dK)
    n = b.nbf
    assert np.max(np.abs(eng.download(dJ, (n, n)) - J)) < 1e-13
    assert np.max(np.abs(eng.download(dK, (n, n)) - K)) < 1e-13


def test_mixed_precision_scf_energy():
    """Mixed-precision SCF (FP32 quartets below 1e-6, FP64 final iterations)
    vs the FP64 SCF on (H2O)8: |dE| <= 1 microhartree."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig
    from qdk_chemistry_amd.molecule import water_cluster
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    mol = water_cluster(8, 3)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    e64 = RHF(mol, b, SCFConfig()).run()["energy"]
    rhf = RHF(mol, b, SCFConfig(eri=ERIConfig(is_mix_precision=True)))
    r = rhf.run()
    print("E64 %.10f Emix %.10f dE %.2e mixed builds %d" % (e64, r["energy"], r["energy"] - e64, r["mixed_builds"]))
    assert r["mixed_builds"] > 0
    assert abs(r["energy"] - e64) <= 1e-6


@pytest.mark.parametrize("mix", [False, True])
def test_device_resident_scf_matches_host_loop(mix):
    """The device-resident iterations (qfe_scf_*) follow the host loop: same
    iteration count, energy within 1e-9, converged F within 1e-8 and orbital
    energies within 1e-8, on (H2O)4/cc-pVDZ, FP64 and mixed precision."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import ERIConfig
    from qdk_chemistry_amd.molecule import water_cluster
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    mol = water_cluster(4, 11)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    out = []
    for dev in (True, False):
        cfg = SCFConfig(convergence_threshold=1e-9, device_resident=dev, eri=ERIConfig(is_mix_precision=mix))
        out.append(RHF(mol, b, cfg).run())
    d, h = out
    print("device %.12f (%d it) host %.12f (%d it)" % (d["energy"], d["iterations"], h["energy"], h["iterations"]),
          flush=True)
    assert abs(d["energy"] - h["energy"]) < 1e-9
    assert abs(d["iterations"] - h["iterations"]) <= 1
    assert np.max(np.abs(d["F"] - h["F"])) < 1e-8
    assert np.max(np.abs(d["eps"] - h["eps"])) < 1e-8
    nocc = mol.nelectrons // 2
    Pd = 2.0 * d["C"][:, :nocc] @ d["C"][:, :nocc].T
    assert np.max(np.abs(Pd - d["P"])) < 1e-6


def test_device_resident_scf_rejects_bad_shapes(water):
    import ctypes
    from qdk_chemistry_amd import _lib
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    eri = create(b, water)
    n = b.nbf
    M = np.eye(n)
    h = ctypes.c_void_p()
    for norb, nocc in ((0, 5), (n + 1, 5), (n, 0), (n, n + 1)):
        rc = _lib.lib().qfe_scf_create(eri.h, _lib.ptr(M), _lib.ptr(M), _lib.ptr(M), norb, nocc, 0.0, 10,
                                       ctypes.byref(h))
        assert rc == _lib.QFE_ERR_INVALID, (norb, nocc)
