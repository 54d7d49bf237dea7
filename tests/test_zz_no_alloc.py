"""qfe_build_jk_device_noalloc: the device-resident build with its per-build
task list, J share slots and survivor masks in the library's fixed scratch
array (no device allocation in the call).  It must give what
qfe_build_jk_device gives, bit for bit (both are deterministic builds), on
two calls in a row, the second with a smaller system, and refuse a build
that needs more than the scratch holds."""
import ctypes

import numpy as np
import pytest

from conftest import random_density

pytestmark = pytest.mark.gpu


def _engine(nwat, seed):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(nwat, seed)
    b = load_basis("cc-pvdz", mol.numbers, mol.coords)
    return b, mol, create(b, mol)


def _build(eng, P, noalloc):
    from qdk_chemistry_amd import _lib
    n = P.shape[0]
    dP, dJ, dK = eng.device_buffers()
    eng.upload(dP, P)
    L = _lib.lib()
    fn = L.qfe_build_jk_device_noalloc if noalloc else L.qfe_build_jk_device
    _lib.check(fn(eng.h, ctypes.c_void_p(dP), ctypes.c_void_p(dJ), ctypes.c_void_p(dK), 1.0, 0.0, 0.0, None))
    return eng.download(dJ, (n, n)), eng.download(dK, (n, n))


def test_noalloc_build_matches_the_device_build_twice():
    from oracle.oracle import OracleERI
    for nwat, seed, dens_seed in ((4, 11, 3), (2, 5, 4)):  # the second call is the smaller system
        b, mol, eng = _engine(nwat, seed)
        P = random_density(b.nbf, dens_seed, 0.3)
        Jn, Kn = _build(eng, P, True)
        Jd, Kd = _build(eng, P, False)
        print(f"(H2O){nwat}: nbf {b.nbf}, quartets {eng.stats()['quartets']}", flush=True)
        assert np.array_equal(Jn, Jd) and np.array_equal(Kn, Kd)
        Jr, Kr = OracleERI(b, mol).build_JK(P)
        assert np.max(np.abs(Jn - Jr)) < 1e-10 and np.max(np.abs(Kn - Kr)) < 1e-10


def test_noalloc_refuses_a_build_beyond_its_scratch():
    from qdk_chemistry_amd import _lib
    L = _lib.lib()
    L.qfe_noalloc_scratch_bytes.restype = ctypes.c_ulonglong
    assert L.qfe_noalloc_scratch_bytes() == 256 << 20
    b, mol, eng = _engine(16, 20261202)
    P = random_density(b.nbf, 5, 1.0)  # a dense density: some 1e8 quartets survive
    with pytest.raises(_lib.QfeError) as e:
        _build(eng, P, True)
    assert e.value.code == _lib.QFE_ERR_INVALID and "scratch" in str(e.value)
    # the engine's own build still runs, and a small one fits again afterwards
    J, K = _build(eng, random_density(b.nbf, 6, 1e-3), False)
    assert np.all(np.isfinite(J))
