"""Host-side pieces of the product: basis conventions, one-electron
integrals, Rys table, structure generators, C-ABI surface (no GPU calls)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")


def test_basis_sizes_and_order(water):
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import linear_alkane, water_cluster
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    assert (b.nbf, b.nsh) == (24, 12)
    # sorted by l, then descending leading exponent (data/basis_set.cpp:147-161)
    keys = [(s.l, -s.exps[0]) for s in b.shells]
    assert keys == sorted(keys)
    assert [s.nprim for s in b.shells] == [8, 8, 3, 3, 1, 1, 1, 3, 1, 1, 1, 1]
    assert load_basis("cc-pvdz", *_nc(water_cluster(16, 20261202))).nbf == 384
    assert load_basis("cc-pvdz", *_nc(water_cluster(64, 20261264))).nbf == 1536
    assert load_basis("cc-pvtz", *_nc(linear_alkane(20, 20261220))).nbf == 1188


def _nc(m):
    return m.numbers, m.coords


def test_structures_match_fixtures():
    from qdk_chemistry_amd.molecule import Structure, linear_alkane, water_cluster
    for name, mol in (("h2o16_seed20261202", water_cluster(16, 20261202)),
                      ("h2o64_seed20261264", water_cluster(64, 20261264)),
                      ("c20h42_seed20261220", linear_alkane(20, 20261220))):
        ref = Structure.from_xyz(open(os.path.join(GOLD, name + ".xyz")).read())
        assert np.array_equal(ref.numbers, mol.numbers)
        assert np.max(np.abs(ref.coords - mol.coords)) < 1e-9


def test_cluster_rejection_rule():
    from qdk_chemistry_amd.molecule import BOHR_TO_ANGSTROM, water_cluster
    m = water_cluster(64, 20261264)
    R = m.coords.reshape(64, 3, 3) * BOHR_TO_ANGSTROM
    for i in range(64):
        for j in range(i):
            assert np.linalg.norm(R[i, 0] - R[j, 0]) >= 2.5
            assert np.min(np.linalg.norm(R[i, 1:, None] - R[j, None, 1:], axis=-1)) >= 1.5


def test_solid_harmonics_orthonormal():
    from qdk_chemistry_amd.basis import _dfact, cart_list, solid_harmonic_matrix
    for l in range(5):
        M = solid_harmonic_matrix(l)
        comps = cart_list(l)
        S = np.zeros((len(comps), len(comps)))
        for i, a in enumerate(comps):
            for j, c in enumerate(comps):
                s = [a[k] + c[k] for k in range(3)]
                if not any(x & 1 for x in s):
                    S[i, j] = _dfact(s[0] - 1) * _dfact(s[1] - 1) * _dfact(s[2] - 1) / _dfact(2 * l - 1)
        assert np.allclose(M @ S @ M.T, np.eye(2 * l + 1), atol=1e-13)


@pytest.mark.parametrize("basis", ["cc-pvdz", "def2-svp", "6-31g*"])
def test_one_body_vs_oracle(water, basis):
    from oracle.oracle import OracleERI
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import one_body
    b = load_basis(basis, water.numbers, water.coords)
    o = OracleERI(b, water)
    for op in range(3):
        assert np.max(np.abs(one_body(b, water, op) - o.onee(op))) < 1e-12
    S = one_body(b, water, 0)
    assert np.allclose(np.diag(S), 1.0, atol=1e-12)  # normalised functions


def test_one_body_gradient_fd(water):
    """sum W dO/dR vs central differences of sum W O (h = 1e-4 bohr)."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import one_body, one_body_grad
    from qdk_chemistry_amd.molecule import Structure
    b = load_basis("cc-pvdz", water.numbers, water.coords)
    rng = np.random.Generator(np.random.PCG64(1))
    W = rng.uniform(-1, 1, (b.nbf, b.nbf))
    W = W + W.T
    h = 1e-4
    for op in range(3):
        g = one_body_grad(b, water, op, W)
        for a in range(3):
            for x in range(3):
                vals = []
                for sgn in (1, -1):
                    R = water.coords.copy()
                    R[a, x] += sgn * h
                    m = Structure(water.numbers, R)
                    bb = load_basis("cc-pvdz", m.numbers, m.coords)
                    vals.append(np.sum(W * one_body(bb, m, op)))
                fd = (vals[0] - vals[1]) / (2 * h)
                assert abs(g[a, x] - fd) < 1e-6, (op, a, x, g[a, x], fd)


def test_rys_table_vs_mpmath():
    import gen_rys_table as g
    from qdk_chemistry_amd._lib import LIB_DIR
    tab = g.load(os.path.join(LIB_DIR, "rys_table.bin"))
    worst = 0.0
    for rec in json.load(open(os.path.join(GOLD, "rys.json"))):
        u, w = g.evaluate(tab, rec["n"], rec["T"])
        ur = np.array([float(x) for x in rec["u"]])
        wr = np.array([float(x) for x in rec["w"]])
        worst = max(worst, np.max(np.abs(u - ur) / ur), np.max(np.abs(w - wr) / wr))
    assert worst < 1e-13


def test_c_abi_exports_every_declared_symbol():
    from qdk_chemistry_amd import _lib
    hdr = open(os.path.join(REPO, "include", "qfe.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(qfe_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 10
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_c_abi_rejects_invalid_basis():
    from qdk_chemistry_amd import _lib
    L = _lib.lib()
    l = np.array([0], dtype=np.int32)
    atom = np.array([0], dtype=np.int32)
    cen = np.zeros(3)
    po = np.array([0, 1], dtype=np.int32)
    exps = np.array([-1.0])
    coefs = np.array([1.0])
    Z = np.array([1.0])
    R = np.zeros(3)
    b = _lib.QfeBasis(1, _lib.ptr(l, ctypes.c_int), _lib.ptr(atom, ctypes.c_int), _lib.ptr(cen),
                      _lib.ptr(po, ctypes.c_int), _lib.ptr(exps), _lib.ptr(coefs), 1, 1, _lib.ptr(Z), _lib.ptr(R))
    h = ctypes.c_void_p()
    rc = L.qfe_create(ctypes.byref(b), None, ctypes.byref(h))
    assert rc == _lib.QFE_ERR_INVALID
    assert b"positive" in L.qfe_last_error()
    rc = L.qfe_build_jk(None, _lib.ptr(np.zeros(1)), None, None, 1.0, 0.0, 0.0, 1)
    assert rc == _lib.QFE_ERR_RUNTIME


def test_eri_multiplexer_routes_like_the_reference():
    """One engine: called once for J and K; two engines: J from the first
    with omega 0 and no K, K from the second with no J (eri_multiplexer.cpp:
    60-71); gradients split the same way unless a gradient engine is given
    (:79-94)."""
    from qdk_chemistry_amd.eri import ERIMultiplexer
    calls = []

    class Fake:
        def __init__(self, tag):
            self.tag = tag

        def build_JK(self, P, alpha, beta, omega, J=True, K=True):
            calls.append((self.tag, omega, J, K))
            return (np.full_like(P, 1.0 if self.tag == "j" else 2.0) if J else None,
                    np.full_like(P, 3.0 if self.tag == "j" else 4.0) if K else None)

        def get_gradients(self, P, alpha, beta, omega, dJ=True, dK=True):
            calls.append(("grad-" + self.tag, omega, dJ, dK))
            return (np.full(3, 5.0) if dJ else None, np.full(3, 6.0) if dK else None)

    P = np.zeros((2, 2))
    J, K = ERIMultiplexer(Fake("j")).build_JK(P, 1.0, 0.0, 0.3)
    assert calls == [("j", 0.3, True, True)] and J[0, 0] == 1.0 and K[0, 0] == 3.0
    calls.clear()
    J, K = ERIMultiplexer(Fake("j"), Fake("k")).build_JK(P, 1.0, 0.0, 0.3)
    assert calls == [("j", 0.0, True, False), ("k", 0.3, False, True)] and J[0, 0] == 1.0 and K[0, 0] == 4.0
    calls.clear()
    ERIMultiplexer(Fake("j"), Fake("k")).get_gradients(P, 1.0, 0.0, 0.3)
    assert calls == [("grad-j", 0.0, True, False), ("grad-k", 0.3, False, True)]
    calls.clear()
    ERIMultiplexer(Fake("j"), Fake("k"), Fake("g")).get_gradients(P, 1.0, 0.0, 0.3)
    assert calls == [("grad-g", 0.3, True, True)]


_INT1E = [(0, "overlap", 1e-9), (1, "kinetic", 1e-9), (2, "nuclear", 1e-8)]


def _teh2hi():
    """TeH2-HI in def2-SVP, cartesian, unsorted shells: the molecule and basis
    of the reference's one-electron integral test (int1e_test.cpp:38-46)."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.molecule import Structure
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "int1e_teh2hi.npz"))
    mol = Structure(f["numbers"], f["coords_angstrom"] / 0.52917721092)
    return f, mol, load_basis("def2-svp", mol.numbers, mol.coords, pure=False, sort=False)


@pytest.mark.parametrize("op,name,tol", _INT1E)
def test_one_body_vs_reference_int1e_fixture(op, name, tol):
    """S, T and V (Te and I with their 28-electron ECP cores: V sees Z - 28,
    core/basis_set.cpp:85-87) against the reference's Psi4 values, with its
    tolerances (int1e_test.cpp:96-99)."""
    from qdk_chemistry_amd.eri import one_body
    f, mol, b = _teh2hi()
    assert b.nbf == f["dm"].shape[0] and list(b.ecp_electrons) == [28, 0, 0, 0, 28]
    assert np.max(np.abs(one_body(b, mol, op) - f[name + "_integral"])) < tol


@pytest.mark.parametrize("op,name,tol", _INT1E)
def test_one_body_gradient_vs_reference_int1e_fixture(op, name, tol):
    """sum_mn D_mn dO_mn/dR for S, T and V against the reference's values
    (int1e_test.cpp:101-105: the same tolerances)."""
    from qdk_chemistry_amd.eri import one_body_grad
    f, mol, b = _teh2hi()
    assert np.max(np.abs(one_body_grad(b, mol, op, f["dm"]) - f[name + "_gradient"])) < tol
