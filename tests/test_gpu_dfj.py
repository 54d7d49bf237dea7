"""Density-fitted J on the GPU (qfe_dfj_*): the reference's RHF-DFJ energy of
water in def2-SVP with def2-universal-jfit, and J against the exact J."""
import json
import os

import numpy as np
import pytest

from conftest import random_density

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _water():
    """make_h2o of the reference's tests (test_common.h), in bohr."""
    from qdk_chemistry_amd.molecule import Structure
    d = json.load(open(os.path.join(GOLD, "reference_fixtures.json")))
    rec = d["h2o"]
    return Structure(np.array(rec["numbers"]), np.array(rec["coords_angstrom"]) / d["bohr_to_angstrom"]), d


def test_dfj_rhf_energy_vs_reference():
    """RHF with density-fitted J and exact K (the multiplexer's J engine and K
    engine, eri_multiplexer.cpp:34-71) reproduces the reference's
    RHF-DFJ/def2-svp/def2-universal-jfit energy (scf_test.cpp:22-25 with
    data/h2o_gauxc.json, 1e-8)."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import DFJ, ERIMultiplexer, create
    from qdk_chemistry_amd.scf import RHF, SCFConfig
    mol, d = _water()
    fx = d["h2o_dfj"]
    b = load_basis(fx["basis"], mol.numbers, mol.coords)
    aux = load_basis(fx["aux_basis"], mol.numbers, mol.coords)
    dfj = DFJ(b, aux, mol)
    eri = ERIMultiplexer(dfj, create(b, mol))
    r = RHF(mol, b, SCFConfig(convergence_threshold=1e-9), eri=eri).run()
    print(f"naux {dfj.naux}: E = {r['energy']:.12f}, reference {fx['energy']:.12f}", flush=True)
    assert abs(r["energy"] - fx["energy"]) < 1e-8


def test_dfj_close_to_exact_j():
    """Fitted J differs from the exact J by the fitting error only: at a
    random density of a water cluster, small against |J| and symmetric."""
    from qdk_chemistry_amd.basis import load_basis
    from qdk_chemistry_amd.eri import DFJ, create
    from qdk_chemistry_amd.molecule import water_cluster
    mol = water_cluster(4, 13)
    b = load_basis("def2-svp", mol.numbers, mol.coords)
    aux = load_basis("def2-universal-jfit", mol.numbers, mol.coords)
    P = random_density(b.nbf, 3, 0.1)
    Jd, _ = DFJ(b, aux, mol).build_JK(P)
    Je, _ = create(b, mol).build_JK(P)
    rel = np.max(np.abs(Jd - Je)) / np.max(np.abs(Je))
    print(f"max|J_df - J| / max|J| = {rel:.2e}", flush=True)
    assert np.max(np.abs(Jd - Jd.T)) < 1e-12 * np.max(np.abs(Jd))
    assert rel < 1e-2
