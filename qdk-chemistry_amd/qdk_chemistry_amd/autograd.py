"""torch.autograd bindings of the Fock-build engine's nuclear gradients.

Both Functions take and return CPU tensors: the GPU work runs inside libqfe
(its own HIP runtime), torch only carries the autograd graph.

* ERIEnergy(coords, P): E2 = 1/2 Tr(P J(P)) - 1/4 Tr(P K(P)) at the
  geometry coords (natoms x 3, bohr) with the density P held in the AO basis
  of that geometry.  d/dcoords is the adjoint kernel (derivative ERIs
  contracted with P x P, the dJ + dK of ERI::get_gradients,
  libint2_direct.cpp:646-890); d/dP = J(P) - 1/2 K(P).
* RHFEnergy(coords): converged RHF energy; backward is the analytic RHF
  gradient (SCFImpl::get_gradients_, scf_impl.cpp:726-843).
"""
from __future__ import annotations

import numpy as np
import torch

from .basis import load_basis
from .eri import ERIConfig, create
from .molecule import Structure
from .scf import RHF, SCFConfig


def _structure(numbers, coords):
    return Structure(np.asarray(numbers, dtype=np.int64), coords.detach().cpu().double().numpy().copy())


class ERIEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, P, numbers, basis_name, eri_cfg=None):
        mol = _structure(numbers, coords)
        b = load_basis(basis_name, mol.numbers, mol.coords)
        eng = create(b, mol, eri_cfg or ERIConfig())
        Pn = P.detach().cpu().double().numpy()
        J, K = eng.build_JK(Pn)
        E2 = 0.5 * float(np.sum(Pn * J)) - 0.25 * float(np.sum(Pn * K))
        dJ, dK = eng.get_gradients(Pn)
        na = mol.natoms
        ctx.save_for_backward(torch.from_numpy((dJ + dK).reshape(3, na).T.copy()),
                              torch.from_numpy(J - 0.5 * K))
        return torch.tensor(E2, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        gR, F2 = ctx.saved_tensors
        return gE * gR, gE * F2, None, None, None


class RHFEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, numbers, basis_name, scf_cfg=None):
        mol = _structure(numbers, coords)
        b = load_basis(basis_name, mol.numbers, mol.coords)
        scf = RHF(mol, b, scf_cfg or SCFConfig())
        r = scf.run()
        ctx.save_for_backward(torch.from_numpy(scf.gradient()))
        return torch.tensor(r["energy"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        (g,) = ctx.saved_tensors
        return gE * g, None, None, None
