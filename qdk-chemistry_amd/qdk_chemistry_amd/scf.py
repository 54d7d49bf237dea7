"""RHF driver around the GPU Fock build.

Restates the Fock-build step of SCFImpl::iterate_
(scf/src/scf/scf_impl.cpp:495-641):
  * P_diff = P for step < incremental_fock_start_step (3) or every
    fock_reset_steps, with F reset to H; otherwise P_diff = P - P_last (:568-578)
  * J, K = build_JK(P_diff, 1, 0, 0) (:581); F += J - 1/2 K (:363-391)
  * E = E_nuc + Tr(P H) + 1/2 Tr(P (F - H)) (:402-420)
  * converged when ||P - P_last||_F / N < density_threshold and
    ||FPS - SPF||_inf / N < og_threshold (scf_algorithm.cpp:440-495), with
    density_threshold = 1e2 * convergence_threshold and og_threshold =
    convergence_threshold (ms/scf.cpp:167-174)
  * orthogonalisation X = U s^-1/2 dropping s < max(s)/1e9 (:422-493)
The orbital update is DIIS (scf_algorithm/diis.cpp) in FP64, on the device
by default (qfe_scf_*: F, P, the DIIS history, X^T F X, the eigensolve and
the new density stay in HBM) or on the host; the reference's default switches to GDM near convergence, which changes the
iteration path but not the converged energy.  After convergence a full
(non-incremental) rebuild gives the final Fock matrix (:612-621).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .eri import ERIConfig, RysERI, create, eigh_gpu, one_body, one_body_gpu, one_body_grad, one_body_grad_gpu


@dataclass
class SCFConfig:
    convergence_threshold: float = 1e-7
    max_iterations: int = 100
    incremental_fock_start_step: int = 3
    fock_reset_steps: int = 1073741824
    diis_size: int = 10
    lindep_threshold: float = -1.0
    # mixed precision: once the orbital gradient drops below this value the
    # driver resets the Fock matrix with a full FP64 build and finishes the
    # SCF with FP64 incremental builds (DIIS, orthogonalisation and
    # diagonalisation are FP64 throughout)
    fp64_switch_gradient: float = 1e-5
    eri: ERIConfig = field(default_factory=ERIConfig)
    # S, T, V and their derivatives on the ERI engine's GPU (else on the host)
    one_body_on_gpu: bool = True
    # diagonalise the orthogonalised Fock matrix on that GPU (rocSOLVER dsyevd)
    eigen_on_gpu: bool = True
    # keep F, P, the DIIS history and the orbital update in the engine's HBM
    # (qfe_scf_*) when the ERI engine is the GPU Rys engine
    device_resident: bool = True


def _one_body(basis, structure, op, cfg):
    if cfg.one_body_on_gpu:
        return one_body_gpu(basis, structure, op, cfg.eri.device_id)
    return one_body(basis, structure, op)


def _one_body_grad(basis, structure, op, W, cfg):
    if cfg.one_body_on_gpu:
        return one_body_grad_gpu(basis, structure, op, W, cfg.eri.device_id)
    return one_body_grad(basis, structure, op, W)


def _fractional_occupations(eps, nelec):
    """Fill orbitals by energy; a degenerate group at the Fermi level shares
    the remaining electrons equally (spherical average of an open shell)."""
    occ = np.zeros_like(eps)
    left = float(nelec)
    i = 0
    n = len(eps)
    while left > 1e-12 and i < n:
        j = i
        while j + 1 < n and abs(eps[j + 1] - eps[i]) < 1e-4:
            j += 1
        g = j - i + 1
        take = min(left, 2.0 * g)
        occ[i:j + 1] = take / g
        left -= take
        i = j + 1
    return occ


def atom_guess(structure, basis_name, pure=True, eri_cfg=None, iterations=30):
    """Superposition of spherically averaged atomic densities (the role of
    atom_guess, scf/src/scf/guess.cpp:23-101): one fractional-occupation SCF
    per element with the same engine, placed block-diagonally in AO order."""
    from .basis import load_basis
    from .molecule import Structure
    mol_basis = load_basis(basis_name, structure.numbers, structure.coords, pure=pure)
    P = np.zeros((mol_basis.nbf, mol_basis.nbf))
    dens = {}
    for z in sorted(set(int(x) for x in structure.numbers)):
        at = Structure(np.array([z]), np.zeros((1, 3)))
        b = load_basis(basis_name, at.numbers, at.coords, pure=pure)
        eri = create(b, at, eri_cfg or ERIConfig())
        dev = (eri_cfg or ERIConfig()).device_id
        S = one_body_gpu(b, at, 0, dev)
        H = one_body_gpu(b, at, 1, dev) + one_body_gpu(b, at, 2, dev)
        X = orthogonalizer(S)
        e, Cp = np.linalg.eigh(X.T @ H @ X)
        C = X @ Cp
        Pa = (C * _fractional_occupations(e, z)) @ C.T
        for _ in range(iterations):
            J, K = eri.build_JK(Pa)
            F = H + J - 0.5 * K
            e, Cp = np.linalg.eigh(X.T @ F @ X)
            C = X @ Cp
            Pn = (C * _fractional_occupations(e, z)) @ C.T
            done = np.max(np.abs(Pn - Pa)) < 1e-7
            Pa = 0.5 * (Pa + Pn) if not done else Pn
            if done:
                break
        dens[z] = Pa
    # scatter: shells of each atom keep their relative order in the molecule
    offs = mol_basis.shell_offsets
    for ia, z in enumerate(structure.numbers):
        idx = []
        for s, sh in enumerate(mol_basis.shells):
            if sh.atom == ia:
                idx.extend(range(offs[s], offs[s + 1]))
        idx = np.array(idx)
        P[np.ix_(idx, idx)] = dens[int(z)]
    return P


def orthogonalizer(S, lindep=-1.0, device_id=None):
    """X = U s^-1/2 over the eigenvalues s >= max(s) / 1e9 (or >= lindep)
    of S (scf_impl.cpp:422-493); the eigensolve on the GPU when device_id
    is given."""
    s, U = np.linalg.eigh(S) if device_id is None else eigh_gpu(S, device_id)
    thr = lindep if lindep >= 0 else s.max() / 1e9
    keep = s >= thr
    return U[:, keep] / np.sqrt(s[keep])


class RHF:
    def __init__(self, structure, basis, cfg: SCFConfig | None = None, eri=None):
        if np.any(getattr(basis, "ecp_electrons", 0)):
            raise NotImplementedError("effective core potentials: the ECP integrals are not part of this package")
        self.structure = structure
        self.basis = basis
        self.cfg = cfg or SCFConfig()
        self.eri = eri or create(basis, structure, self.cfg.eri)
        self.nocc = structure.nelectrons // 2
        if structure.nelectrons % 2:
            raise ValueError("RHF needs an even number of electrons")
        self.S = _one_body(basis, structure, 0, self.cfg)
        self.H = _one_body(basis, structure, 1, self.cfg) + _one_body(basis, structure, 2, self.cfg)
        self.X = orthogonalizer(self.S, self.cfg.lindep_threshold,
                                self.cfg.eri.device_id if self.cfg.eigen_on_gpu else None)
        self.E_nuc = structure.nuclear_repulsion()
        self.build_times = []
        self.build_stats = []

    def _diag(self, F):
        Fo = self.X.T @ F @ self.X
        Fo = 0.5 * (Fo + Fo.T)
        if self.cfg.eigen_on_gpu:
            e, Cp = eigh_gpu(Fo, self.cfg.eri.device_id)
        else:
            e, Cp = np.linalg.eigh(Fo)
        C = self.X @ Cp
        return e, C

    def _density(self, C):
        Co = C[:, : self.nocc]
        return 2.0 * Co @ Co.T

    def _jk(self, P):
        t = time.perf_counter()
        J, K = self.eri.build_JK(P, 1.0, 0.0, 0.0)
        self.build_times.append(time.perf_counter() - t)
        if hasattr(self.eri, "stats"):
            self.build_stats.append(self.eri.stats())
        return J, K

    def energy(self, P, F):
        return self.E_nuc + float(np.sum(P * self.H)) + 0.5 * float(np.sum(P * (F - self.H)))

    def run(self, P0=None, verbose=False, guess="atom"):
        cfg = self.cfg
        n = self.basis.nbf
        if P0 is None and guess == "atom":
            P = atom_guess(self.structure, self.basis.name, self.basis.pure, self.cfg.eri)
        elif P0 is None:
            e, C = self._diag(self.H)
            P = self._density(C)
        else:
            P = np.array(P0, dtype=np.float64)
        if cfg.device_resident and isinstance(self.eri, RysERI):
            return self._run_device(P, verbose)
        conv = cfg.convergence_threshold
        dens_thr, og_thr = conv * 1e2, conv
        F = self.H.copy()
        P_last = np.zeros_like(P)
        P_prev_conv = np.zeros_like(P)
        E_last = 0.0
        diis_F, diis_E = [], []
        converged = False
        mixed = bool(getattr(self.eri, "mixed", False))
        self.mixed_builds = 0
        reset_now = False
        og = np.inf
        for step in range(cfg.max_iterations):
            if mixed and og < cfg.fp64_switch_gradient:
                # final iterations in FP64, starting from a full FP64 rebuild
                self.eri.set_precision(False)
                mixed = False
                reset_now = True
            if step < cfg.incremental_fock_start_step or step % cfg.fock_reset_steps == 0 or reset_now:
                P_diff = P
                F = self.H.copy()
                reset_now = False
            else:
                P_diff = P - P_last
            if mixed:
                self.mixed_builds += 1
            P_last = P.copy()
            J, K = self._jk(P_diff)
            F = F + J - 0.5 * K
            E = self.energy(P, F)
            if not np.isfinite(E):
                raise RuntimeError("NaN or INF Encountered in SCF Energy")
            FPS = F @ P @ self.S
            err = FPS - FPS.T
            og = float(np.max(np.abs(err))) / n
            drms = float(np.linalg.norm(P_prev_conv - P)) / n
            P_prev_conv = P.copy()
            if verbose:
                print(f"Step {step:03d}: E={E:.15e}, DE={E - E_last:+.3e}, |DP|={drms:.3e}, |DG|={og:.3e}")
            E_last = E
            if drms < dens_thr and og < og_thr:
                converged = True
                break
            diis_F.append(F.copy())
            diis_E.append(err)
            if len(diis_F) > cfg.diis_size:
                diis_F.pop(0)
                diis_E.pop(0)
            m = len(diis_F)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(i + 1):
                    B[i, j] = B[j, i] = float(np.sum(diis_E[i] * diis_E[j]))
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                c = np.linalg.solve(B, rhs)[:m]
            except np.linalg.LinAlgError:
                c = np.zeros(m)
                c[-1] = 1.0
            Fd = sum(ci * Fi for ci, Fi in zip(c, diis_F))
            e, C = self._diag(Fd)
            P = self._density(C)
        if not converged:
            raise RuntimeError(f"SCF failed to converge after {cfg.max_iterations} steps")
        if mixed:
            # converged before the switch point: the final rebuild is FP64
            self.eri.set_precision(False)
        # final full rebuild (scf_impl.cpp:612-621)
        J, K = self._jk(P)
        F = self.H + J - 0.5 * K
        E = self.energy(P, F)
        e, C = self._diag(F)
        self.result = dict(energy=E, P=P, F=F, C=C, eps=e, iterations=step + 1, converged=converged,
                           build_seconds=float(np.sum(self.build_times)), mixed_builds=self.mixed_builds)
        return self.result

    def _run_device(self, P, verbose=False):
        """The loop of run() with F, P and DIIS in the engine's HBM: the same
        incremental-build schedule, convergence test, mixed-precision switch
        and final full rebuild; only E, the orbital gradient and the density
        change come back per iteration."""
        import ctypes
        cfg = self.cfg
        L = _lib.lib()
        X = np.ascontiguousarray(self.X, dtype=np.float64)
        H = np.ascontiguousarray(self.H, dtype=np.float64)
        S = np.ascontiguousarray(self.S, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        n, norb = X.shape
        h = ctypes.c_void_p()
        _lib.check(L.qfe_scf_create(self.eri.h, _lib.ptr(H), _lib.ptr(S), _lib.ptr(X), norb, self.nocc,
                                    self.E_nuc, cfg.diis_size, ctypes.byref(h)))
        try:
            _lib.check(L.qfe_scf_set_density(h, _lib.ptr(P)))
            conv = cfg.convergence_threshold
            dens_thr, og_thr = conv * 1e2, conv
            mixed = bool(getattr(self.eri, "mixed", False))
            self.mixed_builds = 0
            reset_now = False
            og = np.inf
            E_last = 0.0
            converged = False
            cE, cog, cdr = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            for step in range(cfg.max_iterations):
                if mixed and og < cfg.fp64_switch_gradient:
                    self.eri.set_precision(False)
                    mixed = False
                    reset_now = True
                full = step < cfg.incremental_fock_start_step or step % cfg.fock_reset_steps == 0 or reset_now
                reset_now = False
                if mixed:
                    self.mixed_builds += 1
                _lib.check(L.qfe_scf_step(h, int(full), ctypes.byref(cE), ctypes.byref(cog), ctypes.byref(cdr)))
                E, og, drms = cE.value, cog.value, cdr.value
                st = self.eri.stats()
                self.build_times.append(st["ms_total"] / 1e3)
                self.build_stats.append(st)
                if verbose:
                    print(f"Step {step:03d}: E={E:.15e}, DE={E - E_last:+.3e}, |DP|={drms:.3e}, |DG|={og:.3e}")
                E_last = E
                if drms < dens_thr and og < og_thr:
                    converged = True
                    break
                _lib.check(L.qfe_scf_update(h))
            if not converged:
                raise RuntimeError(f"SCF failed to converge after {cfg.max_iterations} steps")
            if mixed:
                self.eri.set_precision(False)
            e_out = ctypes.c_double()
            _lib.check(L.qfe_scf_finish(h, ctypes.byref(e_out)))
            st = self.eri.stats()
            self.build_times.append(st["ms_total"] / 1e3)
            self.build_stats.append(st)
            P = np.empty((n, n))
            F = np.empty((n, n))
            C = np.empty((n, norb))
            e = np.empty(norb)
            for what, arr in ((0, P), (1, F), (2, C), (3, e)):
                _lib.check(L.qfe_scf_get(h, what, _lib.ptr(arr)))
        finally:
            L.qfe_scf_destroy(h)
        self.result = dict(energy=e_out.value, P=P, F=F, C=C, eps=e, iterations=step + 1, converged=converged,
                           build_seconds=float(np.sum(self.build_times)), mixed_builds=self.mixed_builds)
        return self.result

    def gradient(self):
        """Nuclear gradient of the converged RHF energy, natoms x 3 (hartree/bohr).

        Restates SCFImpl::get_gradients_ (scf_impl.cpp:726-843) for RHF:
        nuclear repulsion + sum P (T' + V') - sum W S' + dJ + dK, with the
        energy-weighted density W = 1/2 P F P (= 2 sum_occ eps_i C_i C_i^T at
        convergence) and dJ, dK from the adjoint ERI kernel."""
        if not hasattr(self, "result"):
            raise RuntimeError("run() the SCF before asking for its gradient")
        r = self.result
        P, F = r["P"], r["F"]
        W = 0.5 * P @ F @ P  # energy-weighted density (scf_impl.cpp:785-798 holds -W)
        b, m = self.basis, self.structure
        g = m.nuclear_repulsion_gradient()
        c = self.cfg
        g = g + _one_body_grad(b, m, 1, P, c) + _one_body_grad(b, m, 2, P, c) - _one_body_grad(b, m, 0, W, c)
        dJ, dK = self.eri.get_gradients(P, 1.0, 0.0, 0.0)
        na = m.natoms
        self.grad_parts = dict(dJ=dJ.reshape(3, na).T.copy(), dK=dK.reshape(3, na).T.copy())
        return g + (dJ + dK).reshape(3, na).T


class UHF:
    """Open-shell SCF through the engine's two-density build (ndm = 2,
    ERI::build_JK with has_spin_split_density, eri.h:79-80): J and K of the
    alpha and beta densities in one call, F_s = H + J(P_a) + J(P_b) - K(P_s),
    E = E_nuc + 1/2 sum_s Tr(P_s (H + F_s)).  The orbital update is DIIS over
    both spins (errors F_s P_s S - S P_s F_s, one set of coefficients); the
    guess is given densities, the atomic density split evenly, the core
    Hamiltonian, or ("rohf") the converged ROHF densities of the same
    occupation, which steer the iterations to the ground state where the
    other guesses can settle in an excited one.
    Convergence: |dE| below the threshold squared and max|error| below
    the threshold.  n_alpha = (N + M - 1) / 2 and n_beta = (N - M + 1) / 2
    for N electrons and multiplicity M."""

    def __init__(self, structure, basis, multiplicity, cfg: SCFConfig | None = None, eri=None):
        if np.any(getattr(basis, "ecp_electrons", 0)):
            raise NotImplementedError("effective core potentials: the ECP integrals are not part of this package")
        self.structure = structure
        self.basis = basis
        self.cfg = cfg or SCFConfig()
        self.eri = eri or create(basis, structure, self.cfg.eri)
        ne = structure.nelectrons
        if multiplicity < 1 or (ne + multiplicity - 1) % 2 or multiplicity - 1 > ne:
            raise ValueError(f"multiplicity {multiplicity} does not fit {ne} electrons")
        self.na = (ne + multiplicity - 1) // 2
        self.nb = (ne - multiplicity + 1) // 2
        self.S = _one_body(basis, structure, 0, self.cfg)
        self.H = _one_body(basis, structure, 1, self.cfg) + _one_body(basis, structure, 2, self.cfg)
        self.X = orthogonalizer(self.S, self.cfg.lindep_threshold,
                                self.cfg.eri.device_id if self.cfg.eigen_on_gpu else None)
        self.E_nuc = structure.nuclear_repulsion()

    def _eig(self, F):
        Fo = self.X.T @ F @ self.X
        Fo = 0.5 * (Fo + Fo.T)
        e, Cp = eigh_gpu(Fo, self.cfg.eri.device_id) if self.cfg.eigen_on_gpu else np.linalg.eigh(Fo)
        return e, self.X @ Cp

    def _fock(self, Pa, Pb):
        J, K = self.eri.build_JK(np.stack([Pa, Pb]), 1.0, 0.0, 0.0)
        Jt = J[0] + J[1]
        Fa, Fb = self.H + Jt - K[0], self.H + Jt - K[1]
        E = self.E_nuc + 0.5 * float(np.sum(Pa * (self.H + Fa)) + np.sum(Pb * (self.H + Fb)))
        return Fa, Fb, E

    def _guess(self, guess):
        if isinstance(guess, tuple):  # (P_alpha, P_beta)
            return np.array(guess[0], dtype=np.float64), np.array(guess[1], dtype=np.float64)
        if guess == "atom":
            P = atom_guess(self.structure, self.basis.name, self.basis.pure, self.cfg.eri)
            return 0.5 * P, 0.5 * P
        _, C = self._eig(self.H)
        return C[:, :self.na] @ C[:, :self.na].T, C[:, :self.nb] @ C[:, :self.nb].T

    def _update(self, Fa, Fb, Pa, Pb, diis, errs):
        """DIIS-extrapolated Fock matrices -> new densities (and orbitals)."""
        Fa_d, Fb_d = self._diis(diis, errs, [Fa, Fb])
        ea, Ca = self._eig(Fa_d)
        eb, Cb = self._eig(Fb_d)
        return Ca[:, :self.na] @ Ca[:, :self.na].T, Cb[:, :self.nb] @ Cb[:, :self.nb].T, (ea, Ca, eb, Cb)

    def _errors(self, Fa, Fb, Pa, Pb):
        ea = Fa @ Pa @ self.S
        eb = Fb @ Pb @ self.S
        return [ea - ea.T, eb - eb.T]

    def _diis(self, hist, errs, Fs):
        hist.append((Fs, errs))
        if len(hist) > self.cfg.diis_size:
            hist.pop(0)
        m = len(hist)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0.0
        for i in range(m):
            for j in range(i + 1):
                B[i, j] = B[j, i] = sum(float(np.sum(a * b)) for a, b in zip(hist[i][1], hist[j][1]))
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:m]
        except np.linalg.LinAlgError:
            c = np.zeros(m)
            c[-1] = 1.0
        return [sum(ci * h[0][k] for ci, h in zip(c, hist)) for k in range(len(Fs))]

    def run(self, guess="atom", verbose=False):
        cfg = self.cfg
        conv = cfg.convergence_threshold
        if isinstance(guess, str) and guess.startswith("rohf"):
            inner = guess.split(":", 1)[1] if ":" in guess else "atom"
            r = ROHF(self.structure, self.basis, self.na - self.nb + 1, cfg, eri=self.eri).run(guess=inner)
            guess = (r["Pa"], r["Pb"])
        Pa, Pb = self._guess(guess)
        E_last = 0.0
        hist = []
        converged = False
        for step in range(cfg.max_iterations):
            Fa, Fb, E = self._fock(Pa, Pb)
            if not np.isfinite(E):
                raise RuntimeError("NaN or INF Encountered in SCF Energy")
            errs = self._errors(Fa, Fb, Pa, Pb)
            gmax = max(float(np.max(np.abs(e))) for e in errs)
            if verbose:
                print(f"Step {step:03d}: E={E:.15e}, DE={E - E_last:+.3e}, |err|={gmax:.3e}")
            if step > 0 and abs(E - E_last) < conv * conv and gmax < conv:
                converged = True
                break
            E_last = E
            Pa, Pb, orbs = self._update(Fa, Fb, Pa, Pb, hist, errs)
        if not converged:
            raise RuntimeError(f"SCF failed to converge after {cfg.max_iterations} steps")
        self.result = dict(energy=E, Pa=Pa, Pb=Pb, Fa=Fa, Fb=Fb, iterations=step + 1, converged=converged)
        return self.result


class ROHF(UHF):
    """Restricted open-shell HF on the same two-density build: one set of
    orbitals, n_beta doubly and n_alpha - n_beta singly occupied.  The
    orbital update diagonalises the effective Fock matrix whose blocks in the
    current orbitals are (F_a + F_b)/2 within the closed, open and virtual
    spaces and between closed and virtual, F_b between closed and open, and
    F_a between open and virtual; DIIS runs on that matrix with the error
    F_eff P S - S P F_eff, P = P_a + P_b (it vanishes exactly when those
    coupling blocks do).  The converged energy is that of the ROHF stationary
    state, independent of the choice of the diagonal blocks."""

    def _errors(self, Fa, Fb, Pa, Pb):
        Feff = self._effective(Fa, Fb)
        e = Feff @ (Pa + Pb) @ self.S
        self._Feff = Feff
        return [e - e.T]

    def _effective(self, Fa, Fb):
        C = self._C
        na, nb = self.na, self.nb
        Fc = 0.5 * (Fa + Fb)
        Fm = C.T @ Fc @ C
        Fbm = C.T @ Fb @ C
        Fam = C.T @ Fa @ C
        R = Fm.copy()
        R[:nb, nb:na] = Fbm[:nb, nb:na]
        R[nb:na, :nb] = Fbm[nb:na, :nb]
        R[nb:na, na:] = Fam[nb:na, na:]
        R[na:, nb:na] = Fam[na:, nb:na]
        SC = self.S @ C
        return SC @ R @ SC.T

    def _guess(self, guess):
        # one set of orbitals: those of the core Hamiltonian, or of the
        # closed-shell Fock matrix of the atomic density
        F = self.H
        if guess == "atom":
            Pa, Pb = super()._guess(guess)
            J, K = self.eri.build_JK(Pa + Pb, 1.0, 0.0, 0.0)
            F = self.H + J - 0.5 * K
        _, self._C = self._eig(F)
        C = self._C
        return C[:, :self.na] @ C[:, :self.na].T, C[:, :self.nb] @ C[:, :self.nb].T

    def _update(self, Fa, Fb, Pa, Pb, diis, errs):
        (Fd,) = self._diis(diis, errs, [self._Feff])
        e, C = self._eig(Fd)
        self._C = C
        return C[:, :self.na] @ C[:, :self.na].T, C[:, :self.nb] @ C[:, :self.nb].T, (e, C)
