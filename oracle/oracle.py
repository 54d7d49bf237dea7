"""ctypes wrapper and minimal RHF driver around liboracle.so.

TEST INFRASTRUCTURE ONLY (the parity checker).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module.

The RHF driver restates SCFImpl::iterate_ (scf_impl.cpp:495-641):
orthogonalisation X = U s^-1/2 dropping s < max(s)/1e9 (:422-493), F = H + J
- 1/2 K (:363-391), E = E_nuc + Tr(PH) + 1/2 Tr(P(F-H)) (:402-420), P = 2
C_occ C_occ^T (scf_algorithm.cpp:210-226), with plain DIIS on the
orbital-gradient FPS - SPF (scf_algorithm.cpp:401-435) and a tight
convergence so that the converged energy is exact to ~1e-10 Eh.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, _dp, _dp, ctypes.c_int,
                                    ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_nbf.argtypes = [ctypes.c_void_p]
        L.oracle_num_pairs.argtypes = [ctypes.c_void_p]
        L.oracle_schwarz.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_pairs.argtypes = [ctypes.c_void_p, _ip]
        L.oracle_quartet.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_onee.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.oracle_build_jk.restype = ctypes.c_int
        L.oracle_build_jk.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, _dp]
        L.oracle_census.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_jk_incore.restype = ctypes.c_int
        L.oracle_eri_full.restype = ctypes.c_int
        L.oracle_eri_full.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_jk_incore.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.oracle_gradients.restype = ctypes.c_int
        L.oracle_gradients.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.oracle_sph_row.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_set_row_sample.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_set_grad_class.argtypes = [ctypes.c_int]
        L.oracle_set_grad_centre.argtypes = [ctypes.c_int]
        L.oracle_jk_rows.restype = ctypes.c_int
        L.oracle_jk_rows.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _ip, _dp, _dp, ctypes.c_double,
                                     ctypes.c_double]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def boys(mmax: int, T: float) -> np.ndarray:
    out = np.zeros(mmax + 1)
    lib().oracle_boys(mmax, T, _p(out))
    return out


class OracleERI:
    """CPU restatement of the LIBINT2_DIRECT engine (libint2_direct.cpp)."""

    def __init__(self, basis, structure, eri_threshold=1e-12, shell_pair_threshold=1e-12, nthreads=0):
        self.basis = basis
        self.structure = structure
        l, atom, cen, po, exps, coefs = basis.flat()
        self._keep = (l, atom, cen, po, exps, coefs)
        # nuclear charges seen by the electrons (Z minus ECP core electrons)
        Z = np.ascontiguousarray(np.asarray(structure.numbers, dtype=np.float64)
                                 - np.asarray(getattr(basis, "ecp_electrons", 0), dtype=np.float64))
        R = np.ascontiguousarray(structure.coords.astype(np.float64))
        self._keepZR = (Z, R)
        h = lib().oracle_create(basis.nsh, _p(l, ctypes.c_int), _p(atom, ctypes.c_int), _p(cen),
                                _p(po, ctypes.c_int), _p(exps), _p(coefs), int(basis.pure),
                                structure.natoms, _p(Z), _p(R), eri_threshold, shell_pair_threshold)
        if not h:
            raise ValueError("oracle_create failed (unsupported shell)")
        self.h = ctypes.c_void_p(h)
        if nthreads:
            lib().oracle_set_threads(self.h, nthreads)
        self.nbf = basis.nbf
        self.last_stats = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().oracle_destroy(self.h)
        except Exception:
            pass

    def set_threads(self, n):
        lib().oracle_set_threads(self.h, n)

    def schwarz(self):
        Q = np.zeros((self.basis.nsh, self.basis.nsh))
        lib().oracle_schwarz(self.h, _p(Q))
        return Q

    def pairs(self):
        n = lib().oracle_num_pairs(self.h)
        out = np.zeros((n, 2), dtype=np.int32)
        lib().oracle_pairs(self.h, _p(out, ctypes.c_int))
        return out

    def quartet(self, s1, s2, s3, s4, deriv=0):
        b = self.basis
        n = [b.nfun(b.shells[s].l) for s in (s1, s2, s3, s4)]
        out = np.zeros((12 if deriv else 1,) + tuple(n))
        lib().oracle_quartet(self.h, s1, s2, s3, s4, deriv, _p(out))
        return out if deriv else out[0]

    def onee(self, op):
        M = np.zeros((self.nbf, self.nbf))
        lib().oracle_onee(self.h, op, _p(M))
        return M

    def build_JK(self, P, alpha=1.0, beta=0.0, omega=0.0):
        P = np.ascontiguousarray(P, dtype=np.float64)
        ndm = P.size // (self.nbf * self.nbf)
        J = np.zeros_like(P)
        K = np.zeros_like(P)
        st = np.zeros(4)
        rc = lib().oracle_build_jk(self.h, _p(P), _p(J), _p(K), alpha, beta, ndm, _p(st))
        if rc != 0:
            raise RuntimeError("Density matrix contains NaN/Inf values.")
        self.last_stats = dict(quartets=st[0], prim_quartets=st[1], func_quartets=st[2], loop_quartets=st[3])
        return J, K

    def build_JK_sample(self, P, stride, offset=0):
        """build_JK over bra shells s1 % stride == offset only (timing sample)."""
        lib().oracle_set_row_sample(stride, offset)
        try:
            return self.build_JK(P)
        finally:
            lib().oracle_set_row_sample(1, 0)

    def census(self, P, significant=False):
        """Surviving shell quartets and their primitive quartets per class
        [l1, l2, l3, l4] (the reference's loop orientation) of the screened
        canonical loop (libint2_direct.cpp:357-402), screened with the norms
        of P's first nbf x nbf block.  significant=True also returns the
        primitive quartets over significant primitive pairs (|c_a c_b|
        exp(-ab/p AB^2) >= DBL_EPSILON / 1e10, the precision the reference
        builds its shell pairs with, :29,87,146)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        nq = np.zeros(625)
        npq = np.zeros(625)
        nsig = np.zeros(625)
        lib().oracle_census(self.h, _p(P), _p(nq), _p(npq), _p(nsig) if significant else None)
        out = (nq.reshape(5, 5, 5, 5), npq.reshape(5, 5, 5, 5))
        return out + (nsig.reshape(5, 5, 5, 5),) if significant else out

    def jk_rows(self, P, shells, alpha=1.0, beta=0.0):
        """Exact rows of build_JK's (J, K) for the functions of `shells` (in
        that order), each (nfunctions, nbf): every ordered quartet with a row
        shell in the first position, screened as its canonical representative
        is, so the rows equal those of the full screened build (one density)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        if P.shape != (self.nbf, self.nbf):
            raise ValueError("jk_rows takes one nbf x nbf density")
        rows = np.ascontiguousarray(shells, dtype=np.int32)
        off = self.basis.shell_offsets
        nf = int(sum(off[s + 1] - off[s] for s in rows))
        J = np.zeros((nf, self.nbf))
        K = np.zeros((nf, self.nbf))
        if lib().oracle_jk_rows(self.h, _p(P), len(rows), _p(rows, ctypes.c_int), _p(J), _p(K), alpha, beta) != 0:
            raise RuntimeError("Density matrix contains NaN/Inf values.")
        return J, K

    def build_JK_incore(self, P, alpha=1.0, beta=0.0):
        P = np.ascontiguousarray(P, dtype=np.float64)
        J = np.zeros_like(P)
        K = np.zeros_like(P)
        rc = lib().oracle_jk_incore(self.h, _p(P), _p(J), _p(K), alpha, beta)
        if rc != 0:
            raise MemoryError("system too large for the full-tensor oracle")
        return J, K

    def eri_full(self):
        """The dense AO tensor T[i, j, k, l] = (ij|kl), every quartet computed."""
        n = self.nbf
        T = np.zeros((n, n, n, n))
        if lib().oracle_eri_full(self.h, _p(T)) != 0:
            raise RuntimeError("oracle_eri_full failed")
        return T

    def get_gradients(self, P, alpha=1.0, beta=0.0, mode=1, class_key=-1, centre=-1):
        """(dJ, dK), each 3*natoms xyz-major (libint2_direct.cpp:824), for the
        RHF total density or a spin pair [P_alpha, P_beta]; class_key
        (1000 la + 100 lb + 10 lc + ld) restricts the sum to one ERI class,
        centre (0..2) to the derivative by that centre moved against centre D."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        ndm = P.size // (self.nbf * self.nbf)
        lib().oracle_set_grad_class(class_key)
        lib().oracle_set_grad_centre(centre)
        na = self.structure.natoms
        dJ = np.zeros(3 * na)
        dK = np.zeros(3 * na)
        try:
            rc = lib().oracle_gradients(self.h, _p(P), _p(dJ), _p(dK), alpha, beta, mode, ndm)
        finally:
            lib().oracle_set_grad_class(-1)
            lib().oracle_set_grad_centre(-1)
        if rc != 0:
            raise RuntimeError("Density matrix contains NaN/Inf values.")
        return dJ, dK


def quarter_trans(T, C):
    """out[i, j, k, p] = sum_l (ij|kl) C[l, p]: the first quarter of the MO
    transform, nbf^3 x nt with p fastest (ERI::quarter_trans, eri.h:116; the
    in-core form incore_impl.cpp:286-297 is this one GEMM over the dense
    tensor; libint2_direct.cpp:913-1141 accumulates it quartet by quartet)."""
    n = T.shape[0]
    return (T.reshape(n * n * n, n) @ C).reshape(n, n, n, C.shape[1])


def mo_eri(T, Ci, Cj, Ck, Cl):
    """The four-quarter transform of MOERI::compute (moeri.cpp:28-153, CPU
    path :108-152): tmp1 = quarter_trans(Ci); the 2nd quarter contracts the
    third AO index with Cj (:125-133), the 3rd the second with Ck (:135-144),
    the 4th the first with Cl (:146-151).  Row-major result
    out[w, x, y, z] = sum (ab|cd) Cl[a, w] Ck[b, x] Cj[c, y] Ci[d, z]."""
    t1 = quarter_trans(T, Ci)                           # [a, b, c, z]
    t2 = np.einsum("abcz,cy->abyz", t1, Cj, optimize=True)
    t3 = np.einsum("abyz,bx->axyz", t2, Ck, optimize=True)
    return np.einsum("axyz,aw->wxyz", t3, Cl, optimize=True)


def rmp2_correlation(ovov, eps_occ, eps_vir):
    """Closed-shell MP2 correlation energy from (ia|jb) = ovov[i, a, j, b]:
    sum (ia|jb) [2 (ia|jb) - (ib|ja)] / (e_i + e_j - e_a - e_b)."""
    d = (eps_occ[:, None, None, None] - eps_vir[None, :, None, None]
         + eps_occ[None, None, :, None] - eps_vir[None, None, None, :])
    return float(np.sum(ovov * (2.0 * ovov - ovov.transpose(0, 3, 2, 1)) / d))


def orthogonalizer(S, lindep=-1.0):
    s, U = np.linalg.eigh(S)
    thr = lindep if lindep >= 0 else s.max() / 1e9
    keep = s >= thr
    return U[:, keep] / np.sqrt(s[keep])


def rhf(structure, basis, eri=None, conv_energy=1e-10, conv_grad=1e-8, max_iter=100, P0=None,
        jk=None, verbose=False):
    """Plain RHF with DIIS on the oracle's integrals.  Returns dict with the
    energy, density, Fock and orbital data."""
    if eri is None:
        eri = OracleERI(basis, structure)
    jk = jk or eri.build_JK
    S = eri.onee(0)
    H = eri.onee(1) + eri.onee(2)
    X = orthogonalizer(S)
    nocc = structure.nelectrons // 2
    Enuc = structure.nuclear_repulsion()
    if P0 is None:
        e, Cp = np.linalg.eigh(X.T @ H @ X)
        C = X @ Cp
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    else:
        P = P0.copy()
    Fs, Es = [], []
    E_old = 0.0
    for it in range(max_iter):
        J, K = jk(P)
        F = H + J - 0.5 * K
        E = Enuc + np.sum(P * H) + 0.5 * np.sum(P * (F - H))
        err = F @ P @ S - S @ P @ F
        gnorm = np.max(np.abs(err))
        if verbose:
            print(f"it {it:3d} E={E:.12f} dE={E - E_old:+.3e} |g|={gnorm:.3e}")
        if abs(E - E_old) < conv_energy and gnorm < conv_grad:
            break
        E_old = E
        Fs.append(F)
        Es.append(err)
        if len(Fs) > 10:
            Fs.pop(0)
            Es.pop(0)
        m = len(Fs)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0.0
        for i in range(m):
            for j in range(m):
                B[i, j] = np.sum(Es[i] * Es[j])
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            c = np.linalg.solve(B, rhs)[:m]
        except np.linalg.LinAlgError:
            c = np.zeros(m)
            c[-1] = 1.0
        Fd = sum(ci * Fi for ci, Fi in zip(c, Fs))
        e, Cp = np.linalg.eigh(X.T @ Fd @ X)
        C = X @ Cp
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    else:
        raise RuntimeError("oracle RHF did not converge")
    return dict(energy=E, P=P, F=F, H=H, S=S, C=C, eps=e, iterations=it + 1, eri=eri)
