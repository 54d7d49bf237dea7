"""ERI engine interface of the SCF library, MI355X implementation.

Mirrors qdk::chemistry::scf::ERI (scf/include/qdk/chemistry/scf/core/eri.h:
23-176): `build_JK(P, alpha, beta, omega)` (:79-80) and
`get_gradients(P, alpha, beta, omega)` (:100-101), created by `create`
switching on `ERIConfig.method` like ERI::create (scf/src/eri/eri.cpp:36-68).
The engine fills the `ERIMethod.Rys` slot (core/enums.h:22-35); the
reference's other methods (Libint2Direct, Incore, SnK) are CPU/third-party
engines outside this package, and requesting them raises, as an unknown
method does in the reference (eri.cpp:65).
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib


class ERIMethod(enum.IntEnum):
    """Same values as core/enums.h:22-35."""
    Libint2Direct = 0
    Rys = 1
    HGP = 2
    Incore = 3
    SnK = 4


@dataclass
class ERIConfig:
    """Field names of ERIConfig (core/scf.h:65-94) plus the engine knobs."""
    method: ERIMethod = ERIMethod.Rys
    eri_threshold: float = 1e-12
    shell_pair_threshold: float = 1e-12
    # False (the reference's default, core/scf.h:81-82): bitwise-repeatable
    # J/K through order-independent fixed-point sums; True: FP64 atomics
    use_atomics: bool = False
    is_mix_precision: bool = False
    mix_precision_threshold: float = 1e-6
    device_id: int = 0
    verbose: int = 0
    profile: bool = False


class RysERI:
    """GPU Fock-build engine (libqfe.so) for one basis/structure."""

    def __init__(self, basis, structure, cfg: ERIConfig | None = None):
        cfg = cfg or ERIConfig()
        self.cfg = cfg
        self.basis = basis
        self.structure = structure
        self.nbf = basis.nbf
        self._b = _lib.BasisArrays(basis, structure)
        c = _lib.QfeConfig(cfg.eri_threshold, cfg.shell_pair_threshold, cfg.device_id,
                           int(cfg.is_mix_precision), cfg.mix_precision_threshold, cfg.verbose, int(cfg.profile),
                           int(cfg.use_atomics))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().qfe_create(ctypes.byref(self._b.c), ctypes.byref(c), ctypes.byref(h)))
        self.h = h
        self.mixed = bool(cfg.is_mix_precision)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            _lib.lib().qfe_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_JK(self, P, alpha=1.0, beta=0.0, omega=0.0, J=True, K=True):
        """(J, K) for a host density (ndm x nbf x nbf or nbf x nbf).  J=False
        or K=False skips that matrix (None is returned in its place), as the
        reference's build_JK does for a null output (eri_multiplexer.cpp:66-67)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        ndm = P.size // (self.nbf * self.nbf)
        if ndm * self.nbf * self.nbf != P.size or ndm not in (1, 2):
            raise ValueError("build_JK takes 1 or 2 densities of nbf x nbf")
        Jo = np.empty_like(P) if J else None
        Ko = np.empty_like(P) if K else None
        _lib.check(_lib.lib().qfe_build_jk(self.h, _lib.ptr(P), _lib.ptr(Jo) if J else None,
                                           _lib.ptr(Ko) if K else None, alpha, beta, omega, ndm))
        return Jo, Ko

    def device_buffers(self):
        """Engine-owned HBM buffers (P, J, K) as raw device addresses."""
        p, j, k = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.lib().qfe_device_buffers(self.h, ctypes.byref(p), ctypes.byref(j), ctypes.byref(k)))
        return p.value, j.value, k.value

    def upload(self, dst, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        _lib.check(_lib.lib().qfe_copy(self.h, ctypes.c_void_p(dst), host.ctypes.data_as(ctypes.c_void_p),
                                       host.nbytes, 0))

    def download(self, src, shape):
        out = np.empty(shape, dtype=np.float64)
        _lib.check(_lib.lib().qfe_copy(self.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src),
                                       out.nbytes, 1))
        return out

    def build_JK_device(self, dP, dJ, dK, alpha=1.0, beta=0.0, omega=0.0, stream=None):
        """Device-resident variant on raw HIP device addresses (ints)."""
        s = ctypes.c_void_p(stream) if stream is not None else None
        _lib.check(_lib.lib().qfe_build_jk_device(self.h, ctypes.c_void_p(dP),
                                                  ctypes.c_void_p(dJ) if dJ is not None else None,
                                                  ctypes.c_void_p(dK) if dK is not None else None,
                                                  alpha, beta, omega, s))

    def get_gradients(self, P, alpha=1.0, beta=0.0, omega=0.0, dJ=True, dK=True):
        """(dJ, dK): 3*natoms each, xyz-major (libint2_direct.cpp:824).  P is
        the RHF total density or the spin pair [P_alpha, P_beta] (Coulomb term
        of P_alpha + P_beta, exchange summed over the spins with scale -1/2,
        :683-692,866-886).  dJ=False or dK=False skips that term (None in its
        place), as a null output does in the reference."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        ndm = P.size // (self.nbf * self.nbf)
        if ndm * self.nbf * self.nbf != P.size or ndm not in (1, 2):
            raise ValueError("get_gradients takes 1 or 2 densities of nbf x nbf")
        na = self.structure.natoms
        gJ = np.zeros(3 * na) if dJ else None
        gK = np.zeros(3 * na) if dK else None
        _lib.check(_lib.lib().qfe_gradients(self.h, _lib.ptr(P), _lib.ptr(gJ) if dJ else None,
                                            _lib.ptr(gK) if dK else None, alpha, beta, omega, ndm))
        return gJ, gK

    def quarter_trans(self, C):
        """out[i, j, k, p] = sum_l (ij|kl) C[l, p] for C (nbf x nt): the first
        quarter of the MO transform (ERI::quarter_trans, eri.h:116), on the
        GPU over the dense AO tensor in HBM."""
        C = np.ascontiguousarray(C, dtype=np.float64)
        n = self.basis.nbf
        if C.ndim != 2 or C.shape[0] != n:
            raise ValueError("quarter_trans takes C of shape (nbf, nt)")
        nt = C.shape[1]
        out = np.empty((n, n, n, nt))
        _lib.check(_lib.lib().qfe_quarter_trans(self.h, nt, _lib.ptr(C), _lib.ptr(out)))
        return out

    def set_precision(self, mix: bool, threshold: float | None = None):
        """Mixed (FP32 for Q_ab Q_cd < threshold) or all-FP64 builds from now on."""
        thr = self.cfg.mix_precision_threshold if threshold is None else threshold
        _lib.check(_lib.lib().qfe_set_precision(self.h, int(mix), thr))
        self.mixed = bool(mix)

    def class_stats(self, precision=0, executed=False):
        """[((la, lb, lc, ld), quartets, prim_quartets, ms)] of the last build,
        for the FP64 (0) or FP32 (1) tasks; prim_quartets is the census over
        significant primitive pairs, or with executed=True (profiled builds)
        the primitive quartets the kernels evaluated after primitive
        screening."""
        cap = 256
        cls = np.zeros(4 * cap, dtype=np.int32)
        q, pq, epq, ms = np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(cap)
        # the executed count runs a counting kernel over the last task list:
        # only when asked for
        n = _lib.lib().qfe_class_stats(self.h, precision, cap, _lib.ptr(cls, ctypes.c_int), _lib.ptr(q),
                                       _lib.ptr(pq), _lib.ptr(epq) if executed else None, _lib.ptr(ms))
        p = epq if executed else pq
        return [(tuple(int(x) for x in cls[4 * i:4 * i + 4]), float(q[i]), float(p[i]), float(ms[i]))
                for i in range(min(n, cap))]

    def stats(self) -> dict:
        s = _lib.QfeStats()
        _lib.check(_lib.lib().qfe_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.QfeStats._fields_}


class DFJ:
    """Density-fitted Coulomb engine on the GPU (the J of the reference's
    INCORE DF engine, incore_impl_df.cpp): (mu nu|P) over the auxiliary
    basis and the Cholesky-factored metric are built once; build_JK returns
    J = B V^-1 B^T vec(P) and a zero K, for use as the J engine of an
    ERIMultiplexer (eri_multiplexer.cpp:34-71)."""

    def __init__(self, basis, aux_basis, structure, device_id: int = 0):
        self.nbf = basis.nbf
        self._b = _lib.BasisArrays(basis, structure)
        self._a = _lib.BasisArrays(aux_basis, structure)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().qfe_dfj_create(ctypes.byref(self._b.c), ctypes.byref(self._a.c), device_id,
                                             ctypes.byref(h)))
        self.h = h
        self.naux = int(_lib.lib().qfe_dfj_naux(h))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            _lib.lib().qfe_dfj_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_JK(self, P, alpha=1.0, beta=0.0, omega=0.0, J=True, K=True):
        """J of the fitted density; K is not fitted (None when K=False, else
        zeros, as the reference's DF engine leaves it)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        if P.shape != (self.nbf, self.nbf):
            raise ValueError("DFJ.build_JK takes one nbf x nbf density")
        Jo = None
        if J:
            Jo = np.empty_like(P)
            _lib.check(_lib.lib().qfe_dfj_build_j(self.h, _lib.ptr(P), _lib.ptr(Jo)))
        return Jo, (np.zeros_like(P) if K else None)

    def get_gradients(self, P, alpha=1.0, beta=0.0, omega=0.0, dJ=True, dK=True):
        """The reference's DF engine gives DF-J gradients only and refuses an
        exchange gradient (ERI_DF::get_gradients, incore_impl_df.cpp:203-214);
        the DF-J derivative integrals are not built on the GPU here."""
        if abs(omega) > 1e-12:
            raise RuntimeError("ERIINCORE_DF gradients Does Not Support Range-Separated Hybrids")
        if dK and abs(alpha) + abs(beta) > 1e-12:
            raise RuntimeError("ERIINCORE_DF + Exchange gradients is Not Yet Implemented")
        if not dJ:
            raise RuntimeError("ERIINCORE_DF is only valid for DF-J gradients")
        raise NotImplementedError("DF-J nuclear gradients are not built in the MI355X engine")


class MOERI:
    """Four-index MO transform over an engine (MOERI, moeri.cpp:28-153):
    compute(nb, nt, Ci, Cj, Ck, Cl) returns out[w, x, y, z] =
    sum (ab|cd) Cl[a, w] Ck[b, x] Cj[c, y] Ci[d, z], each C of shape (nb, nt);
    with one C for all four, the MO integrals (pq|rs)."""

    def __init__(self, eri: RysERI):
        self.eri = eri

    def compute(self, nb, nt, Ci, Cj, Ck, Cl):
        Cs = [np.ascontiguousarray(C, dtype=np.float64) for C in (Ci, Cj, Ck, Cl)]
        for C in Cs:
            if C.shape != (nb, nt):
                raise ValueError("MOERI.compute takes four C of shape (nb, nt)")
        out = np.empty((nt, nt, nt, nt))
        _lib.check(_lib.lib().qfe_mo_eri(self.eri.h, nb, nt, *[_lib.ptr(C) for C in Cs], _lib.ptr(out)))
        return out


class ERIMultiplexer:
    """Routes J and K to one engine, or J and K to two engines (J with omega 0),
    and gradients to a third when given; restates ERIMultiplexer
    (eri_multiplexer.cpp:10-94): a single engine is called once for both, two
    engines each build only their own matrix (the other output null), and
    gradients follow the same split unless a gradient engine was given."""

    def __init__(self, j_engine, k_engine=None, grad_engine=None):
        self.j_impl = j_engine
        self.k_impl = k_engine if k_engine is not None else j_engine
        self.grad_impl = grad_engine if grad_engine is not None else j_engine

    def build_JK(self, P, alpha=1.0, beta=0.0, omega=0.0):
        if self.j_impl is self.k_impl:
            return self.j_impl.build_JK(P, alpha, beta, omega)
        J, _ = self.j_impl.build_JK(P, alpha, beta, 0.0, K=False)
        _, K = self.k_impl.build_JK(P, alpha, beta, omega, J=False)
        return J, K

    def get_gradients(self, P, alpha=1.0, beta=0.0, omega=0.0):
        if self.grad_impl is not self.j_impl:
            return self.grad_impl.get_gradients(P, alpha, beta, omega)
        if self.j_impl is self.k_impl:
            return self.j_impl.get_gradients(P, alpha, beta, omega)
        dJ, _ = self.j_impl.get_gradients(P, alpha, beta, 0.0, dK=False)
        _, dK = self.k_impl.get_gradients(P, alpha, beta, omega, dJ=False)
        return dJ, dK


def compiled_max_l() -> int:
    """Highest shell angular momentum the ERI kernels of this build handle."""
    L = _lib.lib()
    if not hasattr(L, "qfe_max_l"):  # libraries built before the query existed compile s, p, d
        return 2
    return int(L.qfe_max_l())


def create(basis, structure, cfg: ERIConfig | None = None):
    """ERI::create (eri.cpp:36-68): dispatch on cfg.method."""
    cfg = cfg or ERIConfig()
    if cfg.method == ERIMethod.Rys:
        return RysERI(basis, structure, cfg)
    raise ValueError(f"Invalid ERI method {cfg.method!r} for the MI355X engine")


def one_body(basis, structure, op: int) -> np.ndarray:
    """S (0), T (1) or V (2) on the host."""
    b = _lib.BasisArrays(basis, structure)
    out = np.zeros((basis.nbf, basis.nbf))
    _lib.check(_lib.lib().qfe_one_body(ctypes.byref(b.c), op, _lib.ptr(out)))
    return out


def one_body_grad(basis, structure, op: int, W) -> np.ndarray:
    """sum_mn W_mn dO_mn/dR, natoms x 3."""
    b = _lib.BasisArrays(basis, structure)
    W = np.ascontiguousarray(W, dtype=np.float64)
    g = np.zeros((structure.natoms, 3))
    _lib.check(_lib.lib().qfe_one_body_grad(ctypes.byref(b.c), op, _lib.ptr(W), _lib.ptr(g)))
    return g


def eigh_gpu(A, device_id: int = 0):
    """(w, V) of a symmetric matrix on the GPU (rocSOLVER dsyevd): w ascending,
    eigenvector j in column j, as numpy.linalg.eigh."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("eigh_gpu needs a square matrix")
    w = np.zeros(n)
    V = np.zeros((n, n))
    _lib.check(_lib.lib().qfe_eigh_gpu(n, _lib.ptr(A), _lib.ptr(w), _lib.ptr(V), device_id))
    return w, V


def one_body_gpu(basis, structure, op: int, device_id: int = 0) -> np.ndarray:
    """S (0), T (1) or V (2) computed on the GPU (OneBodyIntegral::
    overlap_integral / kinetic_integral / nuclear_integral, int1e.h)."""
    b = _lib.BasisArrays(basis, structure)
    out = np.zeros((basis.nbf, basis.nbf))
    _lib.check(_lib.lib().qfe_one_body_gpu(ctypes.byref(b.c), op, _lib.ptr(out), device_id))
    return out


def one_body_grad_gpu(basis, structure, op: int, W, device_id: int = 0) -> np.ndarray:
    """sum_mn W_mn dO_mn/dR on the GPU, natoms x 3 (the *_integral_deriv
    contractions of int1e.h)."""
    b = _lib.BasisArrays(basis, structure)
    W = np.ascontiguousarray(W, dtype=np.float64)
    g = np.zeros((structure.natoms, 3))
    _lib.check(_lib.lib().qfe_one_body_grad_gpu(ctypes.byref(b.c), op, _lib.ptr(W), _lib.ptr(g), device_id))
    return g
