"""ctypes binding of libqfe.so (include/qfe.h).

The HIP library is required: importing the engine without it raises, there is
no CPU fallback for the Fock build.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libqfe.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

QFE_OK, QFE_ERR_INVALID, QFE_ERR_RUNTIME, QFE_ERR_DEVICE, QFE_ERR_UNSUPPORTED = 0, 1, 2, 3, 4


class QfeBasis(ctypes.Structure):
    _fields_ = [("nshells", ctypes.c_int), ("l", _ip), ("atom", _ip), ("centers", _dp),
                ("prim_offset", _ip), ("exps", _dp), ("coefs", _dp), ("pure", ctypes.c_int),
                ("natoms", ctypes.c_int), ("charges", _dp), ("coords", _dp)]


class QfeConfig(ctypes.Structure):
    _fields_ = [("eri_threshold", ctypes.c_double), ("shell_pair_threshold", ctypes.c_double),
                ("device_id", ctypes.c_int), ("mix_precision", ctypes.c_int),
                ("mix_threshold", ctypes.c_double), ("verbose", ctypes.c_int), ("profile", ctypes.c_int),
                ("use_atomics", ctypes.c_int)]


class QfeStats(ctypes.Structure):
    _fields_ = [("quartets", ctypes.c_longlong), ("quartets_fp32", ctypes.c_longlong),
                ("prim_quartets", ctypes.c_longlong),
                ("ms_screen", ctypes.c_double), ("ms_eri", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("significant_pairs", ctypes.c_longlong)]


EXPORTS = ["qfe_create", "qfe_build_jk", "qfe_build_jk_device", "qfe_gradients", "qfe_one_body",
           "qfe_one_body_grad", "qfe_one_body_gpu", "qfe_one_body_grad_gpu", "qfe_eigh_gpu",
           "qfe_dfj_create", "qfe_dfj_build_j", "qfe_dfj_naux", "qfe_dfj_destroy",
           "qfe_quarter_trans", "qfe_mo_eri", "qfe_scf_create", "qfe_scf_set_density",
           "qfe_scf_step", "qfe_scf_update", "qfe_scf_finish", "qfe_scf_get", "qfe_scf_destroy", "qfe_get_stats", "qfe_class_stats", "qfe_build_jk_device_noalloc", "qfe_noalloc_scratch_bytes", "qfe_set_precision", "qfe_device_buffers", "qfe_copy", "qfe_class_chunks", "qfe_nbf", "qfe_max_l", "qfe_destroy",
           "qfe_last_error"]

_lib = None


class QfeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libqfe.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python qdk-chemistry_amd/build.py` "
                              "(the Fock build has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        L.qfe_create.argtypes = [ctypes.POINTER(QfeBasis), ctypes.POINTER(QfeConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.qfe_build_jk.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int]
        L.qfe_build_jk_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.qfe_build_jk_device_noalloc.argtypes = L.qfe_build_jk_device.argtypes
        L.qfe_noalloc_scratch_bytes.argtypes = []
        L.qfe_noalloc_scratch_bytes.restype = ctypes.c_ulonglong
        L.qfe_gradients.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int]
        L.qfe_one_body.argtypes = [ctypes.POINTER(QfeBasis), ctypes.c_int, _dp]
        L.qfe_one_body_grad.argtypes = [ctypes.POINTER(QfeBasis), ctypes.c_int, _dp, _dp]
        L.qfe_one_body_gpu.argtypes = [ctypes.POINTER(QfeBasis), ctypes.c_int, _dp, ctypes.c_int]
        L.qfe_one_body_grad_gpu.argtypes = [ctypes.POINTER(QfeBasis), ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.qfe_eigh_gpu.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.qfe_dfj_create.argtypes = [ctypes.POINTER(QfeBasis), ctypes.POINTER(QfeBasis), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p)]
        L.qfe_dfj_build_j.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.qfe_dfj_naux.argtypes = [ctypes.c_void_p]
        L.qfe_dfj_destroy.argtypes = [ctypes.c_void_p]
        L.qfe_dfj_destroy.restype = None
        L.qfe_quarter_trans.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, _dp, _dp]
        L.qfe_mo_eri.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_ulonglong, _dp, _dp, _dp, _dp, _dp]
        L.qfe_scf_create.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.qfe_scf_set_density.argtypes = [ctypes.c_void_p, _dp]
        L.qfe_scf_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.qfe_scf_update.argtypes = [ctypes.c_void_p]
        L.qfe_scf_finish.argtypes = [ctypes.c_void_p, _dp]
        L.qfe_scf_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.qfe_scf_destroy.argtypes = [ctypes.c_void_p]
        L.qfe_scf_destroy.restype = None
        L.qfe_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(QfeStats)]
        L.qfe_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.qfe_device_buffers.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
        L.qfe_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_int]
        L.qfe_class_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        L.qfe_class_chunks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip]
        L.qfe_nbf.argtypes = [ctypes.c_void_p]
        L.qfe_destroy.argtypes = [ctypes.c_void_p]
        L.qfe_destroy.restype = None
        L.qfe_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != QFE_OK:
        raise QfeError(rc, lib().qfe_last_error().decode())


def ptr(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


class BasisArrays:
    """Keeps the numpy arrays alive behind a QfeBasis view."""

    def __init__(self, basis, structure):
        l, atom, cen, po, exps, coefs = basis.flat()
        self.arrays = dict(l=l, atom=atom, cen=cen, po=po, exps=exps, coefs=coefs,
                           Z=np.ascontiguousarray(basis.effective_charges(structure.numbers)),
                           R=np.ascontiguousarray(structure.coords.astype(np.float64)))
        a = self.arrays
        self.c = QfeBasis(basis.nsh, ptr(a["l"], ctypes.c_int), ptr(a["atom"], ctypes.c_int), ptr(a["cen"]),
                          ptr(a["po"], ctypes.c_int), ptr(a["exps"]), ptr(a["coefs"]), int(basis.pure),
                          structure.natoms, ptr(a["Z"]), ptr(a["R"]))
