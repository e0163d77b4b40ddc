"""CPU oracle for the fock engine -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package, and only as the checker (or the timed CPU baseline). The
product library (fock-rs_amd/) never imports, links or calls anything here.

* ``refcpu.c``  -- statement-by-statement C restatement of the reference's
  integral code (boys, basis, electron_tensor); see its header for citations.
* ``oracle64.c`` -- independent accurate, unscreened FP64 MD integrals.
* ``refscf.py`` -- numpy restatement of the reference SCF driver
  (scf/src/lib.rs:73-258, scf/src/utils.rs:25-138).

Functions without an independent pin are listed as "parity unpinned" in the
module that defines them (refscf.py: the eigensolver and matrix inverse of the
reference's third-party dependencies).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Basis(ctypes.Structure):
    """Flat basis view shared by refcpu.c (ref_basis) and oracle64.c (o64_basis)."""
    _fields_ = [("n", ctypes.c_int), ("centers", _dp), ("angular", _ip), ("prim_off", _ip),
                ("exps", _dp), ("coefs", _dp)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i, ll = ctypes.c_double, ctypes.c_int, ctypes.c_longlong
        bp = ctypes.POINTER(Basis)
        sig = {
            "ref_set_policy": (None, [i, i]),
            "ref_boys_order_error": (i, []),
            "ref_powi": (d, [d, i]),
            "ref_boys": (d, [d, i]),
            "ref_hermite_expansion": (d, [i, i, i, d, d, d]),
            "ref_coulomb_auxiliary": (d, [i, i, i, i, d, _dp, d]),
            "ref_overlap_int": (d, [bp, i, i]),
            "ref_kinetic_int": (d, [bp, i, i]),
            "ref_nuclear_int": (d, [bp, i, i, i, _dp, _dp]),
            "ref_eri": (d, [bp, i, i, i, i]),
            "ref_one_electron": (None, [bp, i, _dp, _dp, _dp, _dp, _dp]),
            "ref_packed_size": (ctypes.c_size_t, [i]),
            "ref_electron_tensor": (ll, [bp, i, _dp]),
            "ref_unpack_dense": (None, [i, _dp, _dp]),
            "ref_electron_terms": (None, [i, _dp, _dp]),
            "ref_g_build": (None, [i, _dp, _dp, _dp]),
            "ref_g_build_packed": (None, [i, _dp, _dp, _dp]),
            "ref_eri_sample": (ll, [bp, _dp, ll, ctypes.c_ulonglong, _dp]),
            "ref_eri_diagonal": (None, [bp, _dp]),
            "ref_num_threads": (i, []),
            "ref_set_num_threads": (None, [i]),
            "o64_boys": (d, [d, i]),
            "o64_boys_array": (None, [_dp, i, i, _dp]),
            "o64_eri_packed": (i, [bp, _dp]),
            "o64_shell_quartet": (i, [bp, i, i, i, i, _dp]),
            "o64_jk_packed": (None, [i, _dp, _dp, _dp, _dp]),
            "o64_one_electron": (None, [bp, i, _dp, _dp, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


class BasisView:
    """Keeps the numpy arrays alive while C code holds pointers into them."""

    def __init__(self, fb):
        self.fb = fb
        self.s = Basis(fb.n, dptr(fb.centers), iptr(fb.angular), iptr(fb.prim_offset),
                       dptr(fb.exponents), dptr(fb.coefficients))

    @property
    def ref(self):
        return ctypes.byref(self.s)
