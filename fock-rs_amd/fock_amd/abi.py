"""ctypes binding of the C ABI in include/fock.h (libfock.so).

The library is the only compute path: if libfock.so is missing or no HIP device is
present, every call raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfock.so")

FOCK_OK = 0
FOCK_NOT_CONVERGED = 1
FOCK_DIVERGED = 2
FOCK_INVALID_ARG = 3
FOCK_BOYS_ORDER = 4
FOCK_HIP_ERROR = 5
FOCK_OOM = 6
FOCK_UNSUPPORTED = 7
STATUS_NAMES = {0: "OK", 1: "NOT_CONVERGED", 2: "DIVERGED", 3: "INVALID_ARG", 4: "BOYS_ORDER",
                5: "HIP_ERROR", 6: "OOM", 7: "UNSUPPORTED"}

POLICY_REF_COMPAT = 0
POLICY_ACCURATE = 1
SCREEN_REF_BF_PAIR = 0
SCREEN_SCHWARZ = 1
SCREEN_NONE = 2
PRECISION_FP64 = 0
PRECISION_MIXED = 1
TILES_FP64 = 0
TILES_MIXED = 1
TILES_BATCH = 2
SCHEDULE_COST = 0
SCHEDULE_XCD = 1
TILE_CHOICES = (1, 2, 4, 8, 16, 32, 64)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

EXPORTED = [
    "fock_version", "fock_ctx_create", "fock_ctx_destroy", "fock_ctx_last_error", "fock_ctx_set_stream",
    "fock_scf_opts_default", "fock_scf_run", "fock_rhf_gradient", "fock_scf_batch_run", "fock_boys", "fock_eri_dense", "fock_one_electron",
    "fock_jk_prepare", "fock_jk_build", "fock_jk_build_device", "fock_jk_get_stats", "fock_gemm",
    "fock_sorted_eigs", "fock_diis", "fock_jk_set_timing", "fock_jk_get_timings",
    "fock_jk_autotune", "fock_tiles_get", "fock_tiles_set", "fock_sorted_eigs_batch",
    "fock_mo_grid", "fock_mo_grid_device", "fock_mo_points", "fock_mo_grid_device_noalloc",
]


class BasisDesc(ctypes.Structure):
    _fields_ = [("n_functions", ctypes.c_int32), ("centers", _dp), ("angular", _ip), ("prim_offset", _ip),
                ("exponents", _dp), ("coefficients", _dp)]


class NucleiDesc(ctypes.Structure):
    _fields_ = [("n_atoms", ctypes.c_int32), ("positions", _dp), ("charges", _dp)]


class ScfOpts(ctypes.Structure):
    _fields_ = [("max_iters", ctypes.c_int32), ("epsilon", ctypes.c_double), ("molecule_charge", ctypes.c_int32),
                ("integral_policy", ctypes.c_int32), ("screen", ctypes.c_int32),
                ("schwarz_threshold", ctypes.c_double), ("precision", ctypes.c_int32),
                ("incremental", ctypes.c_int32), ("initial_density", _dp), ("diis_start", ctypes.c_int32),
                ("diis_window", ctypes.c_int32), ("level_shift", ctypes.c_double),
                ("level_shift_until", ctypes.c_double), ("autotune", ctypes.c_int32),
                ("ediis", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2), ("ediis_until", ctypes.c_double),
                ("damping", ctypes.c_double), ("damping_until", ctypes.c_double)]


class ScfIter(ctypes.Structure):
    _fields_ = [("total_energy", ctypes.c_double), ("error", ctypes.c_double), ("fock_build_ms", ctypes.c_double),
                ("quartets", ctypes.c_int64)]


class ScfOut(ctypes.Structure):
    _fields_ = [("coefficients", _dp), ("density", _dp), ("orbital_energies", _dp),
                ("electronic_energy", ctypes.c_double), ("total_energy", ctypes.c_double),
                ("iterations", ctypes.c_int32), ("n_electrons", ctypes.c_int32), ("n_basis", ctypes.c_int32),
                ("last_error", ctypes.c_double), ("history", ctypes.POINTER(ScfIter)),
                ("history_capacity", ctypes.c_int32), ("history_len", ctypes.c_int32),
                ("scf_wall_ms", ctypes.c_double)]


class JKStats(ctypes.Structure):
    _fields_ = [("n_basis", ctypes.c_int32), ("n_shells", ctypes.c_int32), ("n_shell_pairs", ctypes.c_int32),
                ("shell_quartets", ctypes.c_int64 * 6), ("primitive_quartets", ctypes.c_int64 * 6),
                ("integrals", ctypes.c_int64 * 6), ("tasks", ctypes.c_int64 * 6),
                ("algorithmic_flop", ctypes.c_double * 6), ("near_threshold", ctypes.c_int64),
                ("preparations", ctypes.c_int64), ("list_checksum", ctypes.c_uint64)]


class GridDesc(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_double * 3), ("step", ctypes.c_double * 3), ("dims", ctypes.c_int32 * 3)]


class TuneReport(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int32 * 6), ("ms", (ctypes.c_double * 7) * 6), ("schedule", ctypes.c_int32),
                ("schedule_ms", ctypes.c_double * 2), ("tune_ms", ctypes.c_double)]


class FockError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, str(status)), message))
        self.status = status


_lib = None


def load(path: Optional[str] = None):
    """Load libfock.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError("libfock.so not built (%s); run __graft_entry__.build()" % path)
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    i32, i64, d = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    bp, np_ = ctypes.POINTER(BasisDesc), ctypes.POINTER(NucleiDesc)
    sig = {
        "fock_version": (ctypes.c_char_p, []),
        "fock_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "fock_ctx_destroy": (None, [vp]),
        "fock_ctx_last_error": (ctypes.c_char_p, [vp]),
        "fock_ctx_set_stream": (ctypes.c_int, [vp, vp]),
        "fock_scf_opts_default": (None, [ctypes.POINTER(ScfOpts)]),
        "fock_scf_run": (ctypes.c_int, [vp, bp, np_, ctypes.POINTER(ScfOpts), ctypes.POINTER(ScfOut)]),
        "fock_rhf_gradient": (ctypes.c_int, [vp, bp, np_, ctypes.POINTER(ScfOpts), ctypes.POINTER(ScfOut), _dp]),
        "fock_scf_batch_run": (ctypes.c_int, [vp, ctypes.c_int32, bp, np_, ctypes.POINTER(ScfOpts),
                                              ctypes.POINTER(ScfOut), ctypes.POINTER(ctypes.c_int32)]),
        "fock_boys": (ctypes.c_int, [vp, _dp, i64, i32, i32, _dp]),
        "fock_eri_dense": (ctypes.c_int, [vp, bp, i32, i32, d, _dp]),
        "fock_one_electron": (ctypes.c_int, [vp, bp, np_, i32, _dp, _dp, _dp]),
        "fock_jk_prepare": (ctypes.c_int, [vp, bp, i32, i32, d, i32]),
        "fock_jk_build": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "fock_jk_build_device": (ctypes.c_int, [vp, vp, vp, vp]),
        "fock_jk_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(JKStats)]),
        "fock_gemm": (ctypes.c_int, [vp, i32, i32, i32, i32, i32, d, _dp, i32, _dp, i32, d, _dp, i32]),
        "fock_sorted_eigs": (ctypes.c_int, [vp, i32, _dp, _dp, _dp]),
        "fock_sorted_eigs_batch": (ctypes.c_int, [vp, i32, i32, _dp, _dp, _dp]),
        "fock_diis": (ctypes.c_int, [vp, i32, i64, _dp, _dp, _dp]),
        "fock_jk_set_timing": (ctypes.c_int, [vp, i32]),
        "fock_jk_get_timings": (ctypes.c_int, [vp, _dp]),
        "fock_jk_autotune": (ctypes.c_int, [vp, _dp, ctypes.POINTER(TuneReport)]),
        "fock_tiles_get": (ctypes.c_int, [vp, i32, _ip, _ip]),
        "fock_tiles_set": (ctypes.c_int, [vp, i32, _ip, i32]),
        "fock_mo_grid": (ctypes.c_int, [vp, bp, _dp, _ip, i32, ctypes.POINTER(GridDesc), ctypes.POINTER(ctypes.c_float)]),
        "fock_mo_grid_device": (ctypes.c_int, [vp, bp, _dp, _ip, i32, ctypes.POINTER(GridDesc), vp]),
        "fock_mo_points": (ctypes.c_int, [vp, bp, _dp, i32, i64, _dp, _dp]),
        "fock_mo_grid_device_noalloc": (ctypes.c_int, [vp, bp, _dp, _ip, i32, ctypes.POINTER(GridDesc), vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def basis_desc(fb) -> BasisDesc:
    return BasisDesc(fb.n, _d(fb.centers), _i(fb.angular), _i(fb.prim_offset), _d(fb.exponents),
                     _d(fb.coefficients))


def nuclei_desc(fn) -> NucleiDesc:
    return NucleiDesc(len(fn.charges), _d(fn.positions), _d(fn.charges))


class Context:
    """Owns one fock_ctx (one HIP stream, device buffers) on a device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.ptr = ctypes.c_void_p()
        st = self.lib.fock_ctx_create(device, ctypes.byref(self.ptr))
        if st != FOCK_OK:
            raise FockError(st, "fock_ctx_create(device=%d) failed (no HIP device?)" % device)

    def close(self):
        if self.ptr:
            self.lib.fock_ctx_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int, allow=(FOCK_OK,)):
        if st not in allow:
            raise FockError(st, self.lib.fock_ctx_last_error(self.ptr).decode())
        return st

    def set_stream(self, stream_handle: int):
        self.check(self.lib.fock_ctx_set_stream(self.ptr, ctypes.c_void_p(stream_handle)))

    # operator surfaces ---------------------------------------------------------------
    def boys(self, T, n: int, policy: int = POLICY_ACCURATE) -> np.ndarray:
        T = np.ascontiguousarray(T, dtype=np.float64)
        out = np.empty_like(T)
        self.check(self.lib.fock_boys(self.ptr, _d(T), T.size, n, policy, _d(out)))
        return out

    def eri_dense(self, fb, policy: int, screen: int, tau: float = 1e-12) -> np.ndarray:
        n = fb.n
        out = np.zeros(n ** 4)
        bd = basis_desc(fb)
        self.check(self.lib.fock_eri_dense(self.ptr, ctypes.byref(bd), policy, screen, tau, _d(out)))
        return out.reshape(n, n, n, n)

    def one_electron(self, fb, fn, policy: int):
        n = fb.n
        S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
        bd, nd = basis_desc(fb), nuclei_desc(fn)
        self.check(self.lib.fock_one_electron(self.ptr, ctypes.byref(bd), ctypes.byref(nd), policy, _d(S), _d(T),
                                              _d(V)))
        return S, T, V

    def jk_prepare(self, fb, policy: int, screen: int, tau: float = 1e-12, precision: int = PRECISION_FP64):
        bd = basis_desc(fb)
        self.check(self.lib.fock_jk_prepare(self.ptr, ctypes.byref(bd), policy, screen, tau, precision))
        self.n = fb.n

    def jk_build(self, P: np.ndarray):
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = P.shape[0]
        J = np.zeros((n, n)); K = np.zeros((n, n))
        self.check(self.lib.fock_jk_build(self.ptr, _d(P), _d(J), _d(K)))
        return J, K

    def jk_build_device(self, dP: int, dJ: int, dK: int):
        self.check(self.lib.fock_jk_build_device(self.ptr, ctypes.c_void_p(dP), ctypes.c_void_p(dJ),
                                                 ctypes.c_void_p(dK)))

    def jk_set_timing(self, enable: bool = True):
        self.check(self.lib.fock_jk_set_timing(self.ptr, 1 if enable else 0))

    def jk_timings(self) -> np.ndarray:
        ms = np.zeros(6)
        self.check(self.lib.fock_jk_get_timings(self.ptr, _d(ms)))
        return ms

    def jk_autotune(self, P: np.ndarray) -> dict:
        """Tile autotuner on the prepared build at density P: the chosen tile per quartet class
        and the best time (ms) of every candidate tile (-1: not timed)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        rep = TuneReport()
        self.check(self.lib.fock_jk_autotune(self.ptr, _d(P), ctypes.byref(rep)))
        return {"tile": list(rep.tile), "ms": np.array([list(r) for r in rep.ms]), "tune_ms": rep.tune_ms}

    def tiles(self, form: int):
        """(tile per quartet class, batch schedule) of a form (TILES_*)."""
        t = np.zeros(6, dtype=np.int32)
        sc = ctypes.c_int32(0)
        self.check(self.lib.fock_tiles_get(self.ptr, form, _i(t), ctypes.byref(sc)))
        return [int(x) for x in t], sc.value

    def set_tiles(self, form: int, tiles=None, schedule: int = SCHEDULE_COST):
        t = None if tiles is None else np.ascontiguousarray(tiles, dtype=np.int32)
        self.check(self.lib.fock_tiles_set(self.ptr, form, None if t is None else _i(t), schedule))

    def jk_stats(self) -> dict:
        st = JKStats()
        self.check(self.lib.fock_jk_get_stats(self.ptr, ctypes.byref(st)))
        return {
            "n_basis": st.n_basis, "n_shells": st.n_shells, "n_shell_pairs": st.n_shell_pairs,
            "shell_quartets": list(st.shell_quartets), "primitive_quartets": list(st.primitive_quartets),
            "integrals": list(st.integrals), "tasks": list(st.tasks),
            "algorithmic_flop": list(st.algorithmic_flop), "near_threshold": st.near_threshold,
            "preparations": st.preparations, "list_checksum": st.list_checksum,
        }

    def gemm(self, A, B, transA=False, transB=False, alpha=1.0, beta=0.0, C=None):
        A = np.ascontiguousarray(A, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        m = A.shape[1] if transA else A.shape[0]
        k = A.shape[0] if transA else A.shape[1]
        n = B.shape[0] if transB else B.shape[1]
        C = np.zeros((m, n)) if C is None else np.ascontiguousarray(C, dtype=np.float64).copy()
        self.check(self.lib.fock_gemm(self.ptr, int(transA), int(transB), m, n, k, alpha, _d(A), A.shape[1], _d(B),
                                      B.shape[1], beta, _d(C), n))
        return C

    def sorted_eigs(self, M):
        """scf::utils::sorted_eigs: (vectors as columns, ascending values)."""
        M = np.asfortranarray(M, dtype=np.float64)
        n = M.shape[0]
        vec = np.zeros(n * n)
        val = np.zeros(n)
        flat = np.ascontiguousarray(M.T.ravel())  # column-major storage
        self.check(self.lib.fock_sorted_eigs(self.ptr, n, _d(flat), _d(vec), _d(val)))
        return vec.reshape(n, n).T.copy(), val

    def sorted_eigs_batch(self, Ms):
        """Batched sorted_eigs of equal-size symmetric matrices (n <= 128): lists of (vectors as
        columns, ascending values)."""
        Ms = [np.asarray(M, dtype=np.float64) for M in Ms]
        n = Ms[0].shape[0]
        flat = np.ascontiguousarray(np.concatenate([M.T.ravel() for M in Ms]))  # column-major each
        vec = np.zeros(len(Ms) * n * n)
        val = np.zeros(len(Ms) * n)
        self.check(self.lib.fock_sorted_eigs_batch(self.ptr, len(Ms), n, _d(flat), _d(vec), _d(val)))
        return [(vec[i * n * n:(i + 1) * n * n].reshape(n, n).T.copy(), val[i * n:(i + 1) * n]) for i in range(len(Ms))]

    def diis(self, errors, focks):
        """scf::utils::diis; None when the DIIS matrix is singular."""
        E = np.ascontiguousarray(np.stack([np.asarray(e, dtype=np.float64).ravel() for e in errors]))
        F = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.float64).ravel() for f in focks]))
        out = np.zeros(E.shape[1])
        st = self.lib.fock_diis(self.ptr, E.shape[0], E.shape[1], _d(E), _d(F), _d(out))
        if st == FOCK_INVALID_ARG:
            return None
        self.check(st)
        return out.reshape(np.asarray(focks[0]).shape)

    def scf_run(self, fb, fn, opts: ScfOpts, want_history: int = 256):
        n = fb.n
        C = np.zeros(n * n); P = np.zeros(n * n); eps = np.zeros(n)
        hist = (ScfIter * want_history)()
        out = ScfOut()
        out.coefficients = _d(C)
        out.density = _d(P)
        out.orbital_energies = _d(eps)
        out.history = ctypes.cast(hist, ctypes.POINTER(ScfIter))
        out.history_capacity = want_history
        bd, nd = basis_desc(fb), nuclei_desc(fn)
        st = self.lib.fock_scf_run(self.ptr, ctypes.byref(bd), ctypes.byref(nd), ctypes.byref(opts),
                                   ctypes.byref(out))
        self.check(st, allow=(FOCK_OK, FOCK_NOT_CONVERGED, FOCK_DIVERGED))
        history = [(hist[i].total_energy, hist[i].error, hist[i].fock_build_ms, hist[i].quartets)
                   for i in range(out.history_len)]
        # column-major n x n (nalgebra layout) -> numpy with [i, k] = row i, column k
        return st, {
            "coefficients": C.reshape(n, n).T.copy(), "density": P.reshape(n, n).T.copy(),
            "orbital_energies": eps, "electronic_energy": out.electronic_energy,
            "total_energy": out.total_energy, "iterations": out.iterations, "n_electrons": out.n_electrons,
            "n_basis": out.n_basis, "last_error": out.last_error, "history": history,
            "scf_wall_ms": out.scf_wall_ms,
        }

    def scf_batch_run(self, molecules, opts: ScfOpts):
        """Lockstep SCF of many (FlatBasis, FlatNuclei) molecules; returns (overall status,
        per-molecule statuses, per-molecule results as scf_run's)."""
        count = len(molecules)
        bds = (BasisDesc * count)()
        nds = (NucleiDesc * count)()
        outs = (ScfOut * count)()
        keep, arrays = [], []
        for i, (fb, fn) in enumerate(molecules):
            bd, nd = basis_desc(fb), nuclei_desc(fn)
            keep.append((bd, nd, fb, fn))
            bds[i], nds[i] = bd, nd
            n = fb.n
            C, P, eps = np.zeros(n * n), np.zeros(n * n), np.zeros(n)
            arrays.append((C, P, eps))
            outs[i].coefficients, outs[i].density, outs[i].orbital_energies = _d(C), _d(P), _d(eps)
        status = (ctypes.c_int32 * count)()
        st = self.lib.fock_scf_batch_run(self.ptr, count, bds, nds, ctypes.byref(opts), outs, status)
        self.check(st, allow=(FOCK_OK, FOCK_NOT_CONVERGED, FOCK_DIVERGED))
        results = []
        for i, (fb, fn) in enumerate(molecules):
            n = fb.n
            C, P, eps = arrays[i]
            o = outs[i]
            results.append({
                "coefficients": C.reshape(n, n).T.copy(), "density": P.reshape(n, n).T.copy(),
                "orbital_energies": eps, "electronic_energy": o.electronic_energy, "total_energy": o.total_energy,
                "iterations": o.iterations, "n_electrons": o.n_electrons, "n_basis": o.n_basis,
                "last_error": o.last_error, "scf_wall_ms": o.scf_wall_ms})
        return st, list(status), results

    # orbitals at points -------------------------------------------------------------
    @staticmethod
    def _coef_cm(coefficients) -> np.ndarray:
        """numpy [i, k] (row i, orbital k) -> the C ABI's column-major storage."""
        return np.ascontiguousarray(np.asarray(coefficients, dtype=np.float64).T).ravel()

    def mo_grid(self, fb, coefficients, orbitals, origin, step, dims) -> np.ndarray:
        """Orbitals on a grid, float32 [len(orbitals), dims[2], dims[1], dims[0]] (voxel x fastest)."""
        orb = np.ascontiguousarray(np.atleast_1d(orbitals), dtype=np.int32)
        g = GridDesc((ctypes.c_double * 3)(*origin), (ctypes.c_double * 3)(*step), (ctypes.c_int32 * 3)(*dims))
        out = np.zeros((len(orb), dims[2], dims[1], dims[0]), dtype=np.float32)
        C = self._coef_cm(coefficients)
        bd = basis_desc(fb)
        self.check(self.lib.fock_mo_grid(self.ptr, ctypes.byref(bd), _d(C), _i(orb), len(orb), ctypes.byref(g),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def mo_grid_device(self, fb, coefficients, orbitals, origin, step, dims, d_out: int):
        """fock_mo_grid into a device array (e.g. a torch tensor's data_ptr()), asynchronous."""
        orb = np.ascontiguousarray(np.atleast_1d(orbitals), dtype=np.int32)
        g = GridDesc((ctypes.c_double * 3)(*origin), (ctypes.c_double * 3)(*step), (ctypes.c_int32 * 3)(*dims))
        C = self._coef_cm(coefficients)
        bd = basis_desc(fb)
        self.check(self.lib.fock_mo_grid_device(self.ptr, ctypes.byref(bd), _d(C), _i(orb), len(orb),
                                                ctypes.byref(g), ctypes.c_void_p(d_out)))

    def mo_grid_device_noalloc(self, fb, coefficients, orbitals, origin, step, dims, d_out: int):
        """fock_mo_grid_device_noalloc: the same evaluation with no device allocation (records in
        the device's fixed 4 MiB array)."""
        orb = np.ascontiguousarray(np.atleast_1d(orbitals), dtype=np.int32)
        g = GridDesc((ctypes.c_double * 3)(*origin), (ctypes.c_double * 3)(*step), (ctypes.c_int32 * 3)(*dims))
        C = self._coef_cm(coefficients)
        bd = basis_desc(fb)
        self.check(self.lib.fock_mo_grid_device_noalloc(self.ptr, ctypes.byref(bd), _d(C), _i(orb), len(orb),
                                                        ctypes.byref(g), ctypes.c_void_p(d_out)))

    def mo_points(self, fb, coefficients, orbital: int, points) -> np.ndarray:
        """One orbital at points [m, 3] (Bohr), float64."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        out = np.zeros(len(pts))
        C = self._coef_cm(coefficients)
        bd = basis_desc(fb)
        self.check(self.lib.fock_mo_points(self.ptr, ctypes.byref(bd), _d(C), int(orbital), len(pts), _d(pts),
                                           _d(out)))
        return out

    def rhf_gradient(self, fb, fn, opts: ScfOpts, result) -> np.ndarray:
        """dE/dR (Eh/Bohr, shape (n_atoms, 3)) at a converged scf_run result."""
        n = fb.n
        C = np.ascontiguousarray(np.asarray(result["coefficients"], dtype=np.float64).T).ravel()  # column-major
        P = np.ascontiguousarray(np.asarray(result["density"], dtype=np.float64).T).ravel()
        eps = np.ascontiguousarray(result["orbital_energies"], dtype=np.float64)
        out = ScfOut()
        out.coefficients = _d(C)
        out.density = _d(P)
        out.orbital_energies = _d(eps)
        out.n_electrons = int(result["n_electrons"])
        out.n_basis = n
        na = len(fn.charges)
        g = np.zeros(3 * na)
        bd, nd = basis_desc(fb), nuclei_desc(fn)
        st = self.lib.fock_rhf_gradient(self.ptr, ctypes.byref(bd), ctypes.byref(nd), ctypes.byref(opts),
                                        ctypes.byref(out), _d(g))
        self.check(st)
        return g.reshape(na, 3)


def default_opts(**kw) -> ScfOpts:
    o = ScfOpts()
    load().fock_scf_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
