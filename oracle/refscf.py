"""numpy restatement of the reference SCF driver -- TEST INFRASTRUCTURE ONLY.

Follows scf/src/lib.rs:73-258 step by step:
  E_nuc (:102-116), S/T/V (:118-129), ERI tensor (:131-136), X = U s^-1/2 U^T with
  s = diag(U^T S U) (:139-146), GWH guess with 0.875 on the diagonal too (:148-162),
  electron_terms (:167-173), G build (:177-185), F = H + G (:187),
  DIIS error FPS - SPF from the un-extrapolated F and the old P (:189-200),
  window 16, extrapolate from 5 stored matrices (:197-207), F' = X^T F X, sorted
  eigendecomposition, C = X C' (:210-212), P = 2 C_occ C_occ^T (:214-216),
  E_elec = 1/2 tr(P_new (2H + G(P_old))) (:223), exit on err < eps or iter ==
  max_iters (:224) -> result; !err.is_normal() -> None (:248-249).
utils::diis (scf/src/utils.rs:102-138): B_ij = <e_i, e_j>_F bordered by -1,
inverse times (0,...,0,-1), F = sum c_i F_i; None when B is singular.

Parity unpinned (third-party arithmetic of the reference, not reproducible here):
faer-evd 0.9 `compute_hermitian_evd` (scf/src/utils.rs:68-77) and nalgebra 0.30
`SymmetricEigen` / `try_inverse` are replaced by LAPACK (numpy.linalg.eigh / inv);
eigenvector signs and last-ulp rounding differ. Results are judged through
sign-invariant quantities (energies, sorted orbital energies, density matrix).
"""
from __future__ import annotations

import math
from collections import deque
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import BasisView, dptr, lib

TERMS_MAX_N = 120  # n^4 electron_terms array (as the reference) up to 120^4*8 B = 1.6 GB


@dataclass
class ScfResult:
    orbital_energies: np.ndarray
    density_matrix: np.ndarray
    coefficient_matrix: np.ndarray
    electronic_energy: float
    total_energy: float
    iterations: int
    n_electrons: int
    n_basis: int
    errors: list
    energies: list


def is_normal(x: float) -> bool:
    """Rust f64::is_normal: not zero, subnormal, infinite or NaN."""
    return math.isfinite(x) and x != 0.0 and abs(x) >= 2.2250738585072014e-308


def nuclear_repulsion(positions: np.ndarray, charges: np.ndarray) -> float:
    """scf/src/lib.rs:102-116 (sequential sum over i < j)."""
    total = 0.0
    for i in range(len(charges)):
        for j in range(i + 1, len(charges)):
            d = positions[j] - positions[i]
            total += charges[i] * charges[j] / math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    return total


def one_electron(fb, fn, accurate: bool):
    n = fb.n
    S = np.zeros((n, n)); T = np.zeros((n, n)); V = np.zeros((n, n))
    bv = BasisView(fb)
    pos = np.ascontiguousarray(fn.positions); ch = np.ascontiguousarray(fn.charges)
    L = lib()
    if accurate:
        L.o64_one_electron(bv.ref, len(ch), dptr(pos), dptr(ch), dptr(S), dptr(T), dptr(V))
    else:
        L.ref_set_policy(0, 0)
        L.ref_one_electron(bv.ref, len(ch), dptr(pos), dptr(ch), dptr(S), dptr(T), dptr(V))
    return S, T, V


def eri_packed(fb, accurate: bool, screen: bool = True):
    """Packed canonical ERIs: reference restatement (with its 1e-6 screen) or the accurate oracle."""
    L = lib()
    packed = np.zeros(L.ref_packed_size(fb.n))
    bv = BasisView(fb)
    if accurate:
        rc = L.o64_eri_packed(bv.ref, dptr(packed))
        if rc != 0:
            raise ValueError("basis cannot be grouped into shells")
        computed = len(packed)
    else:
        L.ref_set_policy(0, 0)
        computed = L.ref_electron_tensor(bv.ref, 1 if screen else 0, dptr(packed))
    return packed, computed


def dense_from_packed(n: int, packed: np.ndarray) -> np.ndarray:
    out = np.zeros(n ** 4)
    lib().ref_unpack_dense(n, dptr(packed), dptr(out))
    return out.reshape(n, n, n, n)


class GBuilder:
    """G(P) = sum_xy P_xy [(ij|xy) - 1/2 (ix|jy)] as the reference computes it."""

    def __init__(self, n: int, packed: np.ndarray):
        self.n = n
        self.packed = packed
        self.terms = None
        if n <= TERMS_MAX_N:
            self.terms = np.zeros(n ** 4)
            lib().ref_electron_terms(n, dptr(packed), dptr(self.terms))

    def __call__(self, P: np.ndarray) -> np.ndarray:
        P = np.ascontiguousarray(P, dtype=np.float64)
        G = np.zeros((self.n, self.n))
        if self.terms is not None:
            lib().ref_g_build(self.n, dptr(self.terms), dptr(P), dptr(G))
        else:
            lib().ref_g_build_packed(self.n, dptr(self.packed), dptr(P), dptr(G))
        return G


def sorted_eigs(M: np.ndarray):
    """scf/src/utils.rs:59-99 (ascending eigenvalues). Parity unpinned: LAPACK, not faer."""
    w, v = np.linalg.eigh(0.5 * (M + M.T))
    order = np.argsort(w, kind="stable")
    return v[:, order], w[order]


def diis(errors, focks) -> Optional[np.ndarray]:
    """scf/src/utils.rs:102-138. Parity unpinned: LAPACK inverse, not nalgebra try_inverse."""
    n = len(errors)
    B = np.zeros((n + 1, n + 1))
    for i in range(n):
        for j in range(n):
            B[i, j] = float(np.sum(errors[i] * errors[j]))
    B[n, :n] = -1.0
    B[:n, n] = -1.0
    B[n, n] = 0.0
    b = np.zeros(n + 1)
    b[n] = -1.0
    try:
        inv = np.linalg.inv(B)
    except np.linalg.LinAlgError:
        return None
    c = inv @ b
    F = c[0] * focks[0]
    for i in range(1, n):
        F = F + c[i] * focks[i]
    return F


def ediis_coefficients(e: np.ndarray, Q: np.ndarray) -> Optional[np.ndarray]:
    """min e.c + 1/2 c^T Q c over the simplex (c >= 0, sum c = 1) by enumeration: for every subset
    of free indices, the equality-constrained optimum [Q_FF 1; 1^T 0][c_F; lam] = [-e_F; 1]; the
    lowest objective among the feasible ones (c_F >= 0) is the minimum over the simplex when the
    objective is convex there. Exact and independent of the library's active-set method; for the
    small windows of the tests (m <= 10: 1023 subsets)."""
    m = len(e)
    best, best_val = None, np.inf
    for mask in range(1, 1 << m):
        F = [i for i in range(m) if mask >> i & 1]
        f = len(F)
        A = np.zeros((f + 1, f + 1))
        A[:f, :f] = Q[np.ix_(F, F)]
        A[:f, f] = 1.0
        A[f, :f] = 1.0
        rhs = np.concatenate([-e[F], [1.0]])
        try:
            x = np.linalg.solve(A, rhs)
        except np.linalg.LinAlgError:
            continue
        if np.any(x[:f] < -1e-12) or not np.all(np.isfinite(x)):
            continue
        c = np.zeros(m)
        c[F] = np.maximum(x[:f], 0.0)
        c /= c.sum()
        val = float(e @ c + 0.5 * c @ Q @ c)
        if val < best_val - 1e-15:
            best, best_val = c, val
    return best


def density(C: np.ndarray, n_electrons: int) -> np.ndarray:
    occ = C[:, : n_electrons // 2]
    return 2.0 * occ @ occ.T


def try_scf(fb, fn, max_iters: int, epsilon: float, molecule_charge: int = 0, accurate: bool = False,
            screen: bool = True, integrals=None, P0=None, diis_start: int = 5, diis_window: int = 16,
            level_shift: float = 0.0, level_shift_until: float = 0.0,
            exit_plain: Optional[bool] = None, damping: float = 1.0, damping_until: float = 0.0,
            ediis: bool = False, ediis_until: float = 0.1) -> Optional[ScfResult]:
    """scf/src/lib.rs:73-258 on the flat basis `fb` and nuclei `fn`.

    accurate=False: the reference's own integrals (refcpu.c, LUT Boys, x-branch as written,
    1e-6 pair screen when `screen`). accurate=True: oracle64.c, unscreened.
    `integrals` may pass precomputed (S, T, V, packed) to skip their evaluation; `P0` replaces
    the GWH guess by a given density. diis_start / diis_window / level_shift(_until) generalise
    the reference's constants (5, 16, no shift) exactly as the C ABI options do.
    exit_plain (default: `accurate`): on the iteration whose error is below epsilon, diagonalise
    the newest Fock matrix instead of the DIIS extrapolation -- the rule the GPU driver applies
    under its accurate policy (an extrapolation from parallel errors is a singular system whose
    solution follows rounding); the reference's own rule is exit_plain=False.
    damping: the mixing factor F of scf/src/lib.rs:218-219 (the reference fixes F = 1), applied
    while the error exceeds damping_until. ediis: while the error exceeds ediis_until, the Fock
    matrix is the EDIIS combination of the stored ones (ediis_coefficients) instead of DIIS.
    """
    if exit_plain is None:
        exit_plain = accurate
    n = fb.n
    n_electrons = int(round(float(np.sum(fn.charges)))) - molecule_charge
    e_nuc = nuclear_repulsion(fn.positions, fn.charges)
    if integrals is None:
        S, T, V = one_electron(fb, fn, accurate)
        packed, _ = eri_packed(fb, accurate, screen)
    else:
        S, T, V, packed = integrals
    H = T + V
    # X = U diag(1/sqrt((U^T S U)_ii)) U^T (scf/src/lib.rs:139-146)
    _, U = np.linalg.eigh(S)
    D = U.T @ (S @ U)
    X = U @ (np.diag(1.0 / np.sqrt(np.diag(D))) @ U.T)
    # GWH guess (scf/src/lib.rs:150-162)
    if P0 is not None:  # restart from a given density instead of the GWH guess
        P = np.array(P0, dtype=np.float64)
    else:
        hd = np.diag(H)
        Heht = 0.875 * S * (hd[:, None] + hd[None, :])
        Cp, _ = sorted_eigs(X.T @ (Heht @ X))
        C = X @ Cp
        P = density(C, n_electrons)
    gbuild = GBuilder(n, packed)
    focks, errs, dens = deque(), deque(), deque()
    err_hist, e_hist = [], []
    for it in range(max_iters + 1):
        G = gbuild(P)
        F = H + G
        E = (F @ P) @ S - (S @ P) @ F
        err = float(np.max(np.abs(E)))
        errs.append(E)
        focks.append(F)
        dens.append(P)
        if len(errs) > diis_window:
            errs.popleft()
            focks.popleft()
            dens.popleft()
        Fx = None
        if ediis and len(focks) >= 2 and err > ediis_until and not (exit_plain and err < epsilon):
            m_ = len(focks)
            a = np.array([[float(np.sum(dens[i] * focks[j])) for j in range(m_)] for i in range(m_)])
            e = np.array([0.5 * (float(np.sum(dens[i] * H)) + a[i, i]) for i in range(m_)])
            d = np.diag(a)
            Q = -0.5 * (d[:, None] + d[None, :] - a - a.T)
            c = ediis_coefficients(e, Q)
            if c is not None:
                Fx = sum(ci * Fi for ci, Fi in zip(c, focks))
        if Fx is not None:
            pass
        elif len(focks) < diis_start or (exit_plain and err < epsilon):
            Fx = focks[-1]
        else:
            Fx = diis(errs, focks)
            if Fx is None:
                Fx = focks[-1]
        if level_shift > 0.0 and err > level_shift_until:
            Fx = Fx + level_shift * (S - 0.5 * (S @ P) @ S)
        Cp, eps = sorted_eigs(X.T @ (Fx @ X))
        C = X @ Cp
        P_new = density(C, n_electrons)
        # new_density = F new + (1 - F) density (scf/src/lib.rs:218-219)
        P = damping * P_new + (1.0 - damping) * P if (0.0 < damping < 1.0 and err > damping_until) else P_new
        e_elec = 0.5 * float(np.trace(P @ (2.0 * H + G)))
        err_hist.append(err)
        e_hist.append(e_elec + e_nuc)
        if err < epsilon or it == max_iters:
            return ScfResult(eps, P, C, e_elec, e_elec + e_nuc, it, n_electrons, n, err_hist, e_hist)
        elif not is_normal(err):
            return None
    return None


def fd_gradient(fb, fn, h: float = 1e-4, epsilon: float = 1e-13, max_iters: int = 400, **kw) -> np.ndarray:
    """Central finite differences of the accurate (unscreened, oracle64) SCF energy, Eh/Bohr,
    shape (n_atoms, 3): the independent check of the analytic gradient (step h in Bohr). The
    energy expression 1/2 tr(P_new (2H + G(P_old))) is first order in the SCF error, so the SCF is
    converged to epsilon = 1e-13 (energy noise ~1e-12, i.e. ~1e-8 in the difference quotient)."""
    from fock_amd.molecule import displaced
    na = len(fn.charges)
    g = np.zeros((na, 3))
    for a in range(na):
        for x in range(3):
            e = []
            for s in (+1.0, -1.0):
                b2, n2 = displaced(fb, fn, a, x, s * h)
                r = try_scf(b2, n2, max_iters, epsilon, 0, accurate=True, **kw)
                if r is None:
                    raise RuntimeError("SCF failed at a displaced geometry")
                e.append(r.total_energy)
            g[a, x] = (e[0] - e[1]) / (2.0 * h)
    return g
