"""Python mirror of the reference's `scf` crate API over the C ABI.

    SelfConsistentField.try_scf(atoms, max_iters, epsilon, molecule_charge)
        -> Optional[HartreeFockResult]                      (scf/src/lib.rs:48-78)
    HartreeFockResult fields                                (scf/src/lib.rs:27-37)
    MolecularOrbitals / MolecularOrbital.evaluate           (scf/src/molecular_orbitals.rs)

Same argument meaning and outcome rules as the reference: the result is returned
when max|FDS - SDF| < epsilon or at the last iteration (converged or not), and None
when the error stops being a normal float. The work runs in libfock.so on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import abi
from .molecule import Atom, ContractedGaussian, flatten

_CTX = None


def context() -> abi.Context:
    global _CTX
    if _CTX is None:
        _CTX = abi.Context(0)
    return _CTX


def evaluate_function(cg: ContractedGaussian, at) -> float:
    """basis/src/contracted_gaussian.rs:320-331: sum_k c_k (A-r)^{ijk} exp(-a |A-r|^2)."""
    diff = cg.position - np.asarray(at, dtype=np.float64)
    d2 = float(diff @ diff)
    acc = 0.0
    for p in cg.primitives:
        i, j, k = p.angular
        acc += p.coefficient * diff[0] ** i * diff[1] ** j * diff[2] ** k * math.exp(-p.exponent * d2)
    return acc


class MolecularOrbital:
    """scf/src/molecular_orbitals.rs:10-44."""

    def __init__(self, basis_functions: List[ContractedGaussian], coefficients: np.ndarray):
        self.basis_functions = basis_functions
        self.coefficients = coefficients

    def evaluate(self, at) -> float:
        return float(sum(c * evaluate_function(b, at) for c, b in zip(self.coefficients, self.basis_functions)))


class MolecularOrbitals:
    """scf/src/molecular_orbitals.rs:46-99: keeps coefficients with |c| > 0.05 per orbital.

    grid() and points() evaluate orbitals on the GPU (fock_mo_grid / fock_mo_points) with the same
    truncation: the renderers' loops (volumetric/src/main.rs:100-125, pointcloud/src/main.rs:106-121)."""
    MIN_COEFFICIENT_MAGNITUDE = 0.05

    def __init__(self, basis_functions: List[ContractedGaussian], coeff_matrix: np.ndarray):
        self._coefficients = np.ascontiguousarray(coeff_matrix, dtype=np.float64)
        self._flat = None
        self._basis_functions = list(basis_functions)
        self.molecular_orbitals = []
        for k in range(coeff_matrix.shape[1]):
            col = coeff_matrix[:, k]
            idx = [i for i in range(len(col)) if abs(col[i]) > self.MIN_COEFFICIENT_MAGNITUDE]
            self.molecular_orbitals.append(MolecularOrbital([basis_functions[i] for i in idx],
                                                            np.array([col[i] for i in idx])))

    def __getitem__(self, index: int) -> MolecularOrbital:
        return self.molecular_orbitals[index]

    def _flat_basis(self):
        if self._flat is None:
            from .molecule import Atom
            self._flat, _ = flatten([Atom(np.zeros(3), 0, self._basis_functions)])
        return self._flat

    def grid(self, orbitals, n_voxels: int = 200, box: float = 20.0, ctx: Optional[abi.Context] = None) -> np.ndarray:
        """float32 [len(orbitals), N, N, N] (z, y, x) of the renderer's voxel grid: voxel (x, y, z) at
        ((x, y, z) / N - 0.5) * box (volumetric/src/main.rs:111-116)."""
        step = box / n_voxels
        return (ctx or context()).mo_grid(self._flat_basis(), self._coefficients, orbitals, (-0.5 * box,) * 3,
                                          (step,) * 3, (n_voxels,) * 3)

    def points(self, orbital: int, points, ctx: Optional[abi.Context] = None) -> np.ndarray:
        """MolecularOrbital::evaluate of one orbital at many points [m, 3] (Bohr), float64."""
        return (ctx or context()).mo_points(self._flat_basis(), self._coefficients, orbital, points)

    def __len__(self) -> int:
        return len(self.molecular_orbitals)


@dataclass
class HartreeFockResult:
    """scf/src/lib.rs:27-37."""
    orbitals: MolecularOrbitals
    orbital_energies: np.ndarray
    density_matrix: np.ndarray
    coefficient_matrix: np.ndarray
    electronic_energy: float
    total_energy: float
    iterations: int
    n_electrons: int
    n_basis: int
    history: list
    scf_wall_ms: float
    converged: bool
    # options the SCF ran with (integral policy, screen, Schwarz threshold): a nuclear gradient of
    # this result differentiates the same energy
    policy: int = abi.POLICY_ACCURATE
    screen: int = abi.SCREEN_SCHWARZ
    schwarz_threshold: float = 1e-12


def try_scf(atoms: Iterable[Atom], max_iters: int, epsilon: float, molecule_charge: int = 0,
            policy: int = abi.POLICY_REF_COMPAT, screen: int = abi.SCREEN_REF_BF_PAIR,
            schwarz_threshold: float = 1e-12, precision: int = abi.PRECISION_FP64,
            incremental: bool = False, ctx: Optional[abi.Context] = None,
            initial_density: Optional[np.ndarray] = None, diis_start: int = 5, diis_window: int = 16,
            level_shift: float = 0.0, level_shift_until: float = 0.0,
            autotune: bool = False, ediis: bool = False, ediis_until: float = 0.1,
            damping: float = 1.0, damping_until: float = 0.0) -> Optional[HartreeFockResult]:
    """SelfConsistentField::try_scf (scf/src/lib.rs:73-258) on the GPU.

    Defaults reproduce the reference's own arithmetic and schedule (REF_COMPAT integrals, its
    1e-6 basis-function-pair screen, FP64, GWH guess, DIIS from 5 stored matrices in a window
    of 16). autotune: time the J/K tiles at the guess density first (kept in the context).
    ediis / damping: EDIIS above an error of ediis_until, density mixing factor (the reference's F of
    scf/src/lib.rs:218, fixed there at 1) above damping_until."""
    atoms = list(atoms)
    fb, fn = flatten(atoms)
    opts = abi.default_opts(max_iters=max_iters, epsilon=epsilon, molecule_charge=molecule_charge,
                            integral_policy=policy, screen=screen, schwarz_threshold=schwarz_threshold,
                            precision=precision, incremental=1 if incremental else 0,
                            diis_start=diis_start, diis_window=diis_window, level_shift=level_shift,
                            level_shift_until=level_shift_until, autotune=1 if autotune else 0,
                            ediis=1 if ediis else 0, ediis_until=ediis_until, damping=damping,
                            damping_until=damping_until)
    if initial_density is not None:
        P0 = np.ascontiguousarray(initial_density, dtype=np.float64)
        if P0.shape != (fb.n, fb.n):
            raise ValueError("initial density must be n x n")
        opts.initial_density = abi._d(P0)
    ctx = ctx or context()
    st, r = ctx.scf_run(fb, fn, opts)
    if st == abi.FOCK_DIVERGED:
        return None
    basis = [cg for a in atoms for cg in a.basis]
    return HartreeFockResult(
        orbitals=MolecularOrbitals(basis, r["coefficients"]),
        orbital_energies=r["orbital_energies"], density_matrix=r["density"],
        coefficient_matrix=r["coefficients"], electronic_energy=r["electronic_energy"],
        total_energy=r["total_energy"], iterations=r["iterations"], n_electrons=r["n_electrons"],
        n_basis=r["n_basis"], history=r["history"], scf_wall_ms=r["scf_wall_ms"], converged=(st == abi.FOCK_OK),
        policy=policy, screen=screen, schwarz_threshold=schwarz_threshold)


def _result(atoms, r, st, policy, screen, schwarz_threshold) -> HartreeFockResult:
    basis = [cg for a in atoms for cg in a.basis]
    return HartreeFockResult(
        orbitals=MolecularOrbitals(basis, r["coefficients"]),
        orbital_energies=r["orbital_energies"], density_matrix=r["density"],
        coefficient_matrix=r["coefficients"], electronic_energy=r["electronic_energy"],
        total_energy=r["total_energy"], iterations=r["iterations"], n_electrons=r["n_electrons"],
        n_basis=r["n_basis"], history=r.get("history", []), scf_wall_ms=r["scf_wall_ms"],
        converged=(st == abi.FOCK_OK), policy=policy, screen=screen, schwarz_threshold=schwarz_threshold)


def try_scf_batch(molecules: Sequence[Iterable[Atom]], max_iters: int, epsilon: float, molecule_charge: int = 0,
                  policy: int = abi.POLICY_ACCURATE, screen: int = abi.SCREEN_SCHWARZ,
                  schwarz_threshold: float = 1e-12, ctx: Optional[abi.Context] = None,
                  **schedule) -> List[Optional[HartreeFockResult]]:
    """try_scf for many molecules at once (batching form, FP64): the SCFs advance in lockstep with
    one Fock-build launch per quartet class per iteration. None where an SCF diverged."""
    mols = [list(a) for a in molecules]
    flat = [flatten(a) for a in mols]
    opts = abi.default_opts(max_iters=max_iters, epsilon=epsilon, molecule_charge=molecule_charge,
                            integral_policy=policy, screen=screen, schwarz_threshold=schwarz_threshold, **schedule)
    ctx = ctx or context()
    _, statuses, results = ctx.scf_batch_run(flat, opts)
    return [None if st == abi.FOCK_DIVERGED else _result(a, r, st, policy, screen, schwarz_threshold)
            for a, r, st in zip(mols, results, statuses)]


def nuclear_gradient(atoms: Iterable[Atom], result: HartreeFockResult, screen: Optional[int] = None,
                     schwarz_threshold: Optional[float] = None, ctx: Optional[abi.Context] = None) -> np.ndarray:
    """dE/dR in Eh/Bohr, shape (n_atoms, 3), at an accurate-policy try_scf result. The screen and
    Schwarz threshold default to the ones the SCF ran with, so the gradient is that of the
    returned energy. Raises ValueError for a result of the reference-compatible integrals or of
    the reference's 1e-6 pair screen: neither energy is differentiable (the Boys table is
    piecewise linear; the pair screen zeroes single integrals)."""
    if result.policy != abi.POLICY_ACCURATE:
        raise ValueError("nuclear gradients need an SCF with the accurate integral policy")
    screen = result.screen if screen is None else screen
    schwarz_threshold = result.schwarz_threshold if schwarz_threshold is None else schwarz_threshold
    if screen == abi.SCREEN_REF_BF_PAIR:
        raise ValueError("nuclear gradients need the Schwarz screen or none, not the 1e-6 pair screen")
    fb, fn = flatten(list(atoms))
    opts = abi.default_opts(integral_policy=abi.POLICY_ACCURATE, screen=screen, schwarz_threshold=schwarz_threshold)
    r = {"coefficients": result.coefficient_matrix, "density": result.density_matrix,
         "orbital_energies": result.orbital_energies, "n_electrons": result.n_electrons}
    return (ctx or context()).rhf_gradient(fb, fn, opts, r)


class SelfConsistentField:
    """Blanket-impl analogue: SelfConsistentField(atoms).try_scf(max_iters, epsilon, charge)."""

    def __init__(self, atoms: Iterable[Atom]):
        self.atoms = list(atoms)

    def try_scf(self, max_iters: int, epsilon: float, molecule_charge: int = 0, **kw):
        return try_scf(self.atoms, max_iters, epsilon, molecule_charge, **kw)
