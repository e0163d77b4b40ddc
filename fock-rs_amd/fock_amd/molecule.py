"""Molecules, basis sets and geometry generators (host-side data layer).

Mirrors the inputs the reference's `scf` crate consumes:

* ``BasisSet.get`` follows basis_set/src/lib.rs:79-119: every shell of a BSE
  JSON entry is expanded into Cartesian contracted functions, component order
  ``for x in 0..=L, for y in 0..=L, for z in 0..=L, x+y+z == L`` (so p comes
  as z, y, x), SP shells give the s functions before the p functions, each
  primitive coefficient is multiplied by the primitive norm
  (basis/src/primitives.rs:19-27,41-48) and contractions are NOT renormalised.
* ``read_xyz`` follows chemfiles/src/xyz.rs:13-31 (Angstrom x 1.89 -> Bohr).
* ``Atom.point_charge`` / electron counts follow basis_set/src/atom.rs:32-43
  (charge = electrons = atomic number).

The arrays produced by :func:`flatten` are exactly the flat structure-of-arrays
basis descriptor of the C ABI (``fock_basis_desc`` in include/fock.h).
"""
from __future__ import annotations

import json
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Sequence, Tuple

import numpy as np

ANGSTROM_TO_BOHR = 1.89  # chemfiles/src/xyz.rs:23-26 multiplies by 1.89, not 1.8897...

_DATA_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "basis")

SYMBOLS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
_XYZ_SYMBOLS = {"H": 1, "He": 2, "C": 6, "N": 7, "O": 8, "F": 9, "Fe": 26, "B": 5}  # periodic_table.rs:284-296


def powi(x: float, n: int) -> float:
    """Integer power with the square-and-multiply order of LLVM's powi lowering (Rust `f64::powi`)."""
    recip = n < 0
    b = n
    r = 1.0
    a = x
    while True:
        if b & 1:
            r *= a
        b = int(b / 2)  # C division truncates toward zero
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def _prod_range(lo: int, hi: int) -> int:
    p = 1
    for v in range(lo, hi + 1):
        p *= v
    return p


def primitive_norm(angular: Sequence[int], alpha: float) -> float:
    """Primitive normalisation, basis/src/primitives.rs:19-27."""
    i, j, k = angular
    frac_2_pi = 2.0 / math.pi
    a = math.sqrt(math.sqrt(powi(frac_2_pi * alpha, 3)))
    num = powi(8.0 * alpha, i + j + k)
    den = float(_prod_range(i + 1, 2 * i) * _prod_range(j + 1, 2 * j) * _prod_range(k + 1, 2 * k))
    return a * math.sqrt(num / den)


def cartesian_components(L: int) -> List[Tuple[int, int, int]]:
    """Cartesian components of a shell in the reference's order (basis_set/src/lib.rs:87-99)."""
    out = []
    for x in range(L + 1):
        for y in range(L + 1):
            for z in range(L + 1):
                if x + y + z == L:
                    out.append((x, y, z))
    return out


@dataclass
class Primitive:
    angular: Tuple[int, int, int]
    exponent: float
    coefficient: float  # normalised, as GaussianPrimitive::new stores it


@dataclass
class ContractedGaussian:
    position: np.ndarray
    primitives: List[Primitive]


@dataclass
class Atom:
    position: np.ndarray  # Bohr
    z: int
    basis: List[ContractedGaussian] = field(default_factory=list)

    def electron_count(self) -> int:
        return self.z

    def point_charge(self) -> Tuple[np.ndarray, float]:
        return self.position, float(self.z)


class BasisSet:
    """A BSE-JSON basis set (subset of the reference's schema, basis_set/src/lib.rs:15-76)."""

    _cache: Dict[str, "BasisSet"] = {}

    def __init__(self, name: str, elements: dict):
        self.name = name
        self.elements = elements

    @classmethod
    def load(cls, name: str) -> "BasisSet":
        if name not in cls._cache:
            path = os.path.join(_DATA_DIR, name + ".json")
            with open(path) as f:
                data = json.load(f)
            cls._cache[name] = cls(name, data["elements"])
        return cls._cache[name]

    def get(self, position, z: int) -> Atom:
        """basis_set/src/lib.rs:79-119."""
        key = str(int(z))
        if key not in self.elements:
            raise KeyError("Failed to get atom from basis set: element %d not in %s" % (z, self.name))
        position = np.asarray(position, dtype=np.float64).copy()
        funcs: List[ContractedGaussian] = []
        for shell in self.elements[key]["electron_shells"]:
            for idx, L in enumerate(shell["angular_momentum"]):
                for ang in cartesian_components(L):
                    prims = []
                    for s_exp, s_coef in zip(shell["exponents"], shell["coefficients"][idx]):
                        e = float(s_exp)
                        c = float(s_coef)
                        prims.append(Primitive(ang, e, c * primitive_norm(ang, e)))
                    funcs.append(ContractedGaussian(position, prims))
        return Atom(position, int(z), funcs)


def read_xyz(path: str, basis: BasisSet) -> List[Atom]:
    """chemfiles/src/xyz.rs:13-31."""
    pattern = re.compile(r"\s*(\w+)\s+([+\-0-9.]+)\s+([+\-0-9.]+)\s+([0-9.+\-]+).*\b")
    text = open(path).read()
    atoms = []
    for m in pattern.finditer(text):
        sym = m.group(1)
        if sym not in _XYZ_SYMBOLS:
            raise ValueError("not implemented yet: element %r" % sym)
        x = float(m.group(2)) * ANGSTROM_TO_BOHR
        y = float(m.group(3)) * ANGSTROM_TO_BOHR
        z = float(m.group(4)) * ANGSTROM_TO_BOHR
        atoms.append(basis.get((x, y, z), _XYZ_SYMBOLS[sym]))
    return atoms


@dataclass
class FlatBasis:
    """Flat structure-of-arrays view of a basis (the C ABI's basis descriptor)."""
    centers: np.ndarray      # [n, 3] float64, Bohr
    angular: np.ndarray      # [n, 3] int32
    prim_offset: np.ndarray  # [n + 1] int32
    exponents: np.ndarray    # [nprim] float64
    coefficients: np.ndarray  # [nprim] float64 (normalised)

    @property
    def n(self) -> int:
        return int(self.centers.shape[0])


@dataclass
class FlatNuclei:
    positions: np.ndarray  # [N, 3] float64, Bohr
    charges: np.ndarray    # [N] float64


def flatten(atoms: Iterable[Atom]) -> Tuple[FlatBasis, FlatNuclei]:
    """Flatten atoms the way try_scf does (scf/src/lib.rs:79-94)."""
    atoms = list(atoms)
    centers, angular, offs, exps, coefs = [], [], [0], [], []
    for atom in atoms:
        for cg in atom.basis:
            centers.append(cg.position)
            angular.append(cg.primitives[0].angular)
            for p in cg.primitives:
                exps.append(p.exponent)
                coefs.append(p.coefficient)
            offs.append(len(exps))
    fb = FlatBasis(
        np.ascontiguousarray(np.array(centers, dtype=np.float64).reshape(-1, 3)),
        np.ascontiguousarray(np.array(angular, dtype=np.int32).reshape(-1, 3)),
        np.ascontiguousarray(np.array(offs, dtype=np.int32)),
        np.ascontiguousarray(np.array(exps, dtype=np.float64)),
        np.ascontiguousarray(np.array(coefs, dtype=np.float64)),
    )
    fn = FlatNuclei(
        np.ascontiguousarray(np.array([a.position for a in atoms], dtype=np.float64).reshape(-1, 3)),
        np.ascontiguousarray(np.array([float(a.z) for a in atoms], dtype=np.float64)),
    )
    return fb, fn


def at_positions(fb: "FlatBasis", fn: "FlatNuclei", positions) -> Tuple["FlatBasis", "FlatNuclei"]:
    """The same molecule with nucleus i at positions[i] (Bohr); every basis function centred on a
    nucleus moves with it (the basis functions of an Atom share its position,
    basis_set/src/lib.rs:79-119)."""
    positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    if positions.shape != fn.positions.shape:
        raise ValueError("one position per nucleus expected")
    centers = fb.centers.copy()
    for i in range(len(fn.charges)):
        on = np.all(fb.centers == fn.positions[i], axis=1)
        centers[on] = positions[i]
    return (FlatBasis(centers, fb.angular, fb.prim_offset, fb.exponents, fb.coefficients),
            FlatNuclei(np.ascontiguousarray(positions), fn.charges))


def displaced(fb: "FlatBasis", fn: "FlatNuclei", atom: int, axis: int, h: float):
    """Nucleus `atom` (and its basis functions) moved by h Bohr along `axis`."""
    pos = fn.positions.copy()
    pos[atom, axis] += h
    return at_positions(fb, fn, pos)


def n_electrons(atoms: Sequence[Atom], charge: int = 0) -> int:
    """scf/src/lib.rs:91-94."""
    return sum(a.electron_count() for a in atoms) - charge


# --------------------------------------------------------------------------------------------
# Geometry generators (Angstrom in, Bohr out via the reference's 1.89 factor)
# --------------------------------------------------------------------------------------------

def build(symbols_xyz_angstrom: Sequence[Tuple[str, float, float, float]], basis_name: str) -> List[Atom]:
    basis = BasisSet.load(basis_name)
    return [basis.get((x * ANGSTROM_TO_BOHR, y * ANGSTROM_TO_BOHR, z * ANGSTROM_TO_BOHR), SYMBOLS[s])
            for s, x, y, z in symbols_xyz_angstrom]


def water_geometry() -> List[Tuple[str, float, float, float]]:
    """H2O: O at the origin, H at (0, +-0.757, 0.587) Angstrom (yz plane)."""
    return [("O", 0.0, 0.0, 0.0), ("H", 0.0, 0.757, 0.587), ("H", 0.0, -0.757, 0.587)]


def rotate_geometry(geom, R: np.ndarray):
    out = []
    for s, x, y, z in geom:
        v = R @ np.array([x, y, z])
        out.append((s, float(v[0]), float(v[1]), float(v[2])))
    return out


TETRA_HALF = math.acos(-1.0 / 3.0) / 2.0  # half the tetrahedral angle (54.7356 deg)


def alkane_geometry(n_carbon: int, cc: float = 1.54, ch: float = 1.09) -> List[Tuple[str, float, float, float]]:
    """All-anti zigzag n-alkane C_nH_{2n+2} with ideal tetrahedral angles.

    Backbone in the xy plane along x: dx = cc*sin(54.7356 deg) = 1.2574 A,
    alternating dy = +-cc*cos(54.7356 deg) = 0.8891 A. CH2 hydrogens sit in
    the plane perpendicular to the backbone plane, CH3 groups are staggered
    with one hydrogen in the backbone plane (continuing the zigzag).
    """
    if n_carbon < 1:
        raise ValueError("need at least one carbon")
    dx = cc * math.sin(TETRA_HALF)
    dy = cc * math.cos(TETRA_HALF)
    carbons = [np.array([i * dx, (dy if i % 2 else 0.0), 0.0]) for i in range(n_carbon)]

    def virtual(i: int) -> np.ndarray:
        # position a carbon at index i would have (extends the zigzag)
        return np.array([i * dx, (dy if i % 2 else 0.0), 0.0])

    atoms: List[Tuple[str, float, float, float]] = [("C", *map(float, c)) for c in carbons]
    ez = np.array([0.0, 0.0, 1.0])
    for i, c in enumerate(carbons):
        nb_prev = carbons[i - 1] if i > 0 else virtual(i - 1)
        nb_next = carbons[i + 1] if i + 1 < n_carbon else virtual(i + 1)
        u1 = (nb_prev - c) / np.linalg.norm(nb_prev - c)
        u2 = (nb_next - c) / np.linalg.norm(nb_next - c)
        bis = u1 + u2
        bis /= np.linalg.norm(bis)
        hs = []
        for sgn in (+1.0, -1.0):
            d = -bis * math.cos(TETRA_HALF) + sgn * ez * math.sin(TETRA_HALF)
            hs.append(c + ch * d)
        if n_carbon == 1:
            # methane: the two virtual neighbours become hydrogens too
            hs.append(c + ch * u1)
            hs.append(c + ch * u2)
        else:
            if i == 0:
                hs.append(c + ch * u1)
            if i == n_carbon - 1:
                hs.append(c + ch * u2)
        for h in hs:
            atoms.append(("H", float(h[0]), float(h[1]), float(h[2])))
    return atoms


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation from a uniform unit quaternion (Shoemake)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
    ])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster_geometry(k: int, rng: np.random.Generator, min_oo: float = 2.8,
                           radius_scale: float = 2.0) -> List[Tuple[str, float, float, float]]:
    """(H2O)_k: rigid monomers (water_geometry), uniform random orientations,
    O positions uniform in a sphere of radius radius_scale*k^(1/3) A with all O-O >= min_oo."""
    radius = radius_scale * k ** (1.0 / 3.0)
    mono = water_geometry()
    for _attempt in range(10000):
        pos: List[np.ndarray] = []
        tries = 0
        while len(pos) < k and tries < 20000:
            tries += 1
            p = (rng.random(3) * 2.0 - 1.0) * radius
            if p @ p > radius * radius:
                continue
            if all(np.linalg.norm(p - q) >= min_oo for q in pos):
                pos.append(p)
        if len(pos) == k:
            break
    else:  # pragma: no cover
        raise RuntimeError("could not place water molecules")
    out = []
    for p in pos:
        R = _random_rotation(rng)
        for s, x, y, z in mono:
            v = R @ np.array([x, y, z]) + p
            out.append((s, float(v[0]), float(v[1]), float(v[2])))
    return out


def water_cluster_batch(count: int = 64, seed: int = 20261202, kmax: int = 8):
    """The batching workload: `count` clusters (H2O)_k, k ~ U{1..kmax}, seeded."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(1, kmax + 1, size=count)
    return [water_cluster_geometry(int(k), rng) for k in ks]


# Named workloads ------------------------------------------------------------------------------

def h2o_sto3g() -> List[Atom]:
    return build(water_geometry(), "STO-3G")


def octane_631g() -> List[Atom]:
    return build(alkane_geometry(8), "6-31G")


def eicosane_631g() -> List[Atom]:
    return build(alkane_geometry(20), "6-31G")
