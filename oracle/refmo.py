"""numpy restatement of the reference's orbital evaluation -- TEST INFRASTRUCTURE ONLY.

* ``truncate``: MolecularOrbitals::new (scf/src/molecular_orbitals.rs:69-90) keeps, per column of
  the coefficient matrix, the basis functions whose coefficient has |c| > 0.05, in basis order.
* ``evaluate``: MolecularOrbital::evaluate (scf/src/molecular_orbitals.rs:37-43) sums c_f phi_f(r)
  over the kept functions; phi_f is ContractedGaussian::evaluate
  (basis/src/contracted_gaussian.rs:320-331): with d = A - r, sum_k c_k d_x^i d_y^j d_z^k
  exp(-alpha_k |d|^2).
* ``volumetric_points``: the voxel centres of volumetric/src/main.rs:100-125 (voxel i = x + N y +
  N^2 z at ((x, y, z) / N - 0.5) * BOX).

Vectorised over points in float64 (the reference's arithmetic); only the summation order of the
vectorised sums differs from the reference's sequential folds (rounding level).
"""
from __future__ import annotations

import numpy as np

MIN_COEFFICIENT_MAGNITUDE = 0.05  # scf/src/molecular_orbitals.rs:47


def truncate(coefficients: np.ndarray, orbital: int):
    """(indices, coefficients) of the functions MolecularOrbitals::new keeps for one column."""
    col = np.asarray(coefficients[:, orbital], dtype=np.float64)
    idx = np.nonzero(np.abs(col) > MIN_COEFFICIENT_MAGNITUDE)[0]
    return idx, col[idx]


def basis_values(fb, points: np.ndarray, functions) -> np.ndarray:
    """phi_f(r) for the listed functions f at points [m, 3]: array [len(functions), m]."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(functions), len(pts)))
    for row, f in enumerate(functions):
        d = fb.centers[f][None, :] - pts
        d2 = np.sum(d * d, axis=1)
        i, j, k = (int(v) for v in fb.angular[f])
        poly = d[:, 0] ** i * d[:, 1] ** j * d[:, 2] ** k
        acc = np.zeros(len(pts))
        for p in range(fb.prim_offset[f], fb.prim_offset[f + 1]):
            acc += fb.coefficients[p] * poly * np.exp(-fb.exponents[p] * d2)
        out[row] = acc
    return out


def evaluate(fb, coefficients: np.ndarray, orbital: int, points: np.ndarray) -> np.ndarray:
    """MolecularOrbital::evaluate of orbital `orbital` at every point [m, 3]."""
    idx, c = truncate(coefficients, orbital)
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    acc = np.zeros(len(pts))
    for f, cf in zip(idx, c):  # one function at a time: no [functions, points] array
        acc += cf * basis_values(fb, pts, [f])[0]
    return acc


def volumetric_points(n_voxels: int, box: float) -> np.ndarray:
    """Voxel centres in the renderer's order: voxel i = x + N y + N^2 z."""
    v = np.arange(n_voxels, dtype=np.float64) / n_voxels - 0.5
    z, y, x = np.meshgrid(v, v, v, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) * box
