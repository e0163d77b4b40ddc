"""Orbitals at points on the GPU (fock_mo_points, fock_mo_grid): the renderers' hot loop
(volumetric/src/main.rs:100-125, pointcloud/src/main.rs:106-121) over MolecularOrbital::evaluate
(scf/src/molecular_orbitals.rs:37-43) with the |c| > 0.05 truncation (:69-90), against the
oracle's restatement (oracle/refmo.py) at the coefficient matrix of a converged SCF.

Tolerances: FP64 points within 1e-12 (absolute, orbital values are O(1)); the FP32 grid within
one float32 ulp of the oracle's FP64 value rounded to float32, plus 1e-14 absolute (cancellation
near the nodes leaves FP64 rounding of that size).
"""
import numpy as np
import pytest

from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def converged(ctx, atoms, eps=1e-10, **kw):
    from fock_amd.scf import try_scf
    r = try_scf(atoms, 300, eps, 0, policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ, ctx=ctx, **kw)
    assert r is not None and r.converged
    return r


def f32_close(got, want64):
    w = want64.astype(np.float32)
    return np.abs(got.astype(np.float64) - w.astype(np.float64)) <= np.spacing(np.abs(w)).astype(np.float64) + 1e-14


def test_mo_points_all_orbitals_w13(ctx):
    say("mo points W13")
    from oracle import refmo
    atoms = m.build(m.water_geometry(), "6-31G")
    r = converged(ctx, atoms)
    fb, _ = m.flatten(atoms)
    pts = np.random.default_rng(1).uniform(-6.0, 6.0, size=(5000, 3))
    pts[:3] = fb.centers[:3]   # on a nucleus: d = 0 exactly
    for k in range(fb.n):
        got = ctx.mo_points(fb, r.coefficient_matrix, k, pts)
        want = refmo.evaluate(fb, r.coefficient_matrix, k, pts)
        assert np.abs(got - want).max() < 1e-12, (k, np.abs(got - want).max())
        # the facade routes through the same entry point
        assert np.array_equal(r.orbitals.points(k, pts[:10], ctx=ctx), got[:10])


def test_mo_grid_renderer_grid_w13(ctx):
    """The volumetric renderer's grid (200^3 voxels over a 20 Bohr box) for HOMO and LUMO."""
    say("mo grid W13 200^3")
    from oracle import refmo
    atoms = m.build(m.water_geometry(), "6-31G")
    r = converged(ctx, atoms)
    fb, _ = m.flatten(atoms)
    homo = r.n_electrons // 2 - 1
    g = r.orbitals.grid([homo, homo + 1], 200, 20.0, ctx=ctx)
    assert g.shape == (2, 200, 200, 200) and g.dtype == np.float32
    pts = refmo.volumetric_points(200, 20.0)
    for o, k in enumerate((homo, homo + 1)):
        want = refmo.evaluate(fb, r.coefficient_matrix, k, pts)
        ok = f32_close(g[o].ravel(), want)
        assert ok.all(), (k, np.count_nonzero(~ok), np.abs(g[o].ravel() - want).max())
        assert np.abs(g[o]).max() > 0.1


def test_mo_grid_eicosane_sampled(ctx):
    """C20H42/6-31G HOMO on the renderer's grid, every 997th voxel checked against the oracle;
    a second, shifted and anisotropic grid checks origin, steps and voxel order."""
    say("mo grid C20H42")
    from oracle import refmo
    atoms = m.eicosane_631g()
    r = converged(ctx, atoms, eps=1e-8, diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)
    fb, _ = m.flatten(atoms)
    homo = r.n_electrons // 2 - 1
    g = r.orbitals.grid([homo], 200, 40.0, ctx=ctx)[0].ravel()
    pts = refmo.volumetric_points(200, 40.0)
    sel = np.arange(0, len(pts), 997)
    want = refmo.evaluate(fb, r.coefficient_matrix, homo, pts[sel])
    assert f32_close(g[sel], want).all()
    origin, step, dims = (-12.5, -3.0, -4.0), (0.25, 0.2, 0.3), (100, 31, 27)
    g2 = ctx.mo_grid(fb, r.coefficient_matrix, [homo - 3], origin, step, dims)[0]
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    p2 = np.stack([origin[0] + x.ravel() * step[0], origin[1] + y.ravel() * step[1],
                   origin[2] + z.ravel() * step[2]], axis=1)
    want2 = refmo.evaluate(fb, r.coefficient_matrix, homo - 3, p2)
    assert f32_close(g2.ravel(), want2).all()


def test_mo_grid_rejects_bad_arguments(ctx):
    fb, _ = m.flatten(m.h2o_sto3g())
    C = np.eye(fb.n)
    with pytest.raises(abi.FockError):
        ctx.mo_grid(fb, C, [fb.n], (0, 0, 0), (1, 1, 1), (2, 2, 2))      # orbital out of range
    with pytest.raises(abi.FockError):
        ctx.mo_grid(fb, C, [0], (0, 0, 0), (1, 1, 1), (0, 2, 2))         # empty grid
    bad = C.copy(); bad[0, 0] = np.nan
    with pytest.raises(abi.FockError):
        ctx.mo_points(fb, bad, 0, np.zeros((1, 3)))
