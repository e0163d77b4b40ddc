"""Tile autotuner of the blocked J/K digestion (fock_jk_autotune, fock_tiles_get/set, opts.autotune).

A tile is the most wavefront tasks (64 ket pairs each) of one bra pair that one workgroup takes; it
changes the work split and the summation order only. Pins: J/K at every tile equal the default
tile's to rounding (1e-12 relative, FP64; 1e-9, FP32 bulk kernels); tuned SCF and batched SCF runs
equal untuned ones; the batch schedules (by cost, XCD-affine with padding slots) give the same
molecules' results.
"""
import numpy as np
import pytest

from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu

ROBUST = dict(diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)


def cluster(k, seed):
    return m.flatten(m.build(m.water_cluster_geometry(k, np.random.default_rng(seed)), "6-31G"))


def sym_density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.1
    return A + A.T


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_tiles_default_and_validation(ctx):
    for form in (abi.TILES_FP64, abi.TILES_MIXED, abi.TILES_BATCH):
        ctx.set_tiles(form, [8] * 6, abi.SCHEDULE_COST)
        assert ctx.tiles(form) == ([8] * 6, abi.SCHEDULE_COST)
    with pytest.raises(abi.FockError):
        ctx.set_tiles(abi.TILES_FP64, [8, 8, 3, 8, 8, 8])
    with pytest.raises(abi.FockError):
        ctx.set_tiles(abi.TILES_BATCH, [8] * 6, 7)
    with pytest.raises(abi.FockError):
        ctx.set_tiles(5, [8] * 6)


@pytest.mark.parametrize("policy,screen,precision,tol", [
    (abi.POLICY_ACCURATE, abi.SCREEN_SCHWARZ, abi.PRECISION_FP64, 1e-12),
    (abi.POLICY_ACCURATE, abi.SCREEN_SCHWARZ, abi.PRECISION_MIXED, 1e-9),
    (abi.POLICY_REF_COMPAT, abi.SCREEN_REF_BF_PAIR, abi.PRECISION_FP64, 1e-12),
])
def test_every_tile_gives_the_same_jk(ctx, policy, screen, precision, tol):
    fb, _ = cluster(6, 7)
    form = abi.TILES_FP64 if precision == abi.PRECISION_FP64 else abi.TILES_MIXED
    ctx.set_tiles(form, [8] * 6)
    ctx.jk_prepare(fb, policy, screen, 1e-12, precision)
    P = sym_density(fb.n, 3)
    J0, K0 = ctx.jk_build(P)
    for t in abi.TILE_CHOICES:
        ctx.set_tiles(form, [t] * 6)  # re-tiles the prepared build at once
        J, K = ctx.jk_build(P)
        assert rel(J, J0) < tol and rel(K, K0) < tol, (t, rel(J, J0), rel(K, K0))
    ctx.set_tiles(form, [8] * 6)


def test_jk_autotune_report_and_result(ctx):
    fb, _ = cluster(8, 11)
    ctx.set_tiles(abi.TILES_FP64, [8] * 6)
    ctx.jk_prepare(fb, abi.POLICY_ACCURATE, abi.SCREEN_SCHWARZ, 1e-12, abi.PRECISION_FP64)
    P = sym_density(fb.n, 5)
    J0, K0 = ctx.jk_build(P)
    rep = ctx.jk_autotune(P)
    print("tiles", rep["tile"], "tune %.1f ms" % rep["tune_ms"])
    assert all(t in abi.TILE_CHOICES for t in rep["tile"])
    for c in range(6):
        timed = rep["ms"][c] >= 0
        assert timed.any()
        k = abi.TILE_CHOICES.index(rep["tile"][c])
        # the fastest candidate alone, or a tile the whole-build refinement moved to (every
        # candidate that fits was timed alone)
        assert rep["ms"][c][k] >= 0
    assert ctx.tiles(abi.TILES_FP64)[0] == rep["tile"]
    J, K = ctx.jk_build(P)  # the pair-indexed J accumulator was cleared after the timing launches
    assert rel(J, J0) < 1e-12 and rel(K, K0) < 1e-12
    ctx.set_tiles(abi.TILES_FP64, [8] * 6)


def test_facade_try_scf_autotune(ctx):
    from fock_amd.scf import try_scf
    atoms = m.build(m.water_cluster_geometry(3, np.random.default_rng(2)), "6-31G")
    kw = dict(policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ, ctx=ctx, **ROBUST)
    ctx.set_tiles(abi.TILES_FP64, [8] * 6)
    r0 = try_scf(atoms, 200, 1e-10, 0, **kw)
    r1 = try_scf(atoms, 200, 1e-10, 0, autotune=True, **kw)
    assert r0.converged and r1.converged and abs(r0.total_energy - r1.total_energy) < 1e-10
    ctx.set_tiles(abi.TILES_FP64, [8] * 6)


def test_scf_autotune_matches_untuned(ctx):
    fb, fn = cluster(4, 4)
    o = abi.default_opts(max_iters=200, epsilon=1e-10, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12, **ROBUST)
    ctx.set_tiles(abi.TILES_FP64, [8] * 6)
    st, r0 = ctx.scf_run(fb, fn, o)
    o.autotune = 1
    st1, r1 = ctx.scf_run(fb, fn, o)
    assert st == st1 == abi.FOCK_OK
    assert abs(r1["total_energy"] - r0["total_energy"]) < 1e-10
    print("tuned tiles", ctx.tiles(abi.TILES_FP64))
    ctx.set_tiles(abi.TILES_FP64, [8] * 6)


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_COST, abi.SCHEDULE_XCD])
@pytest.mark.parametrize("tile", [2, 16])
def test_batch_schedules_and_tiles_match_single(ctx, schedule, tile):
    mols = [m.flatten(m.build(g, "6-31G")) for g in m.water_cluster_batch(12, seed=31)]
    o = abi.default_opts(max_iters=200, epsilon=1e-9, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12, **ROBUST)
    ctx.set_tiles(abi.TILES_BATCH, [tile] * 6, schedule)
    st, statuses, res = ctx.scf_batch_run(mols, o)
    assert st == abi.FOCK_OK
    for i in (0, 5, 11):
        s1, r = ctx.scf_run(mols[i][0], mols[i][1], o)
        assert abs(res[i]["total_energy"] - r["total_energy"]) < 1e-9
    ctx.set_tiles(abi.TILES_BATCH, [8] * 6, abi.SCHEDULE_COST)


def test_batch_autotune(ctx):
    mols = [m.flatten(m.build(g, "6-31G")) for g in m.water_cluster_batch(16, seed=41)]
    o = abi.default_opts(max_iters=200, epsilon=1e-9, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12, autotune=1, **ROBUST)
    st, statuses, res = ctx.scf_batch_run(mols, o)
    assert st == abi.FOCK_OK
    tiles, sched = ctx.tiles(abi.TILES_BATCH)
    print("batch tiles", tiles, "schedule", sched)
    assert all(t in abi.TILE_CHOICES for t in tiles) and sched in (abi.SCHEDULE_COST, abi.SCHEDULE_XCD)
    o.autotune = 0
    for i in (0, 15):
        s1, r = ctx.scf_run(mols[i][0], mols[i][1], o)
        assert abs(res[i]["total_energy"] - r["total_energy"]) < 1e-9
    ctx.set_tiles(abi.TILES_BATCH, [8] * 6, abi.SCHEDULE_COST)
