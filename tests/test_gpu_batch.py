"""Batching form: many water clusters (H2O)_k, k uniform in 1..8, 6-31G, advanced in lockstep
with one Fock-build launch per quartet class per iteration (fock_scf_batch_run). Each molecule is
checked against the single-molecule path (fock_scf_run) through the same C ABI and, for seeded
(H2O)_2..4, against the oracle.
"""
import numpy as np
import pytest

from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu

ROBUST = dict(diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)


def opts(**kw):
    base = dict(max_iters=200, epsilon=1e-9, integral_policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ,
                schwarz_threshold=1e-12, **ROBUST)
    base.update(kw)
    return abi.default_opts(**base)


def clusters(count, seed):
    return [m.flatten(m.build(g, "6-31G")) for g in m.water_cluster_batch(count, seed=seed)]


def check_against_single(ctx, mols, results, o, which):
    for i in which:
        st, r = ctx.scf_run(mols[i][0], mols[i][1], o)
        b = results[i]
        assert st == abi.FOCK_OK
        assert abs(b["total_energy"] - r["total_energy"]) < 1e-9, (i, b["total_energy"], r["total_energy"])
        assert np.abs(b["orbital_energies"] - r["orbital_energies"]).max() < 1e-7
        assert np.abs(b["density"] - r["density"]).max() < 1e-6


def test_batch_matches_single_molecule_runs(ctx):
    mols = clusters(8, seed=11)
    o = opts()
    st, statuses, res = ctx.scf_batch_run(mols, o)
    assert st == abi.FOCK_OK and all(s == abi.FOCK_OK for s in statuses)
    check_against_single(ctx, mols, res, o, range(len(mols)))


def test_batch_64_clusters(ctx):
    mols = clusters(64, seed=20261202)
    sizes = sorted({fb.n for fb, _ in mols})
    assert sizes[0] >= 13 and sizes[-1] <= 104 and len(sizes) > 4   # uneven sizes
    o = opts()
    st, statuses, res = ctx.scf_batch_run(mols, o)
    assert st == abi.FOCK_OK, statuses
    print("64 clusters: %.1f ms wall" % res[0]["scf_wall_ms"])
    big = max(range(64), key=lambda i: mols[i][0].n)
    small = min(range(64), key=lambda i: mols[i][0].n)
    check_against_single(ctx, mols, res, o, [0, big, small])


def test_batch_above_jacobi_limit_uses_rocsolver(ctx):
    """n-decane/6-31G (n = 134 > 128, the LDS limit of the batched Jacobi solver) with two water
    clusters: the batch falls back to rocSOLVER per size group and still matches single runs."""
    mols = clusters(2, seed=7) + [m.flatten(m.build(m.alkane_geometry(10), "6-31G"))]
    assert mols[-1][0].n == 134
    o = opts()
    st, statuses, res = ctx.scf_batch_run(mols, o)
    assert st == abi.FOCK_OK, statuses
    check_against_single(ctx, mols, res, o, range(len(mols)))


def forms_geometry(d, key):
    return [(str(s), float(x), float(y), float(z)) for s, (x, y, z) in zip(d[key + "_sym"], d[key + "_xyz"])]


def test_batch_matches_oracle_energies(ctx):
    """Seeded (H2O)_2, (H2O)_3, (H2O)_4 in one batch against the accurate, unscreened oracle SCF
    energies (tests/golden/forms.npz): |dE| <= 1e-8 Eh, orbital energies within 1e-6 Eh."""
    import os
    from conftest import REPO
    d = np.load(os.path.join(REPO, "tests", "golden", "forms.npz"))
    mols = [m.flatten(m.build(forms_geometry(d, "w%d" % k), "6-31G")) for k in (2, 3, 4)]
    o = opts(epsilon=1e-10)
    st, statuses, res = ctx.scf_batch_run(mols, o)
    assert st == abi.FOCK_OK, statuses
    for k, r in zip((2, 3, 4), res):
        dE = abs(r["total_energy"] - d["w%d_energy" % k][0])
        print("(H2O)%d batch |dE| vs oracle %.2e" % (k, dE))
        assert dE <= 1e-8
        assert np.abs(r["orbital_energies"] - d["w%d_eps" % k]).max() < 1e-6


def test_batch_rejects_mixed_precision(ctx):
    mols = clusters(2, seed=3)
    with pytest.raises(abi.FockError) as e:
        ctx.scf_batch_run(mols, opts(precision=abi.PRECISION_MIXED))
    assert e.value.status == abi.FOCK_UNSUPPORTED


@pytest.mark.skip(reason="the device-driven batch iteration that runs the mixed form is switched off: "
                         "its first GPU run ended in a GPU memory fault whose cause has not been found")
def test_batch_mixed_precision_matches_oracle(ctx):
    """Mixed form in the batch: FP32 quartet integrals until each molecule's error falls below
    1e-4, FP64 builds from there (the result always comes from an FP64 build): the seeded (H2O)_2..4
    against the oracle's energies, and incremental builds and the 1e-6 pair screen refused."""
    import os
    from conftest import REPO
    d = np.load(os.path.join(REPO, "tests", "golden", "forms.npz"))
    mols = [m.flatten(m.build(forms_geometry(d, "w%d" % k), "6-31G")) for k in (2, 3, 4)]
    st, statuses, res = ctx.scf_batch_run(mols, opts(epsilon=1e-10, precision=abi.PRECISION_MIXED))
    assert st == abi.FOCK_OK, statuses
    for k, r in zip((2, 3, 4), res):
        dE = abs(r["total_energy"] - d["w%d_energy" % k][0])
        print("(H2O)%d mixed batch |dE| vs oracle %.2e, %d iterations" % (k, dE, r["iterations"]))
        assert dE <= 1e-8
    with pytest.raises(abi.FockError):
        ctx.scf_batch_run(mols, opts(precision=abi.PRECISION_MIXED, incremental=1))
    with pytest.raises(abi.FockError):
        ctx.scf_batch_run(mols, opts(precision=abi.PRECISION_MIXED, screen=abi.SCREEN_REF_BF_PAIR))


def test_python_facade_batch_and_gradient(ctx):
    """try_scf_batch / nuclear_gradient: the scf-facade forms of the batched SCF and the gradient."""
    from fock_amd.scf import nuclear_gradient, try_scf, try_scf_batch
    mols = [m.build(g, "6-31G") for g in m.water_cluster_batch(3, seed=5)]
    sched = dict(ROBUST)
    res = try_scf_batch(mols, 200, 1e-10, ctx=ctx, **sched)
    for atoms, r in zip(mols, res):
        one = try_scf(atoms, 200, 1e-10, 0, policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ, ctx=ctx, **sched)
        assert r.converged and abs(r.total_energy - one.total_energy) < 1e-9
        g = nuclear_gradient(atoms, r, ctx=ctx)
        assert g.shape == (len(atoms), 3) and np.abs(g.sum(axis=0)).max() < 1e-10
