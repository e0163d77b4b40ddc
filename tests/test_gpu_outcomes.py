"""Outcome paths of try_scf (scf/src/lib.rs:73-258) and the result it packages, through the C ABI.

* The reference returns Some at iter == max_iters whatever the error (:224): NOT_CONVERGED with the
  outputs filled; with max_iters = 0 that is iteration 0, compared with the oracle's.
* An odd electron count only warns (:98-100): the density keeps N_e / 2 orbitals.
* None on a non-normal error (:248-249): DIVERGED, also for a singular overlap (coincident atoms),
  which the library detects before any eigensolve.
* Non-finite inputs are rejected before anything runs (INVALID_ARG); f shells are beyond the
  compiled kernels (UNSUPPORTED); Boys orders above 10 under the reference policy (BOYS_ORDER).
* HartreeFockResult (:27-37): the coefficient matrix column-major with orbital k in column k, equal
  to the oracle's up to the sign of each column; MolecularOrbitals keeps |c| > 0.05
  (scf/src/molecular_orbitals.rs:69-90) and evaluate (:37-43) matches the oracle's evaluation.
"""
import numpy as np
import pytest

from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def water():
    return m.flatten(m.h2o_sto3g())


def test_not_converged_at_max_iters_zero_matches_oracle(ctx):
    say("outcome: max_iters = 0")
    from oracle import refscf
    fb, fn = water()
    st, r = ctx.scf_run(fb, fn, abi.default_opts(max_iters=0, epsilon=1e-6))
    assert st == abi.FOCK_NOT_CONVERGED
    want = refscf.try_scf(fb, fn, 0, 1e-6, 0, accurate=False)
    assert want is not None and want.iterations == 0 and r["iterations"] == 0
    assert abs(r["total_energy"] - want.total_energy) < 1e-8
    assert np.abs(r["orbital_energies"] - want.orbital_energies).max() < 1e-7
    assert np.abs(r["density"] - want.density_matrix).max() < 1e-7
    # the facade returns the result (the reference's Some), marked unconverged
    from fock_amd.scf import try_scf
    res = try_scf(m.h2o_sto3g(), 0, 1e-6, 0, ctx=ctx)
    assert res is not None and not res.converged and res.iterations == 0


def test_iteration_limit_returns_last_iterate(ctx):
    say("outcome: iteration limit")
    from oracle import refscf
    fb, fn = water()
    st, r = ctx.scf_run(fb, fn, abi.default_opts(max_iters=3, epsilon=1e-12))
    want = refscf.try_scf(fb, fn, 3, 1e-12, 0, accurate=False)
    assert st == abi.FOCK_NOT_CONVERGED and r["iterations"] == 3 == want.iterations
    assert abs(r["total_energy"] - want.total_energy) < 1e-8


def test_odd_electron_count_follows_the_reference(ctx):
    """H2O+ (9 electrons): the reference warns and occupies N_e / 2 = 4 orbitals."""
    say("outcome: odd electron count")
    from oracle import refscf
    fb, fn = water()
    o = abi.default_opts(max_iters=100, epsilon=1e-10, molecule_charge=1, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_NONE)
    st, r = ctx.scf_run(fb, fn, o)
    want = refscf.try_scf(fb, fn, 100, 1e-10, 1, accurate=True)
    assert r["n_electrons"] == 9 == want.n_electrons
    assert st in (abi.FOCK_OK, abi.FOCK_NOT_CONVERGED)
    assert abs(r["total_energy"] - want.total_energy) < 1e-8
    S = ctx.one_electron(fb, fn, abi.POLICY_ACCURATE)[0]
    assert abs(np.trace(r["density"] @ S) - 8.0) < 1e-9  # 2 x 4 occupied orbitals


def coincident_water():
    """H2O with both hydrogens on one point: two equal basis functions, a singular overlap."""
    geom = m.water_geometry()
    geom = [geom[0], geom[1], (geom[1][0],) + tuple(geom[1][1:])]
    return m.flatten(m.build(geom, "STO-3G"))


@pytest.mark.parametrize("policy", [abi.POLICY_REF_COMPAT, abi.POLICY_ACCURATE])
def test_singular_overlap_is_diverged(ctx, policy):
    say("outcome: coincident atoms, policy %d" % policy)
    fb, fn = coincident_water()
    o = abi.default_opts(max_iters=50, epsilon=1e-8, integral_policy=policy,
                         screen=abi.SCREEN_REF_BF_PAIR if policy == 0 else abi.SCREEN_NONE)
    st, r = ctx.scf_run(fb, fn, o)
    assert st == abi.FOCK_DIVERGED
    assert r["n_electrons"] == 10 and r["n_basis"] == 7
    assert "singular" in ctx.lib.fock_ctx_last_error(ctx.ptr).decode()
    from fock_amd.scf import try_scf
    geom = m.water_geometry()
    assert try_scf(m.build([geom[0], geom[1], geom[1]], "STO-3G"), 50, 1e-8, 0, ctx=ctx) is None


def test_batch_marks_singular_molecule_diverged(ctx):
    say("outcome: batch with a singular molecule")
    good = m.flatten(m.build(m.water_geometry(), "6-31G"))
    geom = m.water_geometry()
    bad = m.flatten(m.build([geom[0], geom[1], geom[1]], "6-31G"))
    o = abi.default_opts(max_iters=100, epsilon=1e-9, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12)
    st, statuses, res = ctx.scf_batch_run([good, bad, good], o)
    assert st == abi.FOCK_DIVERGED
    assert statuses == [abi.FOCK_OK, abi.FOCK_DIVERGED, abi.FOCK_OK]
    st1, one = ctx.scf_run(good[0], good[1], o)
    assert st1 == abi.FOCK_OK
    assert abs(res[0]["total_energy"] - one["total_energy"]) < 1e-9
    assert abs(res[2]["total_energy"] - one["total_energy"]) < 1e-9


def _with(fb, **kw):
    d = dict(centers=fb.centers, angular=fb.angular, prim_offset=fb.prim_offset, exponents=fb.exponents,
             coefficients=fb.coefficients)
    d.update(kw)
    return m.FlatBasis(**d)


@pytest.mark.parametrize("what", ["centre", "exponent", "coefficient", "nucleus", "charge"])
def test_non_finite_input_is_rejected(ctx, what):
    say("outcome: non-finite " + what)
    fb, fn = water()
    if what == "centre":
        c = fb.centers.copy(); c[3:6, 1] = np.nan  # the O 2p shell
        fb = _with(fb, centers=np.ascontiguousarray(c))
    elif what == "exponent":
        e = fb.exponents.copy(); e[4] = np.inf
        fb = _with(fb, exponents=e)
    elif what == "coefficient":
        c = fb.coefficients.copy(); c[0] = np.nan
        fb = _with(fb, coefficients=c)
    elif what == "nucleus":
        p = fn.positions.copy(); p[1, 2] = np.inf
        fn = m.FlatNuclei(np.ascontiguousarray(p), fn.charges)
    else:
        q = fn.charges.copy(); q[0] = np.nan
        fn = m.FlatNuclei(fn.positions, q)
    with pytest.raises(abi.FockError) as e:
        ctx.scf_run(fb, fn, abi.default_opts())
    assert e.value.status == abi.FOCK_INVALID_ARG
    with pytest.raises(abi.FockError) as e:
        ctx.one_electron(fb, fn, abi.POLICY_ACCURATE)
    assert e.value.status == abi.FOCK_INVALID_ARG


def test_f_shell_is_unsupported(ctx):
    say("outcome: f shell")
    fb, fn = water()
    comps = m.cartesian_components(3)
    n = len(comps)
    f = m.FlatBasis(np.ascontiguousarray(np.zeros((n, 3))), np.ascontiguousarray(np.array(comps, dtype=np.int32)),
                    np.ascontiguousarray(np.arange(n + 1, dtype=np.int32)), np.ones(n),
                    np.ones(n))
    with pytest.raises(abi.FockError) as e:
        ctx.scf_run(f, fn, abi.default_opts())
    assert e.value.status == abi.FOCK_UNSUPPORTED


def test_gradient_rejects_reference_pair_screen(ctx):
    say("outcome: gradient with the 1e-6 pair screen")
    fb, fn = water()
    o = abi.default_opts(max_iters=100, epsilon=1e-10, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_REF_BF_PAIR)
    st, r = ctx.scf_run(fb, fn, o)
    assert st == abi.FOCK_OK
    with pytest.raises(abi.FockError) as e:
        ctx.rhf_gradient(fb, fn, o, r)
    assert e.value.status == abi.FOCK_INVALID_ARG
    from fock_amd.scf import nuclear_gradient, try_scf
    res = try_scf(m.h2o_sto3g(), 100, 1e-10, 0, policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_REF_BF_PAIR, ctx=ctx)
    with pytest.raises(ValueError):
        nuclear_gradient(m.h2o_sto3g(), res, ctx=ctx)
    res = try_scf(m.h2o_sto3g(), 100, 1e-10, 0, policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ, ctx=ctx)
    g = nuclear_gradient(m.h2o_sto3g(), res, ctx=ctx)
    assert np.abs(g.sum(axis=0)).max() < 1e-10


# ------------------------------------------------------------------ the packaged result

def aligned(C, ref):
    """ref with each column's sign chosen to match C (eigenvector signs are arbitrary)."""
    s = np.sign(np.sum(C * ref, axis=0))
    s[s == 0] = 1.0
    return ref * s


@pytest.mark.parametrize("case", ["w7_ref", "w13_acc", "o108_acc"])
def test_coefficient_matrix_matches_oracle_up_to_column_sign(ctx, case):
    say("result: coefficient matrix " + case)
    from oracle import refscf
    if case == "w7_ref":
        atoms, kw, okw = m.h2o_sto3g(), {}, dict(accurate=False)
        eps = 1e-12
    elif case == "w13_acc":
        atoms = m.build(m.water_geometry(), "6-31G")
        kw, okw, eps = dict(policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_NONE), dict(accurate=True), 1e-12
    else:
        atoms = m.octane_631g()
        kw = dict(policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_NONE)
        okw, eps = dict(accurate=True), 1e-10
    from fock_amd.scf import try_scf
    r = try_scf(atoms, 300, eps, 0, ctx=ctx, **kw)
    fb, fn = m.flatten(atoms)
    want = refscf.try_scf(fb, fn, 300, eps, 0, **okw)
    n = fb.n
    C = r.coefficient_matrix
    assert C.shape == (n, n)
    # the C ABI's column-major layout read back as column k = orbital k: C^T S C = I
    S = ctx.one_electron(fb, fn, kw.get("policy", 0))[0]
    assert np.abs(C.T @ S @ C - np.eye(n)).max() < 1e-9
    ref = aligned(C, want.coefficient_matrix)
    # an orbital is determined up to the rotations its neighbours allow: a perturbation of the
    # Fock matrix of size e moves orbital k by ~e / gap_k (gap_k: distance of eps_k to the nearest
    # other orbital energy), so each column is compared at 1e-8 / gap_k (both SCFs converged to
    # max|FPS - SPF| <= 1e-10); orbitals with gaps below 1e-4 Eh are left to the density checks
    gaps = np.minimum(np.abs(np.diff(want.orbital_energies, prepend=-np.inf)),
                      np.abs(np.diff(want.orbital_energies, append=np.inf)))
    sep = gaps > 1e-4
    err = np.abs(C - ref).max(axis=0)
    assert np.all(err[sep] * gaps[sep] < 1e-8), (err * gaps)[sep].max()
    nocc = want.n_electrons // 2
    assert np.count_nonzero(sep[:nocc]) >= nocc // 2   # most occupied orbitals are tested
    assert np.abs(r.density_matrix - want.density_matrix).max() < 1e-7


def test_molecular_orbitals_truncate_and_evaluate_like_the_oracle(ctx):
    say("result: MolecularOrbitals.evaluate")
    from oracle import refmo, refscf
    from fock_amd.scf import try_scf
    atoms = m.build(m.water_geometry(), "6-31G")
    r = try_scf(atoms, 300, 1e-12, 0, policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_NONE, ctx=ctx)
    fb, fn = m.flatten(atoms)
    want = refscf.try_scf(fb, fn, 300, 1e-12, 0, accurate=True)
    C = r.coefficient_matrix
    ref = aligned(C, want.coefficient_matrix)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-4.0, 4.0, size=(40, 3))
    for k in range(fb.n):
        idx, _ = refmo.truncate(ref, k)
        mo = r.orbitals[k]
        assert len(mo.coefficients) == len(idx)
        # no coefficient sits so close to 0.05 that rounding could move it across
        assert np.all(np.abs(np.abs(C[:, k]) - 0.05) > 1e-6)
        got = np.array([mo.evaluate(p) for p in pts])
        exp = refmo.evaluate(fb, ref, k, pts)
        assert np.abs(got - exp).max() < 1e-6 * max(1.0, np.abs(exp).max()), k
