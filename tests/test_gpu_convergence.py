"""SCF robustness options of fock_scf_run (no reference counterpart beyond the fixed mixing factor
F = 1.0 of scf/src/lib.rs:218-219; the reference README admits convergence struggles,
README.md:13):

* damping: P <- F P_new + (1 - F) P_old while the error exceeds damping_until; the GPU trajectory
  follows the oracle's (oracle/refscf.py with the same factor) iteration by iteration;
* EDIIS (Kudin-Scuseria-Cances): while the error exceeds ediis_until, the convex combination of the
  stored Fock matrices minimising the interpolated energy; the GPU's active-set solution gives the
  oracle's trajectory (whose coefficients come from an exhaustive search over the simplex faces);
* on n-eicosane C20H42/6-31G with the reference's own DIIS schedule (start 5, window 16, no level
  shift), EDIIS reaches the oracle's energy.
"""
import os

import numpy as np
import pytest

from conftest import REPO
from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run(ctx, fb, fn, **kw):
    base = dict(max_iters=200, epsilon=1e-10, integral_policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_NONE)
    base.update(kw)
    return ctx.scf_run(fb, fn, abi.default_opts(**base))


@pytest.mark.parametrize("form", ["damping", "ediis"])
def test_trajectory_matches_oracle(ctx, form):
    say("convergence option: " + form)
    from oracle import refscf
    fb, fn = m.flatten(m.build(m.water_geometry(), "6-31G"))
    if form == "damping":
        kw = dict(damping=0.6, damping_until=1e-3)
    else:
        kw = dict(ediis=1, ediis_until=1e-2, diis_start=2, diis_window=8)
    st, r = run(ctx, fb, fn, **kw)
    assert st == abi.FOCK_OK
    want = refscf.try_scf(fb, fn, 200, 1e-10, 0, accurate=True, **{k: v for k, v in kw.items()})
    got_traj = np.array([h[0] for h in r["history"]])
    want_traj = np.array(want.energies)
    say("  %d iterations (oracle %d), max |dE| along the trajectory %.2e" %
        (r["iterations"], want.iterations, np.abs(got_traj[:len(want_traj)] - want_traj[:len(got_traj)]).max()))
    assert r["iterations"] == want.iterations
    assert np.abs(got_traj - want_traj).max() < 1e-8
    assert abs(r["total_energy"] - want.total_energy) < 1e-9


def test_ediis_converges_eicosane_with_reference_diis_schedule(ctx):
    say("EDIIS on C20H42 with the reference DIIS schedule")
    gold = np.load(os.path.join(REPO, "tests", "golden", "e264.npz"))
    fb, fn = m.flatten(m.eicosane_631g())
    common = dict(max_iters=100, epsilon=1e-8, screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12)
    st0, r0 = run(ctx, fb, fn, **common)             # the reference's schedule alone
    st1, r1 = run(ctx, fb, fn, ediis=1, ediis_until=1e-1, **common)
    say("  DIIS only: %s after %d iterations (error %.2e); EDIIS: %s after %d iterations, |dE| %.2e" %
        (abi.STATUS_NAMES[st0], r0["iterations"], r0["last_error"], abi.STATUS_NAMES[st1], r1["iterations"],
         abs(r1["total_energy"] - gold["e264_acc_energy"][0])))
    assert st1 == abi.FOCK_OK
    assert abs(r1["total_energy"] - gold["e264_acc_energy"][0]) < 1e-7
    if st0 == abi.FOCK_OK:
        assert r1["iterations"] <= r0["iterations"]


def test_batch_rejects_single_molecule_options(ctx):
    mols = [m.flatten(m.h2o_sto3g())]
    for kw in (dict(ediis=1), dict(damping=0.5)):
        o = abi.default_opts(integral_policy=abi.POLICY_ACCURATE, screen=abi.SCREEN_SCHWARZ, **kw)
        with pytest.raises(abi.FockError) as e:
            ctx.scf_batch_run(mols, o)
        assert e.value.status == abi.FOCK_UNSUPPORTED
