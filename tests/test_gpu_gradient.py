"""Analytic RHF nuclear gradient on the GPU (differentiation form) through the C ABI.

Pins: central finite differences (step 1e-4 Bohr) of the accurate oracle energy for water, of
the GPU energy for a seeded (H2O)4 cluster and n-octane (a subset of coordinates), all to
1e-6 Eh/Bohr; translation invariance (exact by construction) and rotation invariance
(torque-free forces of an isolated molecule).
"""
import numpy as np
import pytest

from fock_amd import abi, molecule as m

pytestmark = pytest.mark.gpu

TOL = 1e-6  # Eh/Bohr
# the robust schedule of the large molecules (DIIS from 2 in a window of 10, 0.5 Eh level shift
# until max|FDS - SDF| < 0.05)
ROBUST = dict(diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)


def opts(screen=abi.SCREEN_NONE, tau=0.0, eps=1e-12, **kw):
    return abi.default_opts(max_iters=300, epsilon=eps, integral_policy=abi.POLICY_ACCURATE, screen=screen,
                            schwarz_threshold=tau, **kw)


def energy(ctx, fb, fn, o):
    st, r = ctx.scf_run(fb, fn, o)
    assert st == abi.FOCK_OK, abi.STATUS_NAMES[st]
    return r


def gpu_fd(ctx, fb, fn, o, coords, h=1e-4):
    out = {}
    for a, x in coords:
        e = []
        for s in (+1.0, -1.0):
            b2, n2 = m.displaced(fb, fn, a, x, s * h)
            e.append(energy(ctx, b2, n2, o)["total_energy"])
        out[(a, x)] = (e[0] - e[1]) / (2.0 * h)
    return out


def torque(fn, g):
    c = fn.positions.mean(axis=0)
    return np.cross(fn.positions - c, g).sum(axis=0)


@pytest.mark.parametrize("basis", ["STO-3G", "6-31G"])
def test_gradient_water_vs_oracle_fd(ctx, basis):
    from oracle import refscf
    fb, fn = m.flatten(m.build(m.water_geometry(), basis))
    o = opts(eps=1e-13)
    r = energy(ctx, fb, fn, o)
    g = ctx.rhf_gradient(fb, fn, o, r)
    want = refscf.fd_gradient(fb, fn)
    print(basis, "max |analytic - fd| = %.2e" % np.abs(g - want).max())
    assert np.abs(g - want).max() < TOL
    assert np.abs(g.sum(axis=0)).max() < 1e-10


@pytest.mark.parametrize("key,basis", [("grad_w4", "6-31G"), ("grad_hf", "6-311G")])
def test_gradient_vs_oracle_fd_fixture(ctx, key, basis):
    """The seeded (H2O)4 and HF/6-311G (F exponents to 1.1e4: the gradient's primitive cut scales
    with the largest exponent) against central finite differences of the oracle's accurate,
    unscreened energy (tests/golden/forms.npz, step 1e-4 Bohr)."""
    import os
    from conftest import REPO
    d = np.load(os.path.join(REPO, "tests", "golden", "forms.npz"))
    geom = [(str(s), float(x), float(y), float(z)) for s, (x, y, z) in zip(d[key + "_sym"], d[key + "_xyz"])]
    fb, fn = m.flatten(m.build(geom, basis))
    o = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-13, eps=1e-12, **ROBUST)
    r = energy(ctx, fb, fn, o)
    g = ctx.rhf_gradient(fb, fn, o, r)
    err = np.abs(g - d[key + "_fd"]).max()
    print("%s max |analytic - oracle fd| = %.2e" % (key, err))
    assert err < TOL
    assert np.abs(g.sum(axis=0)).max() < 1e-10


def test_gradient_water_tetramer_vs_gpu_fd(ctx):
    rng = np.random.default_rng(4)
    fb, fn = m.flatten(m.build(m.water_cluster_geometry(4, rng), "6-31G"))
    o = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-13, eps=1e-12)
    r = energy(ctx, fb, fn, o)
    g = ctx.rhf_gradient(fb, fn, o, r)
    fd = gpu_fd(ctx, fb, fn, o, [(a, x) for a in range(len(fn.charges)) for x in range(3)])
    err = max(abs(g[a, x] - v) for (a, x), v in fd.items())
    print("(H2O)4 max |analytic - fd| = %.2e, max |g| = %.2e" % (err, np.abs(g).max()))
    assert err < TOL
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    assert np.abs(torque(fn, g)).max() < 1e-6


def test_gradient_octane_vs_gpu_fd(ctx):
    fb, fn = m.flatten(m.octane_631g())
    o = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-12, eps=1e-11, **ROBUST)
    r = energy(ctx, fb, fn, o)
    g = ctx.rhf_gradient(fb, fn, o, r)
    coords = [(0, 0), (0, 1), (3, 2), (9, 0), (17, 1), (25, 2)]
    fd = gpu_fd(ctx, fb, fn, o, coords, h=1e-3)
    err = max(abs(g[a, x] - v) for (a, x), v in fd.items())
    print("C8H18 max |analytic - fd| = %.2e over %d coordinates" % (err, len(coords)))
    assert err < TOL
    assert np.abs(g.sum(axis=0)).max() < 1e-10
    assert np.abs(torque(fn, g)).max() < 1e-5


def test_gradient_reuses_or_rebuilds_prepared_lists(ctx):
    """The gradient reuses the task lists the SCF prepared when basis and options are the same,
    and prepares again otherwise: A after the SCF of A, B (A displaced) after the SCF of B, then A
    again (lists of B held) and A at another Schwarz threshold give A's gradient."""
    fa, na = m.flatten(m.build(m.water_cluster_geometry(2, np.random.default_rng(9)), "6-31G"))
    fbb, nb = m.displaced(fa, na, 1, 2, 0.05)
    o = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-12, eps=1e-12)
    ra = energy(ctx, fa, na, o)
    ga = ctx.rhf_gradient(fa, na, o, ra)
    rb = energy(ctx, fbb, nb, o)
    gb = ctx.rhf_gradient(fbb, nb, o, rb)
    ga2 = ctx.rhf_gradient(fa, na, o, ra)
    ga3 = ctx.rhf_gradient(fa, na, opts(screen=abi.SCREEN_SCHWARZ, tau=1e-14, eps=1e-12), ra)
    assert np.abs(gb - ga).max() > 1e-3
    assert np.abs(ga2 - ga).max() < 1e-12
    assert np.abs(ga3 - ga).max() < 1e-9
    # the preparation count: a gradient after the SCF of the same basis and options reuses its
    # lists, one after another basis or threshold prepares again
    p0 = ctx.jk_stats()["preparations"]
    energy(ctx, fa, na, o)                       # SCF of A: one preparation
    p1 = ctx.jk_stats()["preparations"]
    ctx.rhf_gradient(fa, na, o, ra)              # reuses A's lists
    p2 = ctx.jk_stats()["preparations"]
    ctx.rhf_gradient(fbb, nb, o, rb)             # B: prepares
    p3 = ctx.jk_stats()["preparations"]
    ctx.rhf_gradient(fbb, nb, opts(screen=abi.SCREEN_SCHWARZ, tau=1e-14, eps=1e-12), rb)  # new threshold
    p4 = ctx.jk_stats()["preparations"]
    assert (p1 - p0, p2 - p1, p3 - p2, p4 - p3) == (1, 0, 1, 1)


def test_gradient_rejects_reference_policy(ctx):
    fb, fn = m.flatten(m.h2o_sto3g())
    o = opts()
    r = energy(ctx, fb, fn, o)
    bad = abi.default_opts(integral_policy=abi.POLICY_REF_COMPAT)
    with pytest.raises(abi.FockError):
        ctx.rhf_gradient(fb, fn, bad, r)


def test_autograd_function_matches_fd(ctx):
    """torch.autograd over the C ABI: backward of the energy is the analytic gradient."""
    import torch
    from fock_amd.autograd import rhf_energy
    fb, fn = m.flatten(m.h2o_sto3g())
    o = opts(eps=1e-13)
    pos = torch.tensor(fn.positions, dtype=torch.float64, requires_grad=True)
    e = rhf_energy(pos, fb, fn, ctx, o)
    e.backward()
    from oracle import refscf
    want = refscf.fd_gradient(fb, fn)
    assert abs(float(e) - refscf.try_scf(fb, fn, 400, 1e-13, 0, accurate=True).total_energy) < 1e-10
    assert np.abs(pos.grad.numpy() - want).max() < TOL


@pytest.mark.parametrize("which", ["water4", "octane"])
def test_gradient_mixed_precision_matches_fp64(ctx, which):
    """Mixed form: derivative integrals of small-bound quartets in FP32 (FP64 above a Schwarz bound
    of 1e-4), the same converged density; within 1e-6 Eh/Bohr of the FP64 gradient."""
    if which == "water4":
        fb, fn = m.flatten(m.build(m.water_cluster_geometry(4, np.random.default_rng(4)), "6-31G"))
        o = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-12, eps=1e-10)
    else:
        fb, fn = m.flatten(m.octane_631g())
        o = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-12, eps=1e-10, **ROBUST)
    r = energy(ctx, fb, fn, o)
    g64 = ctx.rhf_gradient(fb, fn, o, r)
    om = opts(screen=abi.SCREEN_SCHWARZ, tau=1e-12, eps=1e-10, precision=abi.PRECISION_MIXED)
    g32 = ctx.rhf_gradient(fb, fn, om, r)
    err = np.abs(g32 - g64).max()
    print(which, "max |mixed - fp64| = %.2e" % err)
    assert err < TOL
    assert np.abs(g32.sum(axis=0)).max() < 1e-9
