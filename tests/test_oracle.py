"""Pins of the CPU oracle (oracle/refcpu.c, oracle/oracle64.c, oracle/refscf.py).

Each check is chosen so a plausible restatement mistake (dropped term, wrong sign or index,
transposed operand) fails it: reference golden values, closed forms, mpmath at 40 digits,
derivative identities, and agreement between the two independent implementations.
"""
import ctypes
import math
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import REPO
from fock_amd import molecule as m

mp.mp.dps = 40
GOLD = os.path.join(REPO, "tests", "golden")


def boys_mp(T, n):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * n + 1)
    return mp.hyp1f1(n + mp.mpf(1) / 2, n + mp.mpf(3) / 2, -T) / (2 * n + 1)


# ------------------------------------------------------------------ reference goldens

def test_hermite_expansion_doctest_bit_exact(oracle_lib):
    # basis/src/utils.rs:24
    assert oracle_lib.ref_hermite_expansion(1, 2, 3, 2.5, 3.0, 5.0) == 1.9872458203327215e-9


def test_lut_is_f10_on_the_0p15_grid():
    vals = [float(x) for x in open(os.path.join(REPO, "fock-rs_amd", "csrc", "ref_boys_lut.inc")).read()
            .split("*/")[1].replace("\n", "").split(",") if x.strip()]
    assert len(vals) == 200
    for k in range(0, 200, 7):
        want = boys_mp(0.15 * k, 10)
        assert abs(vals[k] - want) / want < 4e-13


# ------------------------------------------------------------------ Boys

def test_ref_boys_regimes(oracle_lib):
    f = oracle_lib.ref_boys
    for n in range(11):
        assert f(0.0, n) == 1.0 / (2 * n + 1)          # t < 1e-16 branch
        for T in (30.0, 31.7, 100.0):                   # asymptotic branch, boys/src/lib.rs:66-69
            df = 1
            for k in range(2 * n - 1, 0, -2):
                df *= k
            want = df * (1.0 / 2 ** (n + 1)) * math.sqrt(math.pi * oracle_lib.ref_powi(T, -2 * n - 1))
            assert f(T, n) == want
    # the quirks the reference documents: ~1e-3 relative error and a jump at T = 30
    errs = [abs(f(T, 0) - float(boys_mp(T, 0))) / float(boys_mp(T, 0)) for T in np.linspace(0.01, 29.8, 60)]
    assert 1e-6 < max(errs) < 2e-2
    jump = abs(f(29.999, 0) - f(30.0, 0)) / f(30.0, 0)
    assert 0.02 < jump < 0.06
    # downward recursion identity F_{n-1} = (2T F_n + e^-T)/(2n-1) holds exactly for the LUT branch
    T = 7.3
    for n in range(1, 11):
        lhs = f(T, n - 1)
        rhs = (2.0 * T * f(T, n) + math.exp(-T)) / (2 * n - 1)
        assert lhs == rhs


def test_oracle64_boys_vs_mpmath(oracle_lib):
    for n in (0, 1, 2, 4, 7, 11):
        for T in (0.0, 1e-9, 0.37, 3.0, 14.9, 29.99, 44.9, 45.1, 80.0, 400.0):
            want = boys_mp(T, n)
            got = oracle_lib.o64_boys(T, n)
            assert abs(got - want) / want < 2e-15, (n, T)


def test_ref_boys_order_limit_flags(oracle_lib):
    assert math.isnan(oracle_lib.ref_boys(1.0, 11))
    assert oracle_lib.ref_boys_order_error() == 1


# ------------------------------------------------------------------ Hermite Coulomb R

def _R(lib, t, u, v, n, p, d):
    arr = (ctypes.c_double * 3)(*d)
    return lib.ref_coulomb_auxiliary(t, u, v, n, p, arr, d[0] ** 2 + d[1] ** 2 + d[2] ** 2)


def test_x_branch_defect_closed_form(oracle_lib):
    """basis/src/utils.rs:98-105 drops (t-1): R_bug(3,0,0) - R_ok(3,0,0) = -R^{(1)}_{100}
    = -X (-2p)^2 F_2(T) for n = 0 (and likewise shifted for n > 0)."""
    L = oracle_lib
    p, d = 0.83, (0.7, -0.4, 1.1)
    T = p * (d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    try:
        for n in (0, 1):
            L.ref_set_policy(1, 0)
            bug = _R(L, 3, 0, 0, n, p, d)
            L.ref_set_policy(1, 1)
            ok = _R(L, 3, 0, 0, n, p, d)
            want = -d[0] * (-2 * p) ** (n + 2) * L.o64_boys(T, n + 2)
            assert abs((bug - ok) - want) < 1e-14
            # t <= 2 and every y/z branch are unaffected
            for tuv in [(2, 0, 0), (1, 1, 0), (0, 0, 3), (0, 3, 0), (2, 1, 1), (0, 2, 2)]:
                L.ref_set_policy(1, 0)
                a = _R(L, *tuv, n, p, d)
                L.ref_set_policy(1, 1)
                b = _R(L, *tuv, n, p, d)
                assert a == b
        # axis permutation symmetry R_400(X,Y,Z) = R_004(Z,Y,X): holds for the correct recursion only
        L.ref_set_policy(1, 1)
        assert abs(_R(L, 4, 0, 0, 0, p, d) - _R(L, 0, 0, 4, 0, p, (d[2], d[1], d[0]))) < 1e-14
        L.ref_set_policy(1, 0)
        assert abs(_R(L, 4, 0, 0, 0, p, d) - _R(L, 0, 0, 4, 0, p, (d[2], d[1], d[0]))) > 1e-3
    finally:
        L.ref_set_policy(0, 0)


# ------------------------------------------------------------------ integrals vs closed forms

def _flat(prims):
    """prims: list of (centre, angular, exponent) with coefficient 1 (unnormalised)."""
    fb = m.FlatBasis(
        np.ascontiguousarray(np.array([c for c, _, _ in prims], dtype=np.float64)),
        np.ascontiguousarray(np.array([a for _, a, _ in prims], dtype=np.int32)),
        np.ascontiguousarray(np.arange(len(prims) + 1, dtype=np.int32)),
        np.ascontiguousarray(np.array([e for _, _, e in prims], dtype=np.float64)),
        np.ascontiguousarray(np.ones(len(prims))),
    )
    return fb


def ssss_mp(A, a, B, b, C, c, D, d):
    A, B, C, D = [mp.matrix(x) for x in (A, B, C, D)]
    a, b, c, d = map(mp.mpf, (a, b, c, d))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    nrm = lambda v: sum(x ** 2 for x in v)
    Kab = mp.e ** (-a * b / p * nrm(A - B))
    Kcd = mp.e ** (-c * d / q * nrm(C - D))
    alpha = p * q / (p + q)
    return 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * Kab * Kcd * boys_mp(alpha * nrm(P - Q), 0)


GEOM = [((0.1, -0.2, 0.3), 1.3), ((1.0, 0.4, -0.5), 0.7), ((-0.6, 0.9, 0.2), 2.1), ((0.3, 1.4, 1.0), 0.45)]


def _oracle_quartet(oracle_lib, angs):
    """One primitive ERI from both oracles: refcpu (per function, accurate switches on) and
    oracle64 (per complete shell, the requested component picked out)."""
    import oracle
    prims, first, comp = [], [], []
    for i, ang in enumerate(angs):
        L = sum(ang)
        comps = m.cartesian_components(L)
        first.append(len(prims))
        comp.append(comps.index(tuple(ang)))
        for c in comps:
            prims.append((GEOM[i][0], c, GEOM[i][1]))
    fb = _flat(prims)
    bv = oracle.BasisView(fb)
    oracle_lib.ref_set_policy(1, 1)
    try:
        ref = oracle_lib.ref_eri(bv.ref, first[0] + comp[0], first[1] + comp[1], first[2] + comp[2],
                                 first[3] + comp[3])
    finally:
        oracle_lib.ref_set_policy(0, 0)
    nf = [len(m.cartesian_components(sum(a))) for a in angs]
    out = np.zeros(int(np.prod(nf)))
    assert oracle_lib.o64_shell_quartet(bv.ref, first[0], first[1], first[2], first[3], oracle.dptr(out)) == 0
    idx = ((comp[0] * nf[1] + comp[1]) * nf[2] + comp[2]) * nf[3] + comp[3]
    return ref, out[idx]


def test_ssss_closed_form(oracle_lib):
    want = ssss_mp(GEOM[0][0], GEOM[0][1], GEOM[1][0], GEOM[1][1], GEOM[2][0], GEOM[2][1], GEOM[3][0], GEOM[3][1])
    ref, o64 = _oracle_quartet(oracle_lib, [(0, 0, 0)] * 4)
    assert abs(ref - want) / want < 1e-13
    assert abs(o64 - want) / want < 1e-13


@pytest.mark.parametrize("angs", [
    [(1, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0)],
    [(0, 1, 0), (0, 0, 1), (0, 0, 0), (0, 0, 0)],
    [(1, 0, 0), (0, 0, 0), (0, 1, 0), (0, 0, 0)],
    [(0, 0, 1), (0, 0, 0), (0, 0, 1), (1, 0, 0)],
    [(1, 0, 0), (1, 0, 0), (1, 0, 0), (1, 0, 0)],
])
def test_p_integrals_are_centre_derivatives(oracle_lib, angs):
    """A reference p function (A_x - x) e^{-a|r-A|^2} equals -(1/2a) d/dA_x of the s primitive,
    so every p-containing primitive ERI is a mixed centre derivative of the (ss|ss) closed form."""
    mp.mp.dps = 30
    centres = [list(map(mp.mpf, g[0])) for g in GEOM]
    exps = [mp.mpf(g[1]) for g in GEOM]
    vars_, orders, scale = [], [], mp.mpf(1)
    for i, ang in enumerate(angs):
        for ax in range(3):
            if ang[ax]:
                vars_.append((i, ax))
                orders.append(ang[ax])
                scale *= -1 / (2 * exps[i])

    def f(*xs):
        C = [list(c) for c in centres]
        for (i, ax), x in zip(vars_, xs):
            C[i][ax] = x
        return ssss_mp(C[0], exps[0], C[1], exps[1], C[2], exps[2], C[3], exps[3])

    x0 = [centres[i][ax] for i, ax in vars_]
    want = float(scale * mp.diff(f, tuple(x0), tuple(orders)))
    ref, o64 = _oracle_quartet(oracle_lib, angs)
    mp.mp.dps = 40
    assert abs(o64 - want) < 1e-12 * max(1.0, abs(want)), (o64, want)
    assert abs(ref - want) < 1e-12 * max(1.0, abs(want)), (ref, want)


def test_szabo_ostlund_h2(oracle_lib):
    """Szabo & Ostlund, Modern Quantum Chemistry, H2/STO-3G at R = 1.4 a0 (4 decimals)."""
    from oracle import refscf
    b = m.BasisSet.load("STO-3G")
    fb, fn = m.flatten([b.get((0, 0, 0), 1), b.get((0, 0, 1.4), 1)])
    S, T, V = refscf.one_electron(fb, fn, True)
    packed, _ = refscf.eri_packed(fb, True)
    D = refscf.dense_from_packed(2, packed)
    assert round(S[0, 1], 4) == 0.6593
    assert round(D[0, 0, 0, 0], 4) == 0.7746
    assert round(D[0, 0, 1, 1], 4) == 0.5697
    assert round(D[0, 1, 0, 1], 4) == 0.2970
    assert round(D[0, 0, 0, 1], 4) == 0.4441
    r = refscf.try_scf(fb, fn, 100, 1e-10, 0, accurate=True)
    assert round(r.total_energy, 4) == -1.1167


def test_refcpu_with_fixed_policies_equals_oracle64(oracle_lib):
    """The restatement with the accurate Boys and the (t-1) factor restored is an independent
    second implementation of the accurate oracle: whole W13 tensor and one-electron matrices."""
    import oracle
    from oracle import refscf
    fb, fn = m.flatten(m.build(m.water_geometry(), "6-31G"))
    oracle_lib.ref_set_policy(1, 1)
    try:
        packed_ref = np.zeros(oracle_lib.ref_packed_size(fb.n))
        oracle_lib.ref_electron_tensor(oracle.BasisView(fb).ref, 0, oracle.dptr(packed_ref))
        S = np.zeros((fb.n, fb.n)); T = np.zeros_like(S); V = np.zeros_like(S)
        pos, ch = np.ascontiguousarray(fn.positions), np.ascontiguousarray(fn.charges)
        oracle_lib.ref_one_electron(oracle.BasisView(fb).ref, len(ch), oracle.dptr(pos), oracle.dptr(ch),
                                    oracle.dptr(S), oracle.dptr(T), oracle.dptr(V))
    finally:
        oracle_lib.ref_set_policy(0, 0)
    packed_o64, _ = refscf.eri_packed(fb, True)
    assert np.abs(packed_ref - packed_o64).max() < 1e-13
    S2, T2, V2 = refscf.one_electron(fb, fn, True)
    assert np.abs(S - S2).max() < 1e-14 and np.abs(T - T2).max() < 1e-13 and np.abs(V - V2).max() < 1e-12


def test_eight_fold_symmetry_and_schwarz(oracle_lib):
    from oracle import refscf
    fb, fn = m.flatten(m.h2o_sto3g())
    packed, _ = refscf.eri_packed(fb, True)
    D = refscf.dense_from_packed(fb.n, packed)
    n = fb.n
    for (i, j, k, l) in [(0, 1, 2, 3), (4, 2, 6, 1), (5, 5, 3, 0)]:
        v = D[i, j, k, l]
        for perm in [(j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)]:
            assert D[perm] == v
    diag = np.array([[D[i, j, i, j] for j in range(n)] for i in range(n)])
    assert np.all(np.abs(D) <= np.sqrt(diag[:, :, None, None] * diag[None, None, :, :]) + 1e-14)


# ------------------------------------------------------------------ SCF restatement

def test_rotation_invariance_accurate_and_defect_ref():
    g = np.load(os.path.join(GOLD, "small.npz"))
    assert abs(g["w7_acc_energy"][0] - g["w7rot_acc_energy"][0]) < 1e-8
    # the reference's x-branch defect makes its energy orientation dependent
    assert abs(g["w7_ref_energy"][0] - g["w7rot_ref_energy"][0]) > 1e-3


def test_scf_invariants_on_fixture():
    g = np.load(os.path.join(GOLD, "small.npz"))
    from oracle import refscf
    fb, fn = m.flatten(m.h2o_sto3g())
    S, _, _ = refscf.one_electron(fb, fn, True)
    P = g["w7_acc_P"]
    assert abs(np.trace(P @ S) - 10.0) < 1e-9        # tr(PS) = N_e
    assert np.abs(P @ S @ P - 2.0 * P).max() < 1e-8  # PSP = 2P (idempotency)
    assert np.all(np.diff(g["w7_acc_eps"]) >= 0)     # ascending orbital energies


def test_small_fixtures_regenerate_from_oracle():
    """The committed small fixtures are what the oracle computes now."""
    from oracle import refscf
    g = np.load(os.path.join(GOLD, "small.npz"))
    fb, fn = m.flatten(m.h2o_sto3g())
    # the fixtures were written at epsilon = 1e-12 (the energy of an earlier exit depends on the
    # exact iterate, not only on the fixed point)
    r = refscf.try_scf(fb, fn, 400, 1e-12, 0, accurate=False)
    assert r.total_energy == pytest.approx(g["w7_ref_energy"][0], abs=1e-10)
    packed, _ = refscf.eri_packed(fb, False)
    assert np.abs(refscf.dense_from_packed(fb.n, packed) - g["w7_ref_eri"]).max() == 0.0


def test_fd_gradient_invariances():
    """The oracle's finite-difference gradient: translation invariance (forces sum to zero) and the
    mirror symmetry of water (y -> -y swaps the hydrogens), for H2O/STO-3G and H2O/6-31G."""
    from oracle import refscf
    for atoms in (m.h2o_sto3g(), m.build(m.water_geometry(), "6-31G")):
        fb, fn = m.flatten(atoms)
        g = refscf.fd_gradient(fb, fn)
        assert np.abs(g.sum(axis=0)).max() < 2e-8
        assert abs(g[1, 1] + g[2, 1]) < 2e-8 and abs(g[1, 2] - g[2, 2]) < 2e-8
        assert np.abs(g[:, 0]).max() < 1e-9            # planar molecule in yz: no x force
        assert np.abs(g).max() > 1e-3                  # not at the minimum


def test_fd_gradient_h2_matches_energy_curve():
    """H2/STO-3G: the force on atom 2 along the bond is dE/dR of the energy curve E(R)."""
    from oracle import refscf
    b = m.BasisSet.load("STO-3G")
    R = 1.4
    fb, fn = m.flatten([b.get((0, 0, 0), 1), b.get((0, 0, R), 1)])
    g = refscf.fd_gradient(fb, fn)
    e = []
    for r in (R + 1e-3, R - 1e-3):
        fb2, fn2 = m.flatten([b.get((0, 0, 0), 1), b.get((0, 0, r), 1)])
        e.append(refscf.try_scf(fb2, fn2, 200, 1e-13, 0, accurate=True).total_energy)
    dEdR = (e[0] - e[1]) / 2e-3
    assert abs(g[1, 2] - dEdR) < 1e-6 and abs(g[0, 2] + dEdR) < 1e-6


# ------------------------------------------------------------------ pins of the J/K contraction,
# the reference screen, DIIS and the orbital evaluator

def _random_symmetric(n, seed):
    A = np.random.default_rng(seed).standard_normal((n, n))
    return np.ascontiguousarray(A + A.T)


@pytest.mark.parametrize("which", ["w7", "w13"])
def test_o64_jk_packed_equals_dense_einsum(oracle_lib, which):
    """o64_jk_packed (which wrote every J/K fixture) against numpy.einsum over the unpacked dense
    tensor: J_ij = sum_kl (ij|kl) P_kl, K_ij = sum_kl (ik|jl) P_kl, at a random symmetric P."""
    import oracle
    from oracle import refscf
    atoms = m.h2o_sto3g() if which == "w7" else m.build(m.water_geometry(), "6-31G")
    fb, _ = m.flatten(atoms)
    n = fb.n
    packed, _ = refscf.eri_packed(fb, True)
    D = refscf.dense_from_packed(n, packed)
    P = _random_symmetric(n, 5)
    J = np.zeros((n, n)); K = np.zeros((n, n))
    oracle_lib.o64_jk_packed(n, oracle.dptr(packed), oracle.dptr(P), oracle.dptr(J), oracle.dptr(K))
    Jw = np.einsum("ijkl,kl->ij", D, P)
    Kw = np.einsum("ikjl,kl->ij", D, P)
    scale = np.abs(D).max() * np.abs(P).sum()
    assert np.abs(J - Jw).max() < 1e-14 * scale
    assert np.abs(K - Kw).max() < 1e-14 * scale
    assert np.abs(J - J.T).max() == 0.0 and np.abs(K - K.T).max() == 0.0


@pytest.mark.parametrize("which", ["w13", "dimer6"])
def test_reference_screen_decision(oracle_lib, which):
    """ref_electron_tensor's 1e-6 rule (basis/src/electron_tensor.rs:89-101, 134-155) on
    H2O/6-31G (where nothing falls below it) and on two waters 6 A apart (where it drops
    quartets): every diagonal quartet (xy|xy) is kept; any other slot is zero exactly when
    sqrt((xy|xy)(zw|zw)) <= 1e-6 and otherwise equals the unscreened value bit for bit."""
    from oracle import refscf
    geom = m.water_geometry()
    if which == "dimer6":
        geom = geom + [(s, x + 6.0, y, z) for s, x, y, z in geom]
    fb, _ = m.flatten(m.build(geom, "6-31G"))
    n = fb.n
    scr, kept_count = refscf.eri_packed(fb, False, screen=True)
    full, all_count = refscf.eri_packed(fb, False, screen=False)
    Ds = refscf.dense_from_packed(n, scr)
    Df = refscf.dense_from_packed(n, full)
    d = np.einsum("ijij->ij", Df)
    est = np.sqrt(d[:, :, None, None] * d[None, None, :, :])
    i, j, k, l = np.indices((n,) * 4)
    diag = ((i == k) & (j == l)) | ((i == l) & (j == k))
    keep = diag | (est > 1e-6)
    assert np.array_equal(Ds[keep], Df[keep])
    assert np.all(Ds[~keep] == 0.0)
    if which == "dimer6":
        assert 0 < np.count_nonzero(~keep)        # the screen does drop quartets here
        assert kept_count < all_count


def test_refscf_diis_is_the_constrained_least_squares_minimum():
    """refscf.diis: the coefficients sum to one (focks all equal to one matrix give that matrix),
    and with focks = errors the extrapolated error is the minimum of |sum c_i e_i| over
    sum c_i = 1, found independently with numpy.linalg.lstsq on the null space of the constraint."""
    from oracle import refscf
    rng = np.random.default_rng(12)
    errs = [rng.standard_normal((6, 6)) * 10.0 ** (-0.5 * k) for k in range(5)]
    ones = [np.full((6, 6), 3.0) for _ in errs]
    assert np.abs(refscf.diis(errs, ones) - 3.0).max() < 1e-12
    got = refscf.diis(errs, errs)
    E = np.stack([e.ravel() for e in errs], axis=1)          # columns e_i
    m_ = E.shape[1]
    c0 = np.full(m_, 1.0 / m_)
    N = np.linalg.svd(np.ones((1, m_)))[2][1:].T            # basis of {t : sum t = 0}
    t, *_ = np.linalg.lstsq(E @ N, -(E @ c0), rcond=None)
    want = (E @ (c0 + N @ t)).reshape(6, 6)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_refmo_basis_values_integrate_to_the_overlap():
    """refmo.basis_values (ContractedGaussian::evaluate restated) integrated on a grid: the
    products of the diffuse valence functions of H2O/6-31G (exponents below 1, so a 0.4 Bohr
    trapezoid grid is exact to ~1e-12 for Gaussians) give the oracle's analytic overlap integrals,
    including the sign of p functions on one centre against s functions on another."""
    from oracle import refmo, refscf
    fb, fn = m.flatten(m.build(m.water_geometry(), "6-31G"))
    diffuse = [f for f in range(fb.n)
               if np.all(fb.exponents[fb.prim_offset[f]:fb.prim_offset[f + 1]] < 1.0)]
    assert any(fb.angular[f].sum() == 1 for f in diffuse) and len(diffuse) >= 5
    h, L = 0.4, 14.0
    ax = np.arange(-L, L + h / 2, h)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    V = refmo.basis_values(fb, pts, diffuse)
    Snum = (V @ V.T) * h ** 3
    S, _, _ = refscf.one_electron(fb, fn, True)
    Sref = S[np.ix_(diffuse, diffuse)]
    assert np.abs(Snum - Sref).max() < 1e-9
    assert np.abs(Sref - np.diag(np.diag(Sref))).max() > 0.1   # off-diagonal overlaps are tested


def test_refmo_truncation_and_voxel_order():
    """MolecularOrbitals::new keeps |c| > 0.05 strictly (molecular_orbitals.rs:79); voxel i of the
    renderer is (i mod N, (i / N) mod N, i / N^2) at (v / N - 0.5) BOX (volumetric/src/main.rs:111-116)."""
    from oracle import refmo
    C = np.array([[0.05, -0.0500001], [0.2, 0.0], [-0.05, 0.049], [0.051, -1.0]])
    idx, c = refmo.truncate(C, 0)
    assert list(idx) == [1, 3] and list(c) == [0.2, 0.051]
    idx, c = refmo.truncate(C, 1)
    assert list(idx) == [0, 3]
    N, box = 5, 20.0
    pts = refmo.volumetric_points(N, box)
    for i in range(N ** 3):
        x, y, z = i % N, (i // N) % N, i // N ** 2
        assert np.array_equal(pts[i], np.array([x, y, z], dtype=np.float64) / N * box - 0.5 * box) or \
            np.abs(pts[i] - (np.array([x, y, z]) / N - 0.5) * box).max() < 1e-14


def test_ediis_coefficients_satisfy_the_simplex_kkt_conditions():
    """oracle.refscf.ediis_coefficients on random convex instances (Q = -1/2 D with D_ij =
    |x_i - x_j|^2, conditionally negative definite, as tr((P_i - P_j)(F_i - F_j)) is): the result
    is on the simplex and certifies optimality -- the gradient g = e + Q c is equal on the support
    and no smaller elsewhere -- and no random simplex point does better."""
    from oracle import refscf
    rng = np.random.default_rng(4)
    for m_ in (2, 3, 5, 8):
        for _ in range(5):
            x = rng.standard_normal((m_, 3))
            D = np.sum((x[:, None] - x[None]) ** 2, axis=-1)
            Q = -0.5 * D
            e = rng.standard_normal(m_) * 2.0
            c = refscf.ediis_coefficients(e, Q)
            assert c is not None and np.all(c >= 0) and abs(c.sum() - 1.0) < 1e-12
            g = e + Q @ c
            sup = c > 1e-10
            assert np.ptp(g[sup]) < 1e-9 * max(1.0, np.abs(g).max())
            assert np.all(g[~sup] >= g[sup].min() - 1e-9)
            f = lambda v: e @ v + 0.5 * v @ Q @ v
            for _ in range(50):
                v = rng.dirichlet(np.ones(m_))
                assert f(v) >= f(c) - 1e-12
