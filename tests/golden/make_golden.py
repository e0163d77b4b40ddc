"""Generate the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py small     # H2, H2O (W7), H2O/6-31G: seconds
    python tests/golden/make_golden.py o108      # n-octane/6-31G: ~1 minute
    python tests/golden/make_golden.py e264      # n-eicosane/6-31G accurate SCF, P, J, K: ~5 minutes
    python tests/golden/make_golden.py forms     # water clusters and gradients: ~5 minutes

Fixtures are plain .npz arrays (no pickles). The oracle is the checker: refcpu.c
(the reference's arithmetic) and oracle64.c (accurate, unscreened), with the SCF
restated in oracle/refscf.py.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "fock-rs_amd"))

from fock_amd import molecule as m  # noqa: E402
from oracle import refscf  # noqa: E402


def rot_y(deg):
    t = np.radians(deg)
    return np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])


def random_density(n, nocc, seed):
    """A valid idempotent-like density: 2 C_occ C_occ^T from random orthonormal columns."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    Q, _ = np.linalg.qr(A)
    C = Q[:, :nocc] * 0.3
    return np.ascontiguousarray(2.0 * C @ C.T)


def scf_both(fb, fn, eps, tag, out, max_iters=400):
    for acc in (False, True):
        t = time.time()
        S, T, V = refscf.one_electron(fb, fn, acc)
        packed, computed = refscf.eri_packed(fb, acc, screen=True)
        r = refscf.try_scf(fb, fn, max_iters, eps, 0, accurate=acc, integrals=(S, T, V, packed))
        key = tag + ("_acc" if acc else "_ref")
        out[key + "_energy"] = np.array([r.total_energy])
        out[key + "_eelec"] = np.array([r.electronic_energy])
        out[key + "_eps"] = r.orbital_energies
        out[key + "_P"] = r.density_matrix
        out[key + "_iters"] = np.array([r.iterations])
        out[key + "_computed"] = np.array([computed])
        out[key + "_traj"] = np.array(r.energies)
        print("%s: E=%.12f iters=%d (%.1fs)" % (key, r.total_energy, r.iterations, time.time() - t), flush=True)
        yield acc, S, T, V, packed


def small():
    out = {}
    # W7: H2O/STO-3G, the reference-CPU oracle case
    fb, fn = m.flatten(m.h2o_sto3g())
    for acc, S, T, V, packed in scf_both(fb, fn, 1e-12, "w7", out):
        k = "w7_acc" if acc else "w7_ref"
        out[k + "_S"], out[k + "_T"], out[k + "_V"] = S, T, V
        out[k + "_eri"] = refscf.dense_from_packed(fb.n, packed)
        if not acc:
            unscreened, _ = refscf.eri_packed(fb, False, screen=False)
            out["w7_ref_eri_unscreened"] = refscf.dense_from_packed(fb.n, unscreened)
    # the same molecule rotated 90 degrees about y: exposes the x-branch defect
    fb, fn = m.flatten(m.build(m.rotate_geometry(m.water_geometry(), rot_y(90.0)), "STO-3G"))
    for acc, S, T, V, packed in scf_both(fb, fn, 1e-12, "w7rot", out):
        k = "w7rot_acc" if acc else "w7rot_ref"
        out[k + "_eri"] = refscf.dense_from_packed(fb.n, packed)
    # H2O/6-31G (n = 13) and H2/STO-3G at 1.4 bohr (Szabo-Ostlund)
    fb, fn = m.flatten(m.build(m.water_geometry(), "6-31G"))
    for acc, S, T, V, packed in scf_both(fb, fn, 1e-12, "w13", out):
        k = "w13_acc" if acc else "w13_ref"
        out[k + "_eri"] = refscf.dense_from_packed(fb.n, packed)
    b = m.BasisSet.load("STO-3G")
    fb, fn = m.flatten([b.get((0, 0, 0), 1), b.get((0, 0, 1.4), 1)])
    for acc, S, T, V, packed in scf_both(fb, fn, 1e-12, "h2", out):
        pass
    np.savez_compressed(os.path.join(HERE, "small.npz"), **out)


def o108():
    out = {}
    fb, fn = m.flatten(m.octane_631g())
    n = fb.n
    P = random_density(n, 33, seed=108)
    out["P_random"] = P
    for acc, S, T, V, packed in scf_both(fb, fn, 1e-11, "o108", out):
        J = np.zeros((n, n)); K = np.zeros((n, n))
        from oracle import dptr, lib
        lib().o64_jk_packed(n, dptr(packed), dptr(P), dptr(J), dptr(K))
        k = "o108_acc" if acc else "o108_ref"
        out[k + "_J"], out[k + "_K"] = J, K
        out[k + "_S"], out[k + "_T"], out[k + "_V"] = S, T, V
    np.savez_compressed(os.path.join(HERE, "o108.npz"), **out)


def e264():
    out = {}
    fb, fn = m.flatten(m.eicosane_631g())
    t = time.time()
    S, T, V = refscf.one_electron(fb, fn, True)
    packed, _ = refscf.eri_packed(fb, True)
    print("e264 accurate integrals %.1fs" % (time.time() - t), flush=True)
    # the reference's schedule (GWH, DIIS from 5 in a window of 16) does not converge for C20H42;
    # the accurate form uses DIIS from 2 in a window of 10 with a 0.5 Eh level shift until
    # max|FDS - SDF| < 0.05 (the same options the GPU driver takes)
    r = refscf.try_scf(fb, fn, 200, 1e-10, 0, accurate=True, integrals=(S, T, V, packed), diis_start=2,
                       diis_window=10, level_shift=0.5, level_shift_until=0.05)
    out["e264_acc_energy"] = np.array([r.total_energy])
    out["e264_acc_eelec"] = np.array([r.electronic_energy])
    out["e264_acc_eps"] = r.orbital_energies
    out["e264_acc_iters"] = np.array([r.iterations])
    out["e264_acc_errors"] = np.array(r.errors)
    print("e264 accurate E=%.12f iters=%d (%.1fs)" % (r.total_energy, r.iterations, time.time() - t), flush=True)
    # the converged density and the unscreened J, K it gives (upper triangles, row-major order of
    # numpy.triu_indices): the reference point of the precision form's max |dF_ij|
    n = fb.n
    P = np.ascontiguousarray(r.density_matrix)
    J = np.zeros((n, n)); K = np.zeros((n, n))
    from oracle import dptr, lib
    lib().o64_jk_packed(n, dptr(packed), dptr(P), dptr(J), dptr(K))
    iu = np.triu_indices(n)
    out["e264_acc_P_triu"] = P[iu]
    out["e264_acc_J_triu"] = J[iu]
    out["e264_acc_K_triu"] = K[iu]
    np.savez_compressed(os.path.join(HERE, "e264.npz"), **out)


ROBUST = dict(diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)
CLUSTER_SEED = 20261203


def forms():
    """Oracle references of the batching and differentiation forms: energies of seeded (H2O)_k,
    k = 2, 3, 4 (6-31G); central finite-difference gradients (1e-4 Bohr) of the accurate,
    unscreened energy for the seeded (H2O)4 of the gradient tests and for HF/6-311G (the largest
    exponents of the shipped s/p sets: F 1.1e4)."""
    out = {}
    rng = np.random.default_rng(CLUSTER_SEED)
    for k in (2, 3, 4):
        t = time.time()
        g = m.water_cluster_geometry(k, rng)
        fb, fn = m.flatten(m.build(g, "6-31G"))
        r = refscf.try_scf(fb, fn, 300, 1e-10, 0, accurate=True, **ROBUST)
        out["w%d_sym" % k] = np.array([a[0] for a in g])
        out["w%d_xyz" % k] = np.array([a[1:] for a in g], dtype=np.float64)
        out["w%d_energy" % k] = np.array([r.total_energy])
        out["w%d_eps" % k] = r.orbital_energies
        print("(H2O)%d E=%.12f iters=%d (%.1fs)" % (k, r.total_energy, r.iterations, time.time() - t), flush=True)
    cases = {"grad_w4": (m.water_cluster_geometry(4, np.random.default_rng(4)), "6-31G"),
             "grad_hf": ([("F", 0.0, 0.0, 0.0), ("H", 0.0, 0.0, 0.917)], "6-311G")}
    for key, (g, basis) in cases.items():
        t = time.time()
        fb, fn = m.flatten(m.build(g, basis))
        out[key + "_sym"] = np.array([a[0] for a in g])
        out[key + "_xyz"] = np.array([a[1:] for a in g], dtype=np.float64)
        out[key + "_fd"] = refscf.fd_gradient(fb, fn, **ROBUST)
        print("%s fd gradient (%.1fs)" % (key, time.time() - t), flush=True)
    np.savez_compressed(os.path.join(HERE, "forms.npz"), **out)


if __name__ == "__main__":
    for stage in sys.argv[1:] or ["small"]:
        {"small": small, "o108": o108, "e264": e264, "forms": forms}[stage]()
