#!/usr/bin/env python3
"""Headline benchmark: direct-SCF Fock builds for n-eicosane C20H42/6-31G on one MI355X.

Workload (BASELINE metric "Fock builds/s + SCF wall s, C20H42/6-31G, mixed vs FP64 GPU at
|dE| < 1e-6 Eh"): the idealised all-anti C20H42 zigzag (C-C 1.54 A, C-H 1.09 A, tetrahedral
angles) in 6-31G, n = 264 basis functions, accurate integrals with Schwarz screening 1e-12.

A step is one complete direct Fock build G(P) = J - K/2 over every screened shell quartet, with
the density, J and K resident in HBM. `value` is the mixed-precision form (FP32 quartet
integrals, FP64 accumulation); the FP64 path is timed the same way beside it. Both full SCFs
(FP64 and mixed with incremental builds + final FP64 build) are run once and reported with
their wall time and |dE| against the FP64 unscreened oracle energy.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "fock-rs_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Fock builds/s + SCF wall s, C20H42/6-31G, mixed vs FP64 GPU at |ΔE|<1e-6 Eh"
CLASS_NAMES = ["(ss|ss)", "(ps|ss)", "(ps|ps)", "(pp|ss)", "(pp|ps)", "(pp|pp)"]
PEAK_FP64 = 78.6   # TFLOP/s, MI355X FP64 vector (= FP64 matrix)
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector (published, packed)
# The reference's own SCF schedule (GWH guess, DIIS from 5 stored matrices in a window of 16) does
# not converge for C20H42 (max|FDS-SDF| stays ~1-2); both precision forms use DIIS from 2 in a
# window of 10 with a 0.5 Eh level shift until the error drops below 0.05.
SCF_SCHEDULE = dict(diis_start=2, diis_window=10, level_shift=0.5, level_shift_until=0.05)
SCF_EPS = 1e-8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_builds(ctx, torch, dP, dJ, dK, steps, warmup):
    """W untimed builds, then exactly K builds bracketed by synchronisation; returns ms/step."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        ctx.jk_build_device(dP.data_ptr(), dJ.data_ptr(), dK.data_ptr())
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        ctx.jk_build_device(dP.data_ptr(), dJ.data_ptr(), dK.data_ptr())
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def per_class_ms(ctx, torch, dP, dJ, dK, steps):
    """Average duration of each class kernel over `steps` builds (HIP events on the launch stream)."""
    import numpy as np
    ctx.jk_set_timing(True)
    acc = np.zeros(6)
    for _ in range(steps):
        ctx.jk_build_device(dP.data_ptr(), dJ.data_ptr(), dK.data_ptr())
        acc += ctx.jk_timings()
    ctx.jk_set_timing(False)
    torch.cuda.synchronize()
    return acc / steps


def load_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel_key)
    except Exception:
        return None


def cpu_baseline(fb, threads):
    """The oracle's restatement of the reference CPU path, on a bounded sample:
    - one-time ERI tensor: a seeded uniform sample of unique basis-function quartets through the
      reference's recursive integral code with its 1e-6 pair screen, extrapolated to all
      n(n+1)/2 (n(n+1)/2 + 1)/2 unique quartets;
    - per-iteration Fock build: the reference's single-threaded n^4/2 contraction of the stored
      electron_terms array, timed at n = 108 and scaled by (264/108)^4 (the n = 264 array
      needs 39 GB)."""
    import numpy as np
    import oracle
    from oracle import dptr
    L = oracle.lib()
    L.ref_set_num_threads(threads)
    L.ref_set_policy(0, 0)
    n = fb.n
    bv = oracle.BasisView(fb)
    t0 = time.time()
    diag = np.zeros((n, n))
    L.ref_eri_diagonal(bv.ref, dptr(diag))
    t_diag = time.time() - t0
    cs = np.zeros(1)
    # size the sample to ~10 s of CPU work
    t0 = time.time()
    probe = 20000
    L.ref_eri_sample(bv.ref, dptr(diag), probe, 12345, dptr(cs))
    rate = probe / max(time.time() - t0, 1e-6)
    count = int(min(max(rate * 10.0, 1e5), 1e9))
    t0 = time.time()
    L.ref_eri_sample(bv.ref, dptr(diag), count, 20261202, dptr(cs))
    t_sample = time.time() - t0
    npair = n * (n + 1) // 2
    unique = npair * (npair + 1) // 2
    eri_tensor_s = t_sample / count * unique + t_diag
    # per-iteration contraction at n = 108 on a synthetic stored array (same loop, same bytes)
    m = 108
    terms = np.random.default_rng(1).standard_normal(m ** 4)
    P = np.random.default_rng(2).standard_normal((m, m))
    P = np.ascontiguousarray(P + P.T)
    G = np.zeros((m, m))
    t0 = time.time()
    L.ref_g_build(m, dptr(terms), dptr(P), dptr(G))
    t_g108 = time.time() - t0
    del terms
    t_g = t_g108 * (n / m) ** 4
    return {
        "value": 1.0 / t_g, "unit": "Fock builds/s", "cores": threads, "kind": "port",
        "sample": ("reference-restatement G contraction (single thread, as the reference) timed at n=108 "
                   "and scaled (n/108)^4 to n=%d: %.3f s/build; one-time ERI tensor: %d of %d unique "
                   "quartets sampled on %d threads (%.1f s) -> %.0f s extrapolated" %
                   (n, t_g, count, unique, threads, t_sample, eri_tensor_s)),
        "g_build_s": t_g, "eri_tensor_s_extrapolated": eri_tensor_s,
        "scf_wall_s_extrapolated_20_iters": eri_tensor_s + 20 * t_g,
    }


def batch_form(ctx, abi, m, np):
    """64 water clusters (H2O)_k, k ~ U{1..8} (seed 20261202), 6-31G, accurate integrals: the
    lockstep batched SCF against the same 64 SCFs run one after the other."""
    mols = [m.flatten(m.build(g, "6-31G")) for g in m.water_cluster_batch(64, seed=20261202)]
    o = abi.default_opts(max_iters=200, epsilon=SCF_EPS, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12, **SCF_SCHEDULE)
    o.autotune = 1
    ctx.scf_batch_run(mols, o)  # warm-up (code objects, allocations) and tile/schedule tuning
    o.autotune = 0
    t0 = time.time()
    st, statuses, res = ctx.scf_batch_run(mols, o)
    t_batch = time.time() - t0
    t0 = time.time()
    single = [ctx.scf_run(fb, fn, o)[1] for fb, fn in mols]
    t_seq = time.time() - t0
    return {
        "workload": "64 (H2O)_k clusters, k in 1..8, 6-31G (n = 13..104), accurate, Schwarz 1e-12",
        "status": abi.STATUS_NAMES[st], "batch_wall_s": t_batch, "sequential_wall_s": t_seq,
        "speedup": t_seq / t_batch, "scf_per_s": 64 / t_batch,
        "iterations_max": int(max(r["iterations"] for r in res)),
        "max_abs_dE_batch_vs_single": float(max(abs(a["total_energy"] - b["total_energy"]) for a, b in zip(res, single))),
        "n_basis_range": [min(fb.n for fb, _ in mols), max(fb.n for fb, _ in mols)],
        "tiles": ctx.tiles(abi.TILES_BATCH)[0],
        "schedule": ["cost", "xcd"][ctx.tiles(abi.TILES_BATCH)[1]],
    }


def mo_grid_form(ctx, abi, m, np, torch, fb, C, nocc, cpu=True):
    """The renderers' loop (volumetric/src/main.rs:100-125): orbitals of the converged C20H42/6-31G
    SCF on the renderer's grid (200^3 voxels over a 20 Bohr box), FP64 arithmetic, float32 texture
    written to HBM (fock_mo_grid_device). Timed: the HOMO alone and all occupied orbitals in one
    call. Algorithmic flop per voxel, counted from the truncated orbital (|c| > 0.05): per shell
    with kept functions 8 (|d|^2), per primitive 2 (alpha d^2, exp counted as one), per kept
    function 2K + L + 2 (radial sum, polynomial, accumulation)."""
    N, box = 200, 20.0
    step = box / N
    origin, steps, dims = (-0.5 * box,) * 3, (step,) * 3, (N,) * 3
    nvox = N ** 3
    stream = torch.cuda.current_stream()

    def flop_per_voxel(k):
        shells, f = {}, 0.0
        for i in range(fb.n):
            if abs(C[i, k]) > 0.05:
                key = tuple(fb.centers[i]) + (fb.prim_offset[i + 1] - fb.prim_offset[i], float(fb.exponents[fb.prim_offset[i]]))
                shells[key] = fb.prim_offset[i + 1] - fb.prim_offset[i]
                f += 2 * shells[key] + int(fb.angular[i].sum()) + 2
        return f + sum(8 + 2 * K for K in shells.values())

    out = {}
    for name, orbs in (("homo", [nocc - 1]), ("occupied", list(range(nocc)))):
        buf = torch.empty(len(orbs) * nvox, dtype=torch.float32, device="cuda:0")
        ctx.mo_grid_device(fb, C, orbs, origin, steps, dims, buf.data_ptr())
        torch.cuda.synchronize()
        reps = 5 if name == "homo" else 2
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            ctx.mo_grid_device(fb, C, orbs, origin, steps, dims, buf.data_ptr())
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        flop = sum(flop_per_voxel(k) for k in orbs) * nvox
        out[name] = {"orbitals": len(orbs), "ms": ms, "voxels_per_s": len(orbs) * nvox / (ms * 1e-3),
                     "achieved_tflops": flop / (ms * 1e-3) / 1e12, "frac_fp64": flop / (ms * 1e-3) / 1e12 / PEAK_FP64}
        del buf
    if cpu:  # the oracle's numpy evaluation (one thread) on 200k voxels of the HOMO
        from oracle import refmo
        pts = refmo.volumetric_points(N, box)[::40][:200000]
        t0 = time.time()
        refmo.evaluate(fb, C, nocc - 1, pts)
        t = time.time() - t0
        out["cpu_baseline"] = {"voxels_per_s": len(pts) / t, "cores": 1, "kind": "port",
                               "sample": "%d voxels of the HOMO through oracle/refmo.py (numpy)" % len(pts)}
        out["speedup_homo_vs_cpu"] = out["homo"]["voxels_per_s"] / out["cpu_baseline"]["voxels_per_s"]
    out["workload"] = "C20H42/6-31G orbitals on the volumetric renderer's grid: 200^3 voxels, 20 Bohr box, float32"
    return out


def gradient_form(ctx, abi, m, np):
    """Analytic RHF gradient of n-octane/6-31G at its converged FP64 SCF."""
    fb, fn = m.flatten(m.octane_631g())
    o = abi.default_opts(max_iters=200, epsilon=1e-10, integral_policy=abi.POLICY_ACCURATE,
                         screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12, **SCF_SCHEDULE)
    st, r = ctx.scf_run(fb, fn, o)
    ctx.rhf_gradient(fb, fn, o, r)  # warm-up
    t0 = time.time()
    g = ctx.rhf_gradient(fb, fn, o, r)
    t = time.time() - t0
    return {"workload": "C8H18/6-31G analytic nuclear gradient (FP64)", "gradient_ms": t * 1e3,
            "scf_wall_ms": r["scf_wall_ms"], "max_abs_force": float(np.abs(g).max()),
            "net_force": float(np.abs(g.sum(axis=0)).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scf", action="store_true", help="skip both SCF runs (builds use a synthetic density)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batching, gradient and orbital-grid forms (builds and the C20H42 SCFs only)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile (8 tasks per workgroup)")
    ap.add_argument("--tiles", default="", metavar="JSON",
                    help='set the tiles instead of tuning, e.g. the "tiles" of an earlier run\'s config: '
                         '{"fp64": {"tile": [...]}, "mixed": {"tile": [...]}} (profiling runs)')
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the Fock build is measured on a single MI355X "
              "(multi-GPU sharding of the quartet task list is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.calibrate_copy:
        # the byte count a plain copy kernel moves, for scaling FETCH_SIZE / WRITE_SIZE
        n_bytes = args.calibrate_copy << 20
        x = torch.ones(n_bytes // 8, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n_bytes, "launches": 3}), flush=True)
        return

    from fock_amd import abi, molecule as m

    torch.cuda.init()
    atoms = m.eicosane_631g()
    fb, fn = m.flatten(atoms)
    n = fb.n
    ctx = abi.Context(0)
    # a dedicated (default-priority) torch stream: the library launches on it and the torch
    # events of the timed region are recorded on it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream.cuda_stream)
    result = {}

    # tile autotuning (setup, once per form, kept in the context for every later prepare): a full
    # build's work does not depend on the density values, so a synthetic density serves
    tiles = {}
    if args.tiles:
        tiles = json.loads(args.tiles)
        for name, form in (("fp64", abi.TILES_FP64), ("mixed", abi.TILES_MIXED)):
            ctx.set_tiles(form, tiles[name]["tile"])
    elif not args.no_autotune:
        for name, prec in (("fp64", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED)):
            ctx.jk_prepare(fb, abi.POLICY_ACCURATE, abi.SCREEN_SCHWARZ, 1e-12, prec)
            rep = ctx.jk_autotune(np.eye(n) * 0.1)
            tiles[name] = {"tile": rep["tile"], "tune_ms": round(rep["tune_ms"], 1)}
            log("tiles %s: %s (%.0f ms)" % (name, rep["tile"], rep["tune_ms"]))

    # reference density: the FP64 SCF (accurate, Schwarz 1e-12) -- also the FP64 SCF wall time
    gold_path = os.path.join(REPO, "tests", "golden", "e264.npz")
    e_oracle = float(np.load(gold_path)["e264_acc_energy"][0]) if os.path.exists(gold_path) else None
    scf = {}
    P_ref = None
    C_fp64 = None
    if not args.no_scf:  # untimed warm-up of both forms (code objects, solver workspaces)
        for prec, incr in ((abi.PRECISION_FP64, 0), (abi.PRECISION_MIXED, 1)):
            ctx.scf_run(fb, fn, abi.default_opts(max_iters=2, epsilon=SCF_EPS, integral_policy=abi.POLICY_ACCURATE,
                                                 screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12,
                                                 precision=prec, incremental=incr, **SCF_SCHEDULE))
    for name, prec, incr in (("fp64", abi.PRECISION_FP64, 0), ("mixed", abi.PRECISION_MIXED, 1)):
        if args.no_scf:
            continue
        opts = abi.default_opts(max_iters=100, epsilon=SCF_EPS, integral_policy=abi.POLICY_ACCURATE,
                                screen=abi.SCREEN_SCHWARZ, schwarz_threshold=1e-12, precision=prec,
                                incremental=incr, **SCF_SCHEDULE)
        log("SCF %s ..." % name)
        try:
            st, r = ctx.scf_run(fb, fn, opts)
        except abi.FockError as e:
            log("SCF %s failed: %s" % (name, e))
            scf[name] = {"error": str(e)}
            continue
        scf[name] = {
            "status": abi.STATUS_NAMES[st], "total_energy": r["total_energy"], "iterations": r["iterations"],
            "scf_wall_s": r["scf_wall_ms"] / 1e3,
            "fock_build_ms_mean": float(np.mean([h[2] for h in r["history"]])) if r["history"] else None,
            "fock_build_ms_total": float(np.sum([h[2] for h in r["history"]])) if r["history"] else None,
            # per iteration: build ms and the shell quartets it evaluated
            "builds": [[round(h[2], 3), int(h[3])] for h in r["history"]],
        }
        # evaluated quartets over the full Schwarz list of the FP64 build, per iteration
        full = max(int(h[3]) for h in r["history"]) if r["history"] else 0
        if full:
            scf[name]["kept_fraction"] = [round(int(h[3]) / full, 3) for h in r["history"]]
        if e_oracle is not None:
            scf[name]["abs_dE_vs_fp64_oracle"] = abs(r["total_energy"] - e_oracle)
        if name == "fp64":
            P_ref = r["density"]
            C_fp64 = r["coefficients"]
            nocc = r["n_electrons"] // 2
        log("SCF %s: %s" % (name, {k: v for k, v in scf[name].items() if k != "builds"}))
    if P_ref is None:  # synthetic density (same screened quartet set: the build does not screen on P)
        P_ref = np.eye(n) * 0.1

    # accuracy of the full builds at the FP64 oracle's converged density P*: max |dF_ij| =
    # max |d(J - K/2)_ij| against the oracle's unscreened J, K (tests/golden/e264.npz)
    dF = {}
    if os.path.exists(gold_path):
        g = np.load(gold_path)
        if "e264_acc_P_triu" in g.files:
            iu = np.triu_indices(n)

            def full(key):
                A = np.zeros((n, n))
                A[iu] = g[key]
                return A + np.triu(A, 1).T

            Pstar, Jstar, Kstar = full("e264_acc_P_triu"), full("e264_acc_J_triu"), full("e264_acc_K_triu")
            for name, prec in (("fp64", abi.PRECISION_FP64), ("fp32", abi.PRECISION_MIXED)):
                ctx.jk_prepare(fb, abi.POLICY_ACCURATE, abi.SCREEN_SCHWARZ, 1e-12, prec)
                J, K = ctx.jk_build(Pstar)
                dF["max_abs_dF_" + name] = float(np.abs((J - 0.5 * K) - (Jstar - 0.5 * Kstar)).max())
            P_ref = Pstar
            log("full-build accuracy at the oracle density: %s" % dF)
    dP = torch.from_numpy(np.ascontiguousarray(P_ref)).to("cuda:0")
    dJ = torch.empty((n, n), dtype=torch.float64, device="cuda:0")
    dK = torch.empty_like(dJ)
    per = {}
    for name, prec in (("fp64", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED)):
        ctx.jk_prepare(fb, abi.POLICY_ACCURATE, abi.SCREEN_SCHWARZ, 1e-12, prec)
        stats = ctx.jk_stats()
        ms = timed_builds(ctx, torch, dP, dJ, dK, args.steps, args.warmup)
        cls_ms = per_class_ms(ctx, torch, dP, dJ, dK, max(3, min(args.steps, 10)))
        per[name] = {"ms": ms, "stats": stats, "class_ms": cls_ms}
        log("%s: %.3f ms/build, class ms %s" % (name, ms, np.round(cls_ms, 3).tolist()))

    mixed, fp64 = per["mixed"], per["fp64"]
    dom = int(np.argmax(mixed["class_ms"]))
    flop = mixed["stats"]["algorithmic_flop"][dom]
    achieved = flop / (mixed["class_ms"][dom] * 1e-3) / 1e12
    total_flop = sum(mixed["stats"]["algorithmic_flop"])
    roofline = {
        "bound": "valu", "kernel": "eri_block_kernel %s fp32" % CLASS_NAMES[dom],
        "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
        "traffic": load_traffic("mixed_" + CLASS_NAMES[dom]),
        # the classes of a build run concurrently: the whole build's algorithmic rate over the
        # same peak (the dominant kernel's own figure above is from a serial launch)
        "whole_build_tflops": total_flop / (mixed["ms"] * 1e-3) / 1e12,
        "whole_build_frac": total_flop / (mixed["ms"] * 1e-3) / 1e12 / PEAK_FP32,
    }
    fp64_dom = int(np.argmax(fp64["class_ms"]))
    fp64_ach = fp64["stats"]["algorithmic_flop"][fp64_dom] / (fp64["class_ms"][fp64_dom] * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": 1e3 / mixed["ms"],
        "unit": "Fock builds/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32 quartets + f64 accumulation",
        "data": "synthetic: generated all-anti C20H42 geometry, 6-31G basis (BSE data)",
        "config": {"workload": "C20H42/6-31G direct Fock build, n=264, accurate integrals, Schwarz 1e-12",
                   "n_basis": n, "shells": mixed["stats"]["n_shells"],
                   "shell_quartets": int(sum(mixed["stats"]["shell_quartets"])),
                   "primitive_quartets": int(sum(mixed["stats"]["primitive_quartets"])),
                   "tiles": tiles or "default 8"},
        "fp64": {"builds_per_s": 1e3 / fp64["ms"], "ms_per_step": fp64["ms"],
                 "dominant_kernel": CLASS_NAMES[fp64_dom], "dominant_achieved_tflops": fp64_ach,
                 "dominant_frac": fp64_ach / PEAK_FP64},
        "speedup_mixed_vs_fp64": fp64["ms"] / mixed["ms"],
        # per class and form: kept shell quartets, primitive quartets and integrals (the FP32
        # kernels drop shell quartets below the FP32 cut), and the algorithmic flop behind roofline
        "class_counts": {name: {CLASS_NAMES[c]: {k: int(per[name]["stats"][k][c]) for k in
                                                 ("shell_quartets", "primitive_quartets", "integrals", "algorithmic_flop")}
                                for c in range(len(CLASS_NAMES))} for name in ("mixed", "fp64")},
        "class_ms": {"mixed": dict(zip(CLASS_NAMES, map(float, mixed["class_ms"]))),
                     "fp64": dict(zip(CLASS_NAMES, map(float, fp64["class_ms"])))},
        "scf": scf, "e_oracle": e_oracle,
        "roofline": roofline,
    }
    result.update(dF)
    if "fp64" in scf and "mixed" in scf and scf["mixed"].get("fock_build_ms_mean"):
        # the SCF-level ratio next to the full-build ratio: mean build of the FP64 SCF over the mean
        # build of the mixed SCF (FP32 bulk + incremental builds + FP64 cleanup builds)
        result["scf_build_ratio_fp64_over_mixed"] = scf["fp64"]["fock_build_ms_mean"] / scf["mixed"]["fock_build_ms_mean"]
        result["scf_mixed_mean_build_over_full_fp32"] = scf["mixed"]["fock_build_ms_mean"] / mixed["ms"]
    if not args.no_scf and not args.no_forms:
        result["batch_form"] = batch_form(ctx, abi, m, np)
        log("batch form: %s" % result["batch_form"])
        result["gradient_form"] = gradient_form(ctx, abi, m, np)
        log("gradient form: %s" % result["gradient_form"])
        if C_fp64 is not None:
            result["mo_grid_form"] = mo_grid_form(ctx, abi, m, np, torch, fb, C_fp64, nocc,
                                                  cpu=not args.no_cpu_baseline)
            log("mo grid form: %s" % result["mo_grid_form"])
    if not args.no_cpu_baseline:
        log("CPU baseline ...")
        cb = cpu_baseline(fb, threads=min(16, os.cpu_count() or 1))
        result["cpu_baseline"] = cb
        result["speedup_vs_cpu_baseline"] = result["value"] / cb["value"]
    else:
        result["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
