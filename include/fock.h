/*
 * fock.h -- C ABI of the MI355X restricted Hartree-Fock engine (libfock.so).
 *
 * Plain C17 types only: pointers, sizes, enums. No HIP or torch types appear in
 * any signature; a HIP stream is passed as an opaque `void *`.
 *
 * Each entry point names the fock-rs interface it replaces (paths relative to
 * the fock-rs repository). A Rust `fock-sys` binding of these declarations is
 * given in INTEGRATION.md.
 *
 * Threading: a context is owned by one host thread at a time and holds no global
 * mutable state; distinct contexts may run concurrently (unlike the reference's
 * `static mut BUFFER`, scf/src/utils.rs:7,63).
 */
#ifndef FOCK_H
#define FOCK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum fock_status {
    FOCK_OK = 0,
    /* Iteration limit reached; outputs are filled. The reference returns
     * Some(result) in this case (scf/src/lib.rs:224). */
    FOCK_NOT_CONVERGED = 1,
    /* max|FDS - SDF| became NaN/inf/subnormal; the reference returns None
     * (scf/src/lib.rs:248-249). Also returned, before any eigensolve, when the overlap matrix is
     * singular (smallest eigenvalue <= n * DBL_EPSILON * largest, e.g. coincident atoms), where
     * the reference's 1/sqrt of (U^T S U)_ii (scf/src/lib.rs:139-146) ends in None. Only
     * n_electrons and n_basis of the outputs are meaningful. */
    FOCK_DIVERGED = 2,
    /* Bad arguments, including non-finite centres, exponents, coefficients, nuclear positions or
     * charges (rejected before anything is launched). */
    FOCK_INVALID_ARG = 3,
    /* Boys order > 10 under the compatibility policy; the reference panics
     * (boys/src/lib.rs:59-62). */
    FOCK_BOYS_ORDER = 4,
    FOCK_HIP_ERROR = 5,
    FOCK_OOM = 6,
    /* Angular momentum beyond the compiled kernels (s and p shells). */
    FOCK_UNSUPPORTED = 7
} fock_status;

/* Integral arithmetic. REF_COMPAT reproduces the reference's own numbers: the
 * Boys lookup table with linear interpolation (boys/src/lib.rs:58-81) and the
 * Hermite-Coulomb x-branch exactly as written (basis/src/utils.rs:98-105).
 * ACCURATE uses a Taylor-table Boys function and the textbook recursion. */
typedef enum fock_integral_policy {
    FOCK_POLICY_REF_COMPAT = 0,
    FOCK_POLICY_ACCURATE = 1
} fock_integral_policy;

/* Integral screening. REF_BF_PAIR: an integral (xy|zw) is kept iff
 * sqrt((xy|xy)(zw|zw)) > 1e-6 per basis-function pair, else it is zero
 * (basis/src/electron_tensor.rs:134-155). SCHWARZ: shell-quartet Schwarz bound
 * Q_ab Q_cd > threshold. NONE: every quartet. */
typedef enum fock_screen {
    FOCK_SCREEN_REF_BF_PAIR = 0,
    FOCK_SCREEN_SCHWARZ = 1,
    FOCK_SCREEN_NONE = 2
} fock_screen;

/* Arithmetic of the quartet integrals. MIXED: quartet integrals and their
 * contraction in FP32, Fock accumulation, Boys for large arguments, density,
 * transforms and diagonalisation in FP64. */
typedef enum fock_precision {
    FOCK_PRECISION_FP64 = 0,
    FOCK_PRECISION_MIXED = 1
} fock_precision;

typedef struct fock_ctx fock_ctx;

/* Flat view of the reference's Vec<ContractedGaussian> (basis/src/contracted_gaussian.rs:9-35),
 * borrowed for the duration of a call. Function f has centre centers[3f..3f+2] (Bohr),
 * angular momentum angular[3f..3f+2] and primitives prim_offset[f] .. prim_offset[f+1]-1 with
 * exponents[] and coefficients[] as GaussianPrimitive::new stores them (already multiplied by
 * the primitive norm, basis/src/primitives.rs:41-48). Functions must come in complete shells in
 * the reference's component order (basis_set/src/lib.rs:87-99: p as z, y, x). */
typedef struct fock_basis_desc {
    int32_t n_functions;
    const double *centers;
    const int32_t *angular;
    const int32_t *prim_offset;
    const double *exponents;
    const double *coefficients;
} fock_basis_desc;

/* Point charges (basis/src/lib.rs:7-11 PointCharge; basis_set/src/atom.rs:32-43). */
typedef struct fock_nuclei_desc {
    int32_t n_atoms;
    const double *positions; /* [n_atoms][3], Bohr */
    const double *charges;   /* [n_atoms]; the electron count is sum(charges) - molecule_charge */
} fock_nuclei_desc;

typedef struct fock_scf_opts {
    int32_t max_iters;        /* try_scf max_iters: iterations 0..=max_iters */
    double epsilon;           /* try_scf epsilon on max|FDS - SDF| */
    int32_t molecule_charge;  /* try_scf molecule_charge */
    int32_t integral_policy;  /* fock_integral_policy */
    int32_t screen;           /* fock_screen */
    double schwarz_threshold; /* for FOCK_SCREEN_SCHWARZ */
    int32_t precision;        /* fock_precision */
    /* MIXED only: build G(P_k) = G(P_{k-1}) + G(P_k - P_{k-1}) with density-weighted screening,
     * then one final full FP64 build at the converged density. */
    int32_t incremental;
    /* Optional starting density (n x n, symmetric) replacing the GWH guess of
     * scf/src/lib.rs:148-162; NULL for the reference's guess. */
    const double *initial_density;
    /* DIIS schedule: extrapolate once diis_start (error, Fock) pairs are stored, keep the last
     * diis_window (<= 16). The reference uses 5 and 16 (scf/src/lib.rs:197-207). */
    int32_t diis_start;
    int32_t diis_window;
    /* Level shift sigma added to the extrapolated Fock matrix as sigma (S - S P S / 2) while
     * max|FDS - SDF| > level_shift_until (0: none, as the reference). */
    double level_shift;
    double level_shift_until;
    /* Nonzero: fock_scf_run / fock_scf_batch_run run the tile autotuner (fock_jk_autotune) at the
     * first iteration's density before iterating; the chosen tiles stay in the context for later
     * calls. 0 (default): the context's current tiles. */
    int32_t autotune;
    /* Nonzero: while max|FDS - SDF| > ediis_until, fock_scf_run extrapolates with EDIIS (energy
     * DIIS, Kudin, Scuseria and Cances, J. Chem. Phys. 116, 8255 (2002)) instead of DIIS: the
     * convex combination sum c_i F_i (c_i >= 0, sum c_i = 1) of the stored Fock matrices that
     * minimises the interpolated energy sum c_i E_i - 1/4 sum c_i c_j tr((P_i - P_j)(F_i - F_j)),
     * E_i = 1/2 tr(P_i (H + F_i)); DIIS below ediis_until. Single-molecule SCF only. */
    int32_t ediis;
    int32_t reserved[2];
    double ediis_until;
    /* Density damping: P <- damping P_new + (1 - damping) P_old while max|FDS - SDF| >
     * damping_until, the mixing of scf/src/lib.rs:218-219 whose factor the reference fixes at 1.
     * Values <= 0 or >= 1 mean no damping (the default). Single-molecule SCF only. */
    double damping;
    double damping_until;
} fock_scf_opts;

/* Per-iteration record (optional output). */
typedef struct fock_scf_iter {
    double total_energy;
    double error;           /* max|FDS - SDF| */
    double fock_build_ms;   /* GPU time of this iteration's Fock build */
    int64_t quartets;       /* shell quartets evaluated in this build */
} fock_scf_iter;

/* Outputs of try_scf (scf/src/lib.rs:27-37 HartreeFockResult). Matrices are caller-allocated,
 * column-major n x n (nalgebra DMatrix layout), rows in the basis-function order of the input;
 * any pointer may be NULL to skip that output. */
typedef struct fock_scf_out {
    double *coefficients;      /* coefficient_matrix: column k = orbital k */
    double *density;           /* density_matrix */
    double *orbital_energies;  /* ascending */
    double electronic_energy;
    double total_energy;
    int32_t iterations;        /* 0-based index of the returning iteration */
    int32_t n_electrons;
    int32_t n_basis;
    double last_error;
    fock_scf_iter *history;    /* optional, capacity history_capacity */
    int32_t history_capacity;
    int32_t history_len;
    double scf_wall_ms;        /* upload of the inputs through the converged energy */
} fock_scf_out;

/* ----------------------------------------------------------------- context */

/* New context on HIP device `device` with its own stream (default priority). */
fock_status fock_ctx_create(int device, fock_ctx **out);
void fock_ctx_destroy(fock_ctx *ctx);
/* Message of the last failing call on this context ("" if none). */
const char *fock_ctx_last_error(const fock_ctx *ctx);
/* Launch on a caller-provided stream (a hipStream_t cast to void*); NULL restores the
 * context's own stream. Every call's work is ordered on that stream: a blocked J/K build forks
 * its class kernels onto an internal side stream (default priority) with events recorded on
 * it and joins them back before it returns, so the caller sees one stream. */
fock_status fock_ctx_set_stream(fock_ctx *ctx, void *hip_stream);
/* Library version string. */
const char *fock_version(void);

/* ------------------------------------------------------------- the SCF drop-in */

/* RefCompat defaults: max_iters 100, epsilon 1e-6, charge 0, REF_COMPAT, REF_BF_PAIR, FP64. */
void fock_scf_opts_default(fock_scf_opts *opts);

/* Restricted Hartree-Fock; replaces scf::SelfConsistentField::try_scf
 * (scf/src/lib.rs:48-78, body :73-258). Returns FOCK_OK (converged), FOCK_NOT_CONVERGED
 * (iteration limit, outputs filled, the reference's Some), FOCK_DIVERGED (the reference's
 * None) or an error. */
fock_status fock_scf_run(fock_ctx *ctx, const fock_basis_desc *basis, const fock_nuclei_desc *nuclei,
                         const fock_scf_opts *opts, fock_scf_out *out);

/* Many SCFs in lockstep (batching form; the reference has no batch API): molecule m is
 * bases[m] / nuclei[m], its result outs[m] (as fock_scf_run's, history not filled) and, when
 * status is not NULL, status[m]. Every iteration builds the Fock matrices of all unconverged
 * molecules in one launch per quartet class (their shell-pair blocks are interleaved by cost) and
 * solves the eigenproblems of equal-size molecules in one batched call. FP64 form only
 * (precision FP64, no incremental builds, no initial density); the SCF schedule and options are
 * those of fock_scf_run. Returns OK when every molecule converged, else NOT_CONVERGED or DIVERGED. */
fock_status fock_scf_batch_run(fock_ctx *ctx, int32_t count, const fock_basis_desc *bases,
                               const fock_nuclei_desc *nuclei, const fock_scf_opts *opts, fock_scf_out *outs,
                               int32_t *status);

/* Analytic RHF nuclear gradient dE/dR in Eh/Bohr, gradient[3 * atom + axis], at a converged
 * fock_scf_run result (density, coefficients, orbital energies and n_electrons of `scf`). The
 * reference has no gradient (differentiation form): dE/dR = sum P dH - sum W dS +
 * 1/2 sum (P P - exchange) d(ab|cd) + dE_nuc, W = 2 sum_occ eps_i C_i C_i^T, with the energy of
 * scf/src/lib.rs:223 and the nuclear repulsion of :102-116. Accurate integral policy only (the
 * reference Boys table is not differentiable); opts->screen / schwarz_threshold select the
 * quartets as in the SCF: FOCK_SCREEN_SCHWARZ or FOCK_SCREEN_NONE (FOCK_SCREEN_REF_BF_PAIR zeroes
 * single integrals of kept quartets, which leaves the energy non-differentiable: FOCK_INVALID_ARG); opts->precision = FOCK_PRECISION_MIXED evaluates the derivative
 * integrals of quartets with Schwarz bound below 1e-4 in FP32 (FP64 accumulation). Every
 * basis-function centre must coincide with a nucleus. The context's prepared J/K state is the
 * one for this basis and options afterwards: kept when the last preparation (fock_jk_prepare or
 * fock_scf_run) had the same basis arrays and options, prepared again otherwise. */
fock_status fock_rhf_gradient(fock_ctx *ctx, const fock_basis_desc *basis, const fock_nuclei_desc *nuclei,
                              const fock_scf_opts *opts, const fock_scf_out *scf, double *gradient);

/* --------------------------------------------------------- operator surfaces */

/* Boys function F_n(T) for nT arguments; replaces boys::boys (boys/src/lib.rs:58).
 * policy: fock_integral_policy. T and out are host arrays. n <= 10 for REF_COMPAT
 * (else FOCK_BOYS_ORDER), n <= 16 for ACCURATE. */
fock_status fock_boys(fock_ctx *ctx, const double *T, int64_t nT, int32_t n, int32_t policy, double *out);

/* Every ERI (ij|kl) of the basis into out[((i n + j) n + k) n + l] (host, n^4 doubles),
 * screened as requested; replaces ElectronTensor::from_basis + Index<(usize,usize,usize,usize)>
 * (basis/src/electron_tensor.rs:81-187). Intended for small n (test hook). */
fock_status fock_eri_dense(fock_ctx *ctx, const fock_basis_desc *basis, int32_t policy, int32_t screen,
                           double schwarz_threshold, double *out_n4);

/* Overlap, kinetic and nuclear-attraction matrices (host, row-major n x n, symmetric);
 * replaces BasisFunction::{overlap_int, kinetic_int, nuclear_attraction_int}
 * (basis/src/contracted_gaussian.rs:209-262) as assembled by scf/src/lib.rs:118-129.
 * Any output pointer may be NULL. */
fock_status fock_one_electron(fock_ctx *ctx, const fock_basis_desc *basis, const fock_nuclei_desc *nuclei,
                              int32_t policy, double *S, double *T, double *V);

/* Prepare direct J/K builds for a basis: shell pairs, Schwarz bounds, screened task lists.
 * The prepared state lives in the context until the next prepare or destroy. */
fock_status fock_jk_prepare(fock_ctx *ctx, const fock_basis_desc *basis, int32_t policy, int32_t screen,
                            double schwarz_threshold, int32_t precision);

/* J_ij = sum_kl (ij|kl) P_kl and K_ij = sum_kl (ik|jl) P_kl, so that G = J - K/2 is the
 * reference's two-electron Fock term (scf/src/lib.rs:167-185). Host arrays, row-major n x n;
 * P must be symmetric. J or K may be NULL. Synchronous. */
fock_status fock_jk_build(fock_ctx *ctx, const double *P, double *J, double *K);

/* Same on device arrays (HBM), enqueued on the context's stream; no host synchronisation. */
fock_status fock_jk_build_device(fock_ctx *ctx, const double *d_P, double *d_J, double *d_K);

/* Dense FP64 product on the matrix cores, the primitive of the SCF's dense steps
 * (nalgebra products at scf/src/lib.rs:141,145,155,157,191,210,212,223):
 * C = alpha op(A) op(B) + beta C with row-major host arrays, op = transpose when trans* != 0;
 * op(A) is m x k, op(B) is k x n, C is m x n. */
fock_status fock_gemm(fock_ctx *ctx, int32_t transA, int32_t transB, int32_t m, int32_t n, int32_t k, double alpha,
                      const double *A, int32_t lda, const double *B, int32_t ldb, double beta, double *C,
                      int32_t ldc);

/* Eigen-decomposition of a real symmetric n x n matrix on the GPU with eigenvalues ascending;
 * replaces scf::utils::sorted_eigs (scf/src/utils.rs:59-99). M, vectors: column-major host
 * arrays (vectors column k = eigenvector k); values: n. Eigenvector signs are arbitrary. */
fock_status fock_sorted_eigs(fock_ctx *ctx, int32_t n, const double *M, double *vectors, double *values);

/* count eigen-decompositions of symmetric n x n matrices (n <= 128) in one launch, the batching
 * form of fock_sorted_eigs (no reference counterpart; the solver of fock_scf_batch_run): one
 * workgroup per matrix, one-sided Jacobi in LDS. M: count column-major matrices one after the
 * other; vectors, values: as fock_sorted_eigs for each. */
fock_status fock_sorted_eigs_batch(fock_ctx *ctx, int32_t count, int32_t n, const double *M, double *vectors,
                                   double *values);

/* DIIS extrapolation; replaces scf::utils::diis (scf/src/utils.rs:102-138): B_ij = <e_i, e_j>
 * (Frobenius) bordered by -1, solve B c = (0,...,0,-1), out = sum_i c_i F_i. errors and focks
 * hold m matrices of len doubles each (host). Returns FOCK_INVALID_ARG when B is singular (the
 * reference's None). */
fock_status fock_diis(fock_ctx *ctx, int32_t m, int64_t len, const double *errors, const double *focks, double *out);

/* Per-class kernel timing of J/K builds: when enabled, every class launch is bracketed by HIP
 * events recorded on the launch stream, and the classes run one after another on that stream
 * (otherwise they run concurrently). fock_jk_get_timings returns the milliseconds of each of
 * the six class kernels of the most recent build (waits for that build to finish). */
fock_status fock_jk_set_timing(fock_ctx *ctx, int32_t enable);
fock_status fock_jk_get_timings(fock_ctx *ctx, double ms[6]);

/* Statistics of the prepared J/K build, per quartet class (ss|ss), (ps|ss), (ps|ps), (pp|ss),
 * (pp|ps), (pp|pp). Counts are of the work one full build of the prepared precision evaluates:
 * shell quartets of the task lists (the FP32 kernels of the mixed form also drop shell quartets
 * with Q_ab Q_cd < 1e-11), their integrals, and the primitive quartets the accurate policy's
 * primitive screen keeps (1e-15 FP64, 1e-11 for the FP32 kernels of the mixed form). */
typedef struct fock_jk_stats {
    int32_t n_basis, n_shells, n_shell_pairs;
    int64_t shell_quartets[6];
    int64_t primitive_quartets[6];
    int64_t integrals[6];        /* contracted Cartesian integrals (shell quartets x components) */
    int64_t tasks[6];            /* wavefront tasks */
    double algorithmic_flop[6];  /* canonical McMurchie-Davidson count: f_prim(class) per primitive
                                  * quartet + 12 per contracted integral (J: 2 FMA, K: 4 FMA) */
    int64_t near_threshold;      /* REF_BF_PAIR: |estimate - 1e-6| < 1e-12 pairs met */
    int64_t preparations;        /* task-list preparations this context has made so far (a gradient
                                  * that reuses the SCF's lists leaves it unchanged) */
    uint64_t list_checksum;      /* FNV-1a of the prepared pair records, primitive-pair data, task
                                  * and block lists: equal for equal inputs, run to run */
} fock_jk_stats;
fock_status fock_jk_get_stats(fock_ctx *ctx, fock_jk_stats *stats);

/* --------------------------------------------------------- orbitals at points */

/* A regular grid: voxel (x, y, z), 0 <= x < dims[0] etc., is point origin + (x step[0], y step[1],
 * z step[2]) (Bohr) and number x + dims[0] (y + dims[1] z). The volumetric renderer's grid
 * (volumetric/src/main.rs:100-125: N_VOXELS = 200, BOX_SIZE = 20) is dims = 200, origin = -10,
 * step = 0.1 on every axis. */
typedef struct fock_grid_desc {
    double origin[3];
    double step[3];
    int32_t dims[3];
} fock_grid_desc;

/* Molecular orbitals on a grid: out[o * nvox + v] = (float) psi_k(r_v), k = orbitals[o], nvox =
 * dims[0] dims[1] dims[2]. psi_k is MolecularOrbital::evaluate (scf/src/molecular_orbitals.rs:37-43)
 * over the functions f with |C_fk| > 0.05 (MolecularOrbitals::new, :69-90), phi_f as
 * ContractedGaussian::evaluate (basis/src/contracted_gaussian.rs:320-331), evaluated in FP64 and
 * rounded to float as the renderer's texture (volumetric/src/main.rs:115-117). coefficients:
 * column-major n x n host array (fock_scf_out's coefficient matrix); out: host array of
 * n_orbitals * nvox floats. Replaces the rayon loop of create_wave_texture_data
 * (volumetric/src/main.rs:100-125). */
fock_status fock_mo_grid(fock_ctx *ctx, const fock_basis_desc *basis, const double *coefficients,
                         const int32_t *orbitals, int32_t n_orbitals, const fock_grid_desc *grid, float *out);
/* Same into a device array d_out (HBM); the orbital records are uploaded synchronously, the
 * evaluation is enqueued on the context's stream without a host synchronisation. */
fock_status fock_mo_grid_device(fock_ctx *ctx, const fock_basis_desc *basis, const double *coefficients,
                                const int32_t *orbitals, int32_t n_orbitals, const fock_grid_desc *grid,
                                float *d_out);
/* fock_mo_grid_device without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, a renderer's frame loop, a
 * serving loop, harnesses that forbid hipMalloc). The orbital records go into a fixed array of
 * 4 MiB per device, held until the process ends (about 100 orbitals of a 264-function basis);
 * larger inputs are FOCK_INVALID_ARG. Calls on one device share that array: calls ordered on one
 * stream are safe, calls that overlap on one device (other streams or contexts) race. */
fock_status fock_mo_grid_device_noalloc(fock_ctx *ctx, const fock_basis_desc *basis, const double *coefficients,
                                        const int32_t *orbitals, int32_t n_orbitals, const fock_grid_desc *grid,
                                        float *d_out);
/* psi_k at n_points arbitrary points (points[3i..3i+2], Bohr) in FP64: out[i] (host arrays);
 * the pointcloud renderer's MolecularOrbital::evaluate calls (pointcloud/src/main.rs:106-121). */
fock_status fock_mo_points(fock_ctx *ctx, const fock_basis_desc *basis, const double *coefficients, int32_t orbital,
                           int64_t n_points, const double *points, double *out);

/* ------------------------------------------------------------------ autotuner */

/* The blocked J/K digestion runs one workgroup per block of up to `tile` wavefront tasks (64 ket
 * pairs each) of one bra pair; the tile sets the workgroup's LDS and so its occupancy, against
 * the per-block cost of staging the bra pair, the Boys slab and the K rows. Tiles are chosen per
 * quartet class and per form; batched builds also choose the block schedule. The reference has no
 * counterpart (its builds are CPU loops); these entry points serve the batching form. */
typedef enum fock_tile_form {
    FOCK_TILES_FP64 = 0,   /* one molecule, FP64 kernels */
    FOCK_TILES_MIXED = 1,  /* one molecule, mixed form (FP32 bulk kernels) */
    FOCK_TILES_BATCH = 2   /* fock_scf_batch_run */
} fock_tile_form;
typedef enum fock_batch_schedule {
    FOCK_SCHEDULE_COST = 0,  /* the blocks of all molecules by cost, descending */
    FOCK_SCHEDULE_XCD = 1    /* each molecule's blocks on one XCD (L2) of the device: workgroup i
                              * runs on XCD i mod 8, molecules are balanced over the XCDs by cost */
} fock_batch_schedule;
typedef struct fock_tune_report {
    int32_t tile[6];         /* chosen tile per quartet class */
    double ms[6][7];         /* per class and candidate tile 1, 2, 4, 8, 16, 32, 64: best of three
                              * launches in ms (batched: under the chosen schedule); -1 not timed */
    int32_t schedule;        /* batched builds: chosen fock_batch_schedule */
    double schedule_ms[2];   /* batched builds: sum over classes of the best time per schedule */
    double tune_ms;          /* wall time of the tuning */
} fock_tune_report;

/* Time every candidate tile of each class on the prepared J/K build (fock_jk_prepare) at density
 * P (host, row-major n x n), keep the fastest for the prepared build and for this context's later
 * builds of the prepared form (FP64 or mixed). The class kernels of a build run concurrently, so
 * the per-class choice is then refined on whole builds: starting from the fastest of the per-class
 * choice, every class at 32 and every class at 64 (where they fit), each class, longest first,
 * tries every tile with the others fixed and keeps the one with the fastest whole build (timed,
 * best of five), for up to two passes; the result and the best start are then timed again side by
 * side and the faster kept (report->ms holds the per-class times, report->tile the final choice).
 * report may be NULL. */
fock_status fock_jk_autotune(fock_ctx *ctx, const double *P, fock_tune_report *report);
/* Read or set the context's tiles of a form (fock_tile_form); tiles are 1, 2, 4, 8, 16, 32 or 64 (default
 * 8). schedule applies to FOCK_TILES_BATCH (fock_batch_schedule; default COST). tile may be NULL. */
fock_status fock_tiles_get(fock_ctx *ctx, int32_t form, int32_t tile[6], int32_t *schedule);
fock_status fock_tiles_set(fock_ctx *ctx, int32_t form, const int32_t tile[6], int32_t schedule);

#ifdef __cplusplus
}
#endif
#endif /* FOCK_H */
