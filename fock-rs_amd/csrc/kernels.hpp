// kernels.hpp -- launch interface between the host driver and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fock_device.hpp"

namespace fock {

enum { MODE_JK = 0, MODE_DENSE = 1 };

// Quartet classes (bra pair class, ket pair class) with L(bra) >= L(ket):
// 0 (ss|ss), 1 (ps|ss), 2 (ps|ps), 3 (pp|ss), 4 (pp|ps), 5 (pp|pp)
constexpr int NUM_CLASSES = 6;
constexpr int CLASS_BRA[NUM_CLASSES] = {0, 1, 1, 2, 2, 2};  // pair classes: 0 ss, 1 ps, 2 pp
constexpr int CLASS_KET[NUM_CLASSES] = {0, 0, 1, 0, 1, 2};

struct JKArgs {
    const Task *tasks;
    int ntasks;
    int nbf;
    const PairInfo *pairs;
    const void *prims;  // PrimPair<double> or PrimPair<float>
    const double *P;
    double *J;
    double *K;
    const double *D;  // (xy|xy) of every basis-function pair (reference screen)
    double dthr;
    BoysTables bt;
    double *dense;          // MODE_DENSE output, n^4
    const int *ket_list;    // compacted tasks: ket pair = ket_list[ket_begin + lane] (null: contiguous)
    const int *d_ntasks;    // compacted tasks: number of valid tasks on the device (null: ntasks)
    const Block *blocks;    // blocked digestion: groups of tasks sharing one bra pair
    int nblocks;
    int tile;               // blocked digestion: most tasks per block (LDS is sized for it)
    // density-weighted screening (incremental builds): a lane is evaluated only when
    // Q_ab Q_cd max(2|P_ab|, 2|P_cd|, |P_ac|, |P_ad|, |P_bc|, |P_bd|) lies in [bound_lo, bound_hi)
    const int2 *pair_shell;  // null: no density screening
    const float *pairQ;
    const float *Pm;         // [nshell][nshell] max |P| per shell block
    const float *Pm_pair;    // blocked kernel: Pm of every shell pair's own block, by pair
    int nshell;
    float bound_lo, bound_hi;
    // blocked kernel with density screening: a block whose bound pairQ[bra] block_qmax 2 pm_max
    // (an upper bound of every lane's) lies below bound_lo returns at once (null: no block test)
    const float *block_qmax;  // per block: the largest Q of its kets
    const float *pm_max;      // max over all shell blocks of Pm
    unsigned long long *kept_count;  // blocked kernel: += the kets each block evaluates (null: none)
    // without density screening: a lane is evaluated only when Q_ab Q_cd >= ket_tau (0: all;
    // the FP32 full builds cut at their primitive threshold)
    float ket_tau;
    // primitive screening (accurate policy): the primitive quartet (ij|kl) is skipped when
    // q_ij q_kl <= prim_thr, q = sqrt(max_component (ij|ij)) of the primitive pair including its
    // contraction coefficients (a Cauchy-Schwarz bound on its contribution). Primitive pairs of
    // each shell pair are stored in descending q. null: every primitive quartet is evaluated.
    const float *prim_q;
    float prim_thr;
    // ket-side primitive data (blocked kernel): per (class, K) bucket of nb pairs, field f of
    // primitive k of the bucket's i-th pair at soa[start + (k * PRIM_FIELDS + f) * nb], with
    // start = soa_pairs[j].pp_start and nb = soa_stride[j]; fields p, Px, Py, Pz, Kc, xa(3),
    // xb(3), 1/2p, q
    const void *soa;
    const PairInfo *soa_pairs;
    const int *soa_stride;
    // ket-side J of the blocked kernel, by pair and component: J[fc + c][fd + d] of ket pair j
    // accumulates in Jk[(c * ND + d) * npair + j] (consecutive lanes hold consecutive pairs, so
    // every atomic instruction covers a few cache lines), scattered into J after the build
    double *Jk;
    int npair;
    // batched launch (many molecules): block i of the launch is block bblocks[i].y of record
    // bblocks[i].x, whose arrays replace the ones above; null: one molecule
    const struct BatchRec *batch;
    const int2 *bblocks;
};
// Iteration state of one molecule of the device-driven batched SCF: read and written by the kernels
// of the iteration graph (the host reads it after the last iteration).
struct BatchState {
    int active;          // 1 while the molecule iterates; every kernel of the iteration skips it at 0
    int status;          // fock_status of the exit
    int iter;            // index of the iteration being run
    int iter_done;       // index of the returning iteration
    int ring_len, next_slot;
    int ring[16];        // ring slots of the stored (error, Fock) pairs, oldest first
    int ncoef;           // the Fock matrix of this iteration: sum_i coef[i] focks[cslot[i]]
    int cslot[16];
    int shift;           // 1: the level shift applies this iteration
    int phase;           // 0: FP32 quartet integrals (mixed form), 1: FP64
    int stall;
    int pad_;
    double coef[16];
    double best_err;
    double err, e_elec;
    double maxabs;       // max |FPS - SPF| of this iteration (accumulated by the error kernel)
    double dots[16];     // <e_new, e_slot> by ring slot
    double B[16 * 16];   // <e_i, e_j> of the stored pairs in ring order
};
// one molecule's arrays for one quartet class of a batched Fock build
struct BatchRec {
    const Task *tasks;
    const Block *blocks;
    const PairInfo *pairs;
    const void *prims;
    const float *prim_q;
    const void *soa;
    const PairInfo *soa_pairs;
    const int *soa_stride;
    const double *D;
    const double *P;
    double *J, *K, *Jk;
    int npair, nbf;
    int active;  // 0: the molecule has converged, its blocks return at once
    // device-driven iteration: the blocks of a molecule run only while state->active and when the
    // molecule's phase matches the kernel's precision (FP32 kernels: phase 0); null: host-driven
    const BatchState *state;
    const void *prims32, *soa32;  // FP32 copies of prims and soa (mixed form)
    const float *pairQ;           // shell-pair Schwarz factors (the FP32 kernels' shell-quartet cut)
};
constexpr int PRIM_FIELDS = 13;
// largest primitive-pair count of a shell pair (contractions of up to 6 primitives)
constexpr int MAX_PRIM_PAIRS = 36;
// tasks of one bra pair per workgroup of the blocked digestion (one wave per task at a time):
// the tile, chosen per quartet class from TILE_CHOICES (default DEFAULT_TILE) by the autotuner
constexpr int DEFAULT_TILE = 8;
constexpr int MAX_TILE = 64;
constexpr int NUM_TILE_CHOICES = 7;
constexpr int TILE_CHOICES[NUM_TILE_CHOICES] = {1, 2, 4, 8, 16, 32, 64};
// largest basis for the blocked digestion: (3 + 3) K rows of n doubles in LDS next to the staged
// Boys slab (at most 33 KB) and the block's ket records (0.3 KB per task of the tile)
constexpr int BLOCKED_MAX_NBF = 600;
// whether the blocked kernel of total angular momentum L stages its Boys slab in LDS (accurate policy)
constexpr bool block_slab_in_lds(int L, bool fp32) { return fp32 || L == 0; }

// density-weighted screening of the static task list (compact.hip)
struct ScreenArgs {
    const Task *tasks;
    int ntasks;
    const int2 *pair_shell;  // shells (a, b) of each pair
    const float *pairQ;      // sqrt(max_component (ab|ab))
    const float *Pm;         // [nshell][nshell] max |dP| per shell block
    int nshell;
    float tau, tau64;
    int *flag32, *flag64;    // per slot (task * 64 + lane)
};
struct CompactBuffers {
    int *pos, *kets, *head, *tid, *kbra, *kend, *kcost;
    Task *tasks;
    int *count;
    void *tmp;
    size_t tmp_bytes;
};
// preparation (prepare.hip): primitive pairs reordered by the host's permutation src (with their
// exponent pairs), converted to FP32, and laid out as the ket-side bucket SoA
hipError_t permute_prims(int64_t n, const int *src, const PrimPair<double> *in, const double2 *ab_in,
                         PrimPair<double> *out, double2 *ab_out, hipStream_t s);
hipError_t prims_to_f32(int64_t n, const PrimPair<double> *in, PrimPair<float> *out, hipStream_t s);
hipError_t build_soa(int64_t n, const int *prim_pair, const int *prim_k, const PairInfo *sinfo, const int *stride,
                     const PrimPair<double> *prims, const float *q, double *soa64, float *soa32, hipStream_t s);

hipError_t screen_flags(const ScreenArgs &a, hipStream_t s);
size_t compact_temp_bytes(int64_t slots);
hipError_t compact_tasks(const Task *tasks, int ntasks, const int *gfirst, const int *glast, const int *flag,
                         const CompactBuffers &b, hipStream_t s);
// Gacc += (Jb + Jb^T) - 0.5 (Kb + Kb^T)
hipError_t accumulate_g(int n, const double *Jb, const double *Kb, double *Gacc, hipStream_t s);
// one molecule of a batched scatter_jk
struct ScatterRec {
    int npair, n;
    const int2 *pair_shell;
    const int *shell_f0, *shell_nf;
    double *Jk, *J;
    const int *flag;  // skipped when *flag == 0 (null: always)
};
hipError_t scatter_jk_batch(int nrec, const ScatterRec *recs, int max_npair, hipStream_t s);
// J[f0(a) + c][f0(b) + d] += Jk[(c * nf(b) + d) * npair + j] for every pair j = (a, b), then Jk = 0
hipError_t scatter_jk(int npair, const int2 *pair_shell, const int *shell_f0, const int *shell_nf, double *Jk,
                      double *J, int n, hipStream_t s);
// Batched dense steps of the batched SCF: one launch per step for all molecules, one record each.
struct GemmRec {  // C = alpha op(A) op(B) + beta C, op(A) m x k, op(B) k x n, row-major with leading dims
    int m, n, k, lda, ldb, ldc;
    const double *A, *B;
    double *C;
    const int *flag;  // skipped when *flag == 0 (null: always)
};
hipError_t gemm_f64_batch(bool transA, bool transB, double alpha, double beta, const GemmRec *recs, int count,
                          int max_m, int max_n, hipStream_t s);
struct DiisRec {  // one molecule's DIIS error step
    int n;
    const double *FPS;   // F P S
    double *e;           // out: FPS - (FPS)^T (the ring slot)
    const double *F;
    double *f_slot;      // out: copy of F (the ring slot)
    const double *errs;  // the 16 error slots, n * n apart
    double *dots;        // out: dots[s] = <e, errs[s]> (16), dots[16] = max |e| (zeroed beforehand)
};
hipError_t diis_error_batch(const DiisRec *recs, int count, int max_n, hipStream_t s);
struct CombRec {  // out = sum_i c[i] base[slots[i]] (matrices of len doubles)
    int64_t len;
    const double *base;
    const int *slots;
    const double *c;
    int count;
    double *out;
};
hipError_t lincomb_batch(const CombRec *recs, int count, int64_t max_len, hipStream_t s);
struct ShiftRec {  // F += sigma (S - SPS / 2)
    int64_t len;
    double *F;
    const double *S, *SPS;
};
hipError_t level_shift_batch(const ShiftRec *recs, int count, int64_t max_len, double sigma, hipStream_t s);
struct TraceRec {  // out = 1/2 sum P (2H + G)
    int n;
    const double *P, *H, *G;
    double *out;
};
hipError_t energy_trace_batch(const TraceRec *recs, int count, hipStream_t s);

// ----- device-driven batched SCF (batch_scf.hip): every step one launch over the molecules,
// controlled by each molecule's BatchState, so that one iteration is one captured graph
struct BatchParams {
    int max_iters, diis_start, diis_window;
    int accurate;  // accurate policy: the returning iteration diagonalises the newest Fock matrix
    int mixed;     // FP32 quartets until the error falls below max(10 eps, mixed_switch) or stalls
    double eps, level_shift, level_shift_until, mixed_switch, mixed_stall;
};
struct BatchMolRec {  // one molecule's matrices for the state-driven steps
    int n;
    BatchState *st;
    double *Jb, *Kb;          // J/K accumulators of the build
    const double *H, *S;
    double *G, *F, *Fx;
    const double *FPS;        // F P S (the DIIS error), later S P S (the level shift)
    double *errs, *focks;     // 16 ring slots each, n * n apart
    const double *U;          // eigenvectors turned back with the previous basis: copied to E
    double *E;
    const double *P;
};
// zero J/K accumulators and the per-iteration maxima of the active molecules; *active_count = 0
hipError_t batch_begin(const BatchMolRec *recs, int count, int max_n, int *active_count, hipStream_t s);
// G = (Jb + Jb^T) - (Kb + Kb^T) / 2, F = H + G
hipError_t batch_fock_from_jk(const BatchMolRec *recs, int count, int max_n, hipStream_t s);
// e = FPS - (FPS)^T into the next ring slot (and F into its Fock slot), max |e|, <e, e_s> for the 16 slots
hipError_t batch_diis_error(const BatchMolRec *recs, int count, int max_n, hipStream_t s);
// ring bookkeeping, DIIS solve (or the newest matrix), level-shift decision, divergence
hipError_t batch_diis_update(const BatchMolRec *recs, int count, BatchParams p, hipStream_t s);
// Fx = sum coef_i focks[cslot_i]; Fx += sigma (S - SPS / 2) where the shift applies; E = U
hipError_t batch_lincomb(const BatchMolRec *recs, int count, int max_n, hipStream_t s);
hipError_t batch_level_shift(const BatchMolRec *recs, int count, int max_n, double sigma, hipStream_t s);
hipError_t batch_copy_basis(const BatchMolRec *recs, int count, int max_n, hipStream_t s);
// E_elec = 1/2 tr(P (2H + G)) of every active molecule into its state
hipError_t batch_energy(const BatchMolRec *recs, int count, hipStream_t s);
// exit rule of scf/src/lib.rs:224-249 per molecule, phase changes of the mixed form, active count
hipError_t batch_exit(const BatchMolRec *recs, int count, BatchParams p, int *active_count, hipStream_t s);

// one matrix of a batched symmetric eigenproblem (n = 0: skipped)
struct EigRec {
    int n;
    double *A;  // in: symmetric n x n; out: eigenvector k (of the k-th smallest eigenvalue) at A[k * n]
    double *W;  // out: eigenvalues, ascending
    const int *flag;  // skipped when *flag == 0 (null: always)
};
constexpr int JACOBI_MAX_N = 128;  // the matrix is held in LDS (128 KB at the limit)
hipError_t jacobi_eig_batch(const EigRec *recs, int count, int max_n, hipStream_t s);
// out = F + sigma (S - 0.5 SPS)
hipError_t level_shift_fock(int64_t len, const double *F, const double *S, const double *SPS, double sigma,
                            double *out, hipStream_t s);

struct DiagArgs {
    const PairInfo *pairs;
    int first, npairs, nbf;
    const PrimPair<double> *prims;
    double *D;
    BoysTables bt;
};

// q of every primitive pair of one pair class: one thread per primitive pair
struct PrimDiagArgs {
    const int *prim_index;  // primitive-pair indices of this class
    int count;
    const PrimPair<double> *prims;
    float *q;  // q[prim_index[t]] = sqrt(max_component (ij|ij))
    BoysTables bt;
};
// surviving primitive quartets of a task list (statistics): out[0] += count
struct PrimCountArgs {
    const Task *tasks;
    int ntasks;
    const PairInfo *pairs;
    const float *prim_q;
    float prim_thr;
    const float *pairQ;  // shell-quartet cut Q_ab Q_cd >= ket_tau when ket_tau > 0 (FP32 kernels)
    float ket_tau;
    unsigned long long *out;        // kept primitive quartets
    unsigned long long *out_shell;  // kept shell quartets
};

// nuclear gradient, two-electron part: one lane per unique shell quartet of the task lists
struct GradEriArgs {
    const Task *tasks;
    int ntasks;
    const PairInfo *pairs;
    const PrimPair<double> *prims;
    const PrimPair<float> *prims32;  // mixed form
    const double2 *prim_ab;  // exponents (alpha_a, alpha_b) of every primitive pair
    // mixed form (pairQ not null): tasks whose bound Q_bra Q_ket(first) reaches tau64 run in FP64,
    // the others in FP32; null: every task in FP64
    const float *pairQ;
    float tau64;
    const int2 *pair_shell;
    const int *shell_atom;
    const double *P;
    int nbf, natoms;
    BoysTables bt;
    double *grad;  // [natoms][3], accumulated
    int split;     // waves per task, each taking every split-th primitive quartet (set by the launcher)
    // primitive screen: the primitive quartet (ij|kl) is skipped when q_ij q_kl <= prim_thr (q as in
    // JKArgs::prim_q); null: every primitive quartet is evaluated
    const float *prim_q;
    float prim_thr;
};
// nuclear gradient, one-electron part: one thread per (shell pair, term) of one pair class
struct GradOneArgs {
    int nbf, natoms, npairs;
    const int *pair_a, *pair_b, *shell_f0, *shell_atom;
    const double *centers;
    const int *prim_off;
    const double *exps, *coefs;
    const double *nuc_pos, *nuc_charge;
    const double *P, *W;
    BoysTables bt;
    double *grad;
};

hipError_t launch_eri_tasks(int cls, int mode, const JKArgs &a, int policy, bool refscreen, bool fp32,
                            hipStream_t s);
hipError_t launch_grad_eri(int cls, const GradEriArgs &a, hipStream_t s);
hipError_t launch_grad_one(int pair_class, const GradOneArgs &a, hipStream_t s);
hipError_t energy_weighted_density(int n, int nocc, const double *C, const double *eps, double *W, hipStream_t s);
hipError_t launch_prim_diag(int pair_class, const PrimDiagArgs &a, hipStream_t s);
hipError_t launch_prim_count(const PrimCountArgs &a, hipStream_t s);
// mo_grid.hip: orbitals at points (regular grid or a list); the records of every orbital are built
// on the host from the basis and the coefficient matrix (kept functions grouped by shell)
constexpr int MO_MAXK = 8;  // most primitives per contracted function
struct MoShellRec {
    double A[3];
    double amin;     // smallest exponent of the shell
    int K, nf;       // primitives; kept functions of this orbital on the shell
    int alpha;       // dbl[alpha .. alpha + K): exponents
    int fun;         // funs[fun .. fun + nf)
};
struct MoFunRec {
    int ang[3];
    int w;           // dbl[w .. w + K): C_fk * coef_p
};
struct MoArgs {
    const MoShellRec *shells;
    const MoFunRec *funs;
    const double *dbl;
    const int *shell_off;  // orbital o: shells[shell_off[o] .. shell_off[o + 1])
    int norb;
    int64_t npoints;
    double origin[3], step[3];  // grid form: point (x, y, z) = origin + (x, y, z) * step
    int dims[3];                // grid form: point number x + dims[0] (y + dims[1] z)
    const double *points;       // list form: [npoints][3]
    void *out;                  // float or double, out[o * out_stride + v]
    int64_t out_stride;
};
hipError_t launch_mo_eval(const MoArgs &a, bool grid, bool f32, hipStream_t s);
// the allocation-free orbital evaluation keeps its records in a fixed device array of this size
// (one per device); its address on `device` (which must be current), looked up once per device
constexpr size_t MO_SCRATCH_BYTES = (size_t)4 << 20;
constexpr int MO_SCRATCH_MAX_DEVICES = 64;
hipError_t mo_scratch_address(int device, void **out);

hipError_t launch_eri_diag(int pair_class, const DiagArgs &a, int policy, hipStream_t s);
hipError_t launch_eri_blocks(int cls, const JKArgs &a, int policy, bool refscreen, bool fp32, hipStream_t s);
// dynamic LDS of one workgroup of the blocked kernel of class cls
size_t block_lds_bytes(int cls, int nbf, int policy, bool fp32, int tile);

// one_electron.hip
struct OneElArgs {
    int nbf, natoms;
    const double *centers;
    const int *angular;
    const int *prim_off;
    const double *exps;
    const double *coefs;
    const double *nuc_pos;
    const double *nuc_charge;
    double *S, *T, *V;
    BoysTables bt;
};
hipError_t launch_one_electron_class(int pair_class, const OneElArgs &a, const int *shell_f0, const int *pair_a,
                                     const int *pair_b, int npairs, int policy, hipStream_t s);

// boys_kernel.hip
hipError_t launch_boys(const double *T, int64_t nT, int n, int policy, const BoysTables &bt, double *out,
                       hipStream_t s);

// dense.hip: FP64 MFMA GEMM and small dense helpers (row-major n x n unless stated)
// C = alpha * op(A) op(B) + beta * C, op = transpose if flag set; A: m x k, B: k x n, C: m x n
hipError_t gemm_f64(bool transA, bool transB, int m, int n, int k, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *C, int ldc, hipStream_t s);
// out = x + x^T (n x n)
hipError_t symmetrize_add(int n, const double *x, double *out, hipStream_t s);
// F = H + J - 0.5 K (with J, K the non-symmetric accumulation buffers, symmetrised here)
hipError_t fock_from_jk(int n, const double *H, const double *Jb, const double *Kb, double *G, double *F,
                        double scaleG, hipStream_t s);
// e = M - M^T, and reduce max|e| into *maxabs (device scalar, zero-initialised by the kernel)
hipError_t antisym_maxabs(int n, const double *M, double *e, double *maxabs, hipStream_t s);
// dots[i] = <a, b_i> (Frobenius) for i < count, b_i = base + i * stride
hipError_t frob_dots(int64_t len, const double *a, const double *base, int64_t stride, int count, double *dots,
                     hipStream_t s);
// out = sum_i c[i] * (base + slot[i] * stride), c and slots on device
hipError_t lincomb(int64_t len, const double *base, int64_t stride, const int *slots, const double *c, int count,
                   double *out, hipStream_t s);
// out[i][j] = 2 * sum_{k<nocc} C[i][k] C[j][k]  (C row-major n x n)
hipError_t density_from_coeffs(int n, int nocc, const double *C, double *P, hipStream_t s);
// 0.5 * sum_ij P_ij (2 H_ij + G_ij) = the sum of parts[0 .. ENERGY_PARTS), added in order
constexpr int ENERGY_PARTS = 32;
hipError_t energy_trace(int n, const double *P, const double *H, const double *G, double *parts, hipStream_t s);
// x = a - b elementwise (len)
hipError_t sub_f64(int64_t len, const double *a, const double *b, double *out, hipStream_t s);
hipError_t add_f64(int64_t len, const double *a, const double *b, double *out, hipStream_t s);
// out = lam a + (1 - lam) b (density damping)
hipError_t mix_f64(int64_t len, double lam, const double *a, const double *b, double *out, hipStream_t s);
hipError_t transpose_f64(int n, const double *in, double *out, hipStream_t s);
hipError_t fill_f64(int64_t len, double v, double *out, hipStream_t s);
hipError_t copy_f64(int64_t len, const double *in, double *out, hipStream_t s);
// trace-correcting purification helpers: Gershgorin bounds {lo, hi} of a symmetric matrix,
// X = (hi I - A)/(hi - lo), out = {tr X, tr X2, grow} (grow: tr(2X - X2) closer to target than
// tr X2), X = 2X - X2 if *grow else X2 (the choice stays on the device)
hipError_t gershgorin_bounds(int n, const double *A, double *out, hipStream_t s);
// Newton-Schulz S^{-1/2}: Y = S / bound[1], Z = I; T = (3 I - ZY) / 2 with *res = max(*res, max|I - ZY|);
// X = (Z + Z^T) / (2 sqrt(bound[1]))
hipError_t ns_init(int n, const double *S, const double *bound, double *Y, double *Z, hipStream_t s);
hipError_t ns_update(int n, const double *ZY, double *T, double *res, hipStream_t s);
hipError_t ns_finish(int n, const double *Z, const double *bound, double *X, hipStream_t s);
// Gershgorin interval of every row: rowb[2i], rowb[2i + 1]
hipError_t gershgorin_rows(int n, const double *A, double *rowb, hipStream_t s);
// {lo, hi} = the union of the row intervals (also written to bounds[0..1]), X = (hi I - A) / (hi - lo),
// per-workgroup (tr, sum of squares) partials in part[tc2_init_parts(n)][2]
int tc2_init_parts(int n);
hipError_t tc2_init(int n, const double *A, const double *rowb, double *bounds, double *X, double *part,
                    hipStream_t s);
// one purification step X -> Xout (X^2 or 2X - X^2, chosen on the device from part_in), the new
// partials in part_out[tc2_step_parts(n)][2], hist = {tr X, tr X^2, grow} of the input
int tc2_step_parts(int n);
hipError_t tc2_step(int n, const double *X, double *Xout, const double *part_in, int npart_in, double *part_out,
                    double nocc, double *hist, hipStream_t s);
// GWH: out_ij = 0.875 S_ij (H_ii + H_jj)
hipError_t gwh_guess(int n, const double *S, const double *H, double *out, hipStream_t s);
// X = U diag(1/sqrt(D_ii)) U^T with D = U^T S U: scale columns of U by dinv (device vector)
hipError_t scale_columns(int n, const double *U, const double *dvec_diag_source, int ld_diag, double *out,
                         hipStream_t s);
// per shell-pair max |P| blocks for density-weighted screening
// out[j] = Pm[pair_shell[j].x][pair_shell[j].y] for every shell pair j
hipError_t pair_block_max(int npair, const int2 *pair_shell, const float *Pm, int nshell, float *out, hipStream_t s);
// *out = max of x[0..len) (non-negative values; NaN propagates), one workgroup
hipError_t max_f32(int64_t len, const float *x, float *out, hipStream_t s);
hipError_t shell_block_maxabs(int nshell, const int *shell_f0, const int *shell_nf, int n, const double *P,
                              float *out, hipStream_t s);

}  // namespace fock
