// eri_kernels.hip -- direct-SCF shell-quartet ERI kernels with fused J/K digestion.
//
// Replaces the dense electron tensor (basis/src/electron_tensor.rs:81-176) and the
// per-iteration n^4 contraction (scf/src/lib.rs:167-185) by a direct build.
//
// Work decomposition: one wavefront (64 lanes) per task. A task is one bra shell pair
// and up to 64 ket shell pairs of one quartet class, one ket pair per lane. Every lane
// of a wavefront shares the bra pair, so the bra primitive data are wave-uniform loads
// and the J_ab contribution is reduced across the wavefront before a single atomic.
// Ket pairs of a task have the same primitive-pair count (the host buckets them), so
// the primitive loops are uniform and no lane idles inside them.
//
// Digestion of a unique quartet (ab|cd) with value v and degeneracy weight
// w = (a==b ? 1/2 : 1)(c==d ? 1/2 : 1)(ab==cd ? 1/2 : 1) into non-symmetric buffers:
//   J[a][b] += 2w v P[c][d]   J[c][d] += 2w v P[a][b]
//   K[a][c] += w v P[b][d]    K[a][d] += w v P[b][c]   K[b][c] += w v P[a][d]   K[b][d] += w v P[a][c]
// and J = J + J^T, K = K + K^T afterwards, which sums every one of the 8 symmetric
// images of each integral exactly once.
#include "fock_device.hpp"
#include "kernels.hpp"

namespace fock {

// Contracted integrals of one lane's quartet (bra pair of the task, ket pair of the lane), with
// the reference's pair screen applied when REFSCREEN; inactive lanes return zeros.
template <int LA, int LB, int LC, int LD, int POLICY, bool REFSCREEN, typename Real>
__device__ __forceinline__ void lane_integrals(const JKArgs &args, const PairInfo &bra, const PairInfo &ket,
                                               bool active,
                                               Real I[ncart(LA) * ncart(LB)][ncart(LC) * ncart(LD)]) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    const PrimPair<Real> *__restrict__ pp = reinterpret_cast<const PrimPair<Real> *>(args.prims);
#pragma unroll
    for (int x = 0; x < NAB; ++x)
#pragma unroll
        for (int y = 0; y < NCD; ++y) I[x][y] = Real(0);

    if (args.prim_q) {
        // primitive pairs are in descending q: once no lane keeps (i, k), none keeps (i, k' > k)
        for (int i = 0; i < bra.K; ++i) {
            const PrimPair<Real> b = pp[bra.pp_start + i];
            const float qb = args.prim_q[bra.pp_start + i];
            for (int k = 0; k < ket.K; ++k) {
                const bool keep = qb * args.prim_q[ket.pp_start + k] > args.prim_thr;
                if (__ballot(keep) == 0) break;
                if (keep) {
                    const PrimPair<Real> kq = pp[ket.pp_start + k];
                    accumulate_prim_quartet<LA, LB, LC, LD, POLICY, Real>(b, kq, args.bt, I);
                }
            }
        }
    } else {
        for (int i = 0; i < bra.K; ++i) {
            const PrimPair<Real> b = pp[bra.pp_start + i];
            for (int k = 0; k < ket.K; ++k) {
                const PrimPair<Real> kq = pp[ket.pp_start + k];
                accumulate_prim_quartet<LA, LB, LC, LD, POLICY, Real>(b, kq, args.bt, I);
            }
        }
    }

    const int n = args.nbf;
    const int fa = bra.fa, fb = bra.fb, fc = ket.fa, fd = ket.fb;
    if constexpr (REFSCREEN) {
        // the reference's basis-function pair screen: keep (xy|zw) iff (xy|xy)(zw|zw) > thr,
        // thr = the largest double whose square root does not exceed 1e-6
        // (electron_tensor.rs:134-155); the diagonal integrals (xy|xy) themselves are computed
        // unconditionally in its first pass (:89-101) and are never screened
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const double dab = args.D[(size_t)(fa + a) * n + (fb + b)];
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        const double dcd = args.D[(size_t)(fc + c) * n + (fd + d)];
                        const bool diag = (fa + a == fc + c && fb + b == fd + d) || (fa + a == fd + d && fb + b == fc + c);
                        if (!diag && !(dab * dcd > args.dthr)) I[a * NB + b][c * ND + d] = Real(0);
                    }
            }
    }
    if (!active) {
#pragma unroll
        for (int x = 0; x < NAB; ++x)
#pragma unroll
            for (int y = 0; y < NCD; ++y) I[x][y] = Real(0);
    }
}

// Primitive k of a ket pair from the bucket-SoA copy: base = soa + start, stride = bucket size.
template <typename Real>
__device__ __forceinline__ PrimPair<Real> soa_prim(const Real *__restrict__ base, int k, int stride) {
    const Real *f = base + (size_t)k * PRIM_FIELDS * stride;
    PrimPair<Real> q;
    q.p = f[0];
    q.Px = f[stride];
    q.Py = f[2 * stride];
    q.Pz = f[3 * stride];
    q.Kc = f[4 * stride];
    q.xax = f[5 * stride];
    q.xay = f[6 * stride];
    q.xaz = f[7 * stride];
    q.xbx = f[8 * stride];
    q.xby = f[9 * stride];
    q.xbz = f[10 * stride];
    q.inv2p = f[11 * stride];
    return q;
}

// Bra primitive pair i of a record read through the scalar data path (every lane of the workgroup
// reads the same record).
template <typename Real>
__device__ __forceinline__ PrimPair<Real> scalar_prim(ConstPtr<Real> base, int i) {
    const ConstPtr<Real> f = base + i * 12;
    PrimPair<Real> b;
    b.p = f[0];
    b.Px = f[1];
    b.Py = f[2];
    b.Pz = f[3];
    b.Kc = f[4];
    b.xax = f[5];
    b.xay = f[6];
    b.xaz = f[7];
    b.xbx = f[8];
    b.xby = f[9];
    b.xbz = f[10];
    b.inv2p = f[11];
    return b;
}

// Issues the scalar loads of a bra primitive's q and record at this point: an empty asm that takes
// them as operands, so that the compiler waits once for all of them before the screen test instead
// of once for q and again for the record inside the branch that uses it (the loop's two scalar-load
// waits per bra primitive were the FP32 primitive loop's longest stalls). LAB = 0 reads no Hermite
// data.
template <int LAB, typename Real>
__device__ __forceinline__ void bra_loads_here(float q, const PrimPair<Real> &b) {
    asm volatile("" ::"s"(q), "s"(b.p), "s"(b.Px), "s"(b.Py), "s"(b.Pz), "s"(b.Kc));
    if constexpr (LAB > 0)
        asm volatile("" ::"s"(b.xax), "s"(b.xay), "s"(b.xaz), "s"(b.xbx), "s"(b.xby), "s"(b.xbz), "s"(b.inv2p));
}

// Same integrals with the Boys slab in LDS, the ket primitive pairs read from the bucket-SoA copy
// (the 64 lanes of a task read consecutive entries of every field: each ket primitive pair is
// loaded once per lane and reused for every bra primitive), and the bra primitive pairs and their
// q read with scalar loads (wave-uniform: SGPR operands, no LDS traffic). ket.pp_start is the
// pair's SoA start, ket_stride its bucket size.
template <int LA, int LB, int LC, int LD, int POLICY, bool REFSCREEN, typename Real, bool SLAB>
__device__ __forceinline__ void lane_integrals_lds(const JKArgs &args, const PairInfo &bra, const PairInfo &ket,
                                                   int ket_stride, bool active, LdsPtr<Real> slice,
                                                   Real I[ncart(LA) * ncart(LB)][ncart(LC) * ncart(LD)]) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    const Real *__restrict__ kbase = reinterpret_cast<const Real *>(args.soa) + ket.pp_start;
    // the bra pair is the workgroup's: its range is uniform
    const int bstart = __builtin_amdgcn_readfirstlane(bra.pp_start), bK = __builtin_amdgcn_readfirstlane(bra.K);
    const ConstPtr<Real> cb = (ConstPtr<Real>)(reinterpret_cast<const Real *>(args.prims) + (size_t)bstart * 12);
#pragma unroll
    for (int x = 0; x < NAB; ++x)
#pragma unroll
        for (int y = 0; y < NCD; ++y) I[x][y] = Real(0);
    if (args.prim_q) {
        // primitive pairs are in descending q (bra and ket): the kept bra primitives of a ket
        // primitive form a prefix, and the wave stops at the first (k, i) no lane keeps
        const ConstPtr<float> cq = (ConstPtr<float>)(args.prim_q + bstart);
        const float q0 = cq[0], thr = args.prim_thr;
        if constexpr (sizeof(Real) == 4) {
            // FP32: ket primitive k + 1 is loaded while primitive k is evaluated (every pair has
            // K >= 1); in FP64 the extra registers cost the larger classes more than it saves
            float qn = active ? (float)kbase[(size_t)12 * ket_stride] : 0.0f;
            PrimPair<Real> kn = soa_prim(kbase, 0, ket_stride);
            for (int k = 0; k < ket.K; ++k) {
                const float qk = qn;
                const PrimPair<Real> kq = kn;
                if (k + 1 < ket.K) {
                    qn = active ? (float)kbase[((size_t)(k + 1) * PRIM_FIELDS + 12) * ket_stride] : 0.0f;
                    kn = soa_prim(kbase, k + 1, ket_stride);
                }
                if (__ballot(qk * q0 > thr) == 0) break;
                for (int i = 0; i < bK; ++i) {
                    // the record is read with its q (one scalar-load wait per bra primitive, not
                    // one for q and another for the record after the test)
                    const PrimPair<Real> bp = scalar_prim(cb, i);
                    const float qi = cq[i];
                    bra_loads_here<LA + LB>(qi, bp);
                    const bool keep = qi * qk > thr;
                    if (__ballot(keep) == 0) break;
                    if (keep) accumulate_prim_quartet<LA, LB, LC, LD, POLICY, Real, SLAB>(bp, kq, args.bt, I, slice);
                }
            }
        } else {
            for (int k = 0; k < ket.K; ++k) {
                const float qk = active ? (float)kbase[((size_t)k * PRIM_FIELDS + 12) * ket_stride] : 0.0f;
                if (__ballot(qk * q0 > thr) == 0) break;
                const PrimPair<Real> kq = soa_prim(kbase, k, ket_stride);
                for (int i = 0; i < bK; ++i) {
                    const PrimPair<Real> bp = scalar_prim(cb, i);
                    const float qi = cq[i];
                    bra_loads_here<LA + LB>(qi, bp);
                    const bool keep = qi * qk > thr;
                    if (__ballot(keep) == 0) break;
                    if (keep) accumulate_prim_quartet<LA, LB, LC, LD, POLICY, Real, SLAB>(bp, kq, args.bt, I, slice);
                }
            }
        }
    } else {
        for (int k = 0; k < ket.K; ++k) {
            const PrimPair<Real> kq = soa_prim(kbase, k, ket_stride);
            for (int i = 0; i < bK; ++i)
                accumulate_prim_quartet<LA, LB, LC, LD, POLICY, Real, SLAB>(scalar_prim(cb, i), kq, args.bt, I, slice);
        }
    }
    const int n = args.nbf;
    const int fa = bra.fa, fb = bra.fb, fc = ket.fa, fd = ket.fb;
    if constexpr (REFSCREEN) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const double dab = args.D[(size_t)(fa + a) * n + (fb + b)];
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        const double dcd = args.D[(size_t)(fc + c) * n + (fd + d)];
                        const bool diag = (fa + a == fc + c && fb + b == fd + d) || (fa + a == fd + d && fb + b == fc + c);
                        if (!diag && !(dab * dcd > args.dthr)) I[a * NB + b][c * ND + d] = Real(0);
                    }
            }
    }
    if (!active) {
#pragma unroll
        for (int x = 0; x < NAB; ++x)
#pragma unroll
            for (int y = 0; y < NCD; ++y) I[x][y] = Real(0);
    }
}

template <int LA, int LB, int LC, int LD, int POLICY, int MODE, bool REFSCREEN, typename Real>
__global__ __launch_bounds__(256) void eri_task_kernel(JKArgs args) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= args.ntasks) return;  // wave-uniform
    if (args.d_ntasks && wave >= *args.d_ntasks) return;
    const Task task = args.tasks[wave];
    const bool active = lane < task.ket_count;
    const PairInfo bra = args.pairs[task.bra];
    const int slot = task.ket_begin + (active ? lane : 0);
    const int ket_idx = args.ket_list ? args.ket_list[slot] : slot;
    const PairInfo ket = args.pairs[ket_idx];
    Real I[NAB][NCD];
    lane_integrals<LA, LB, LC, LD, POLICY, REFSCREEN, Real>(args, bra, ket, active, I);
    const int n = args.nbf;
    const int fa = bra.fa, fb = bra.fb, fc = ket.fa, fd = ket.fb;

    if constexpr (MODE == MODE_DENSE) {
        if (!active) return;
        double *out = args.dense;
        const size_t N = (size_t)n;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        const size_t i = fa + a, j = fb + b, k = fc + c, l = fd + d;
                        const double v = (double)I[a * NB + b][c * ND + d];
                        out[((i * N + j) * N + k) * N + l] = v;
                        out[((j * N + i) * N + k) * N + l] = v;
                        out[((i * N + j) * N + l) * N + k] = v;
                        out[((j * N + i) * N + l) * N + k] = v;
                        out[((k * N + l) * N + i) * N + j] = v;
                        out[((l * N + k) * N + i) * N + j] = v;
                        out[((k * N + l) * N + j) * N + i] = v;
                        out[((l * N + k) * N + j) * N + i] = v;
                    }
        return;
    } else {
        Real w = Real(1);
        if (fa == fb) w *= Real(0.5);
        if (fc == fd) w *= Real(0.5);
        if (task.bra == ket_idx) w *= Real(0.5);
        const double *__restrict__ P = args.P;
        double *__restrict__ J = args.J;
        double *__restrict__ K = args.K;
        const size_t N = (size_t)n;

        // J[a][b]: wave-uniform destination, reduce over lanes first
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                Real s = Real(0);
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) s += I[a * NB + b][c * ND + d] * (Real)P[(fc + c) * N + (fd + d)];
                double sd = wave_sum((double)(s * w));
                if (lane == 0) unsafeAtomicAdd(&J[(fa + a) * N + (fb + b)], 2.0 * sd);
            }
        if (!active) return;
        // J[c][d]
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                Real s = Real(0);
#pragma unroll
                for (int a = 0; a < NA; ++a)
#pragma unroll
                    for (int b = 0; b < NB; ++b) s += I[a * NB + b][c * ND + d] * (Real)P[(fa + a) * N + (fb + b)];
                unsafeAtomicAdd(&J[(fc + c) * N + (fd + d)], 2.0 * (double)(s * w));
            }
        // K[a][c] += sum_bd v P[b][d]
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                Real s = Real(0);
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int d = 0; d < ND; ++d) s += I[a * NB + b][c * ND + d] * (Real)P[(fb + b) * N + (fd + d)];
                unsafeAtomicAdd(&K[(fa + a) * N + (fc + c)], (double)(s * w));
            }
        // K[a][d] += sum_bc v P[b][c]
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                Real s = Real(0);
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int c = 0; c < NC; ++c) s += I[a * NB + b][c * ND + d] * (Real)P[(fb + b) * N + (fc + c)];
                unsafeAtomicAdd(&K[(fa + a) * N + (fd + d)], (double)(s * w));
            }
        // K[b][c] += sum_ad v P[a][d]
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                Real s = Real(0);
#pragma unroll
                for (int a = 0; a < NA; ++a)
#pragma unroll
                    for (int d = 0; d < ND; ++d) s += I[a * NB + b][c * ND + d] * (Real)P[(fa + a) * N + (fd + d)];
                unsafeAtomicAdd(&K[(fb + b) * N + (fc + c)], (double)(s * w));
            }
        // K[b][d] += sum_ac v P[a][c]
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                Real s = Real(0);
#pragma unroll
                for (int a = 0; a < NA; ++a)
#pragma unroll
                    for (int c = 0; c < NC; ++c) s += I[a * NB + b][c * ND + d] * (Real)P[(fa + a) * N + (fc + c)];
                unsafeAtomicAdd(&K[(fb + b) * N + (fd + d)], (double)(s * w));
            }
    }
}

// (ab|ab) for every shell pair of one class, one thread per pair: the diagonal the
// reference computes first (basis/src/electron_tensor.rs:89-101) and from which both its
// 1e-6 screen and the Schwarz bounds are taken. Writes D[x][y] = D[y][x] = (xy|xy).
template <int LA, int LB, int POLICY>
__global__ __launch_bounds__(256) void eri_diag_kernel(DiagArgs args) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= args.npairs) return;
    const PairInfo pr = args.pairs[args.first + t];
    const PrimPair<double> *__restrict__ pp = args.prims;
    double I[NAB][NAB];
#pragma unroll
    for (int x = 0; x < NAB; ++x)
#pragma unroll
        for (int y = 0; y < NAB; ++y) I[x][y] = 0.0;
    for (int i = 0; i < pr.K; ++i) {
        const PrimPair<double> b = pp[pr.pp_start + i];
        for (int k = 0; k < pr.K; ++k) {
            const PrimPair<double> kq = pp[pr.pp_start + k];
            accumulate_prim_quartet<LA, LB, LA, LB, POLICY, double>(b, kq, args.bt, I);
        }
    }
    const size_t N = (size_t)args.nbf;
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double v = I[a * NB + b][a * NB + b];
            args.D[(pr.fa + a) * N + (pr.fb + b)] = v;
            args.D[(pr.fb + b) * N + (pr.fa + a)] = v;
        }
}

// The same diagonal with one wave per pair: the lanes take the K^2 primitive quartets in turn and
// the diagonal components are summed over the wave. Accurate policy only; the reference-compatible
// policy keeps the serial order of eri_diag_kernel (its 1e-6 pair screen compares these values
// against a threshold).
template <int LA, int LB>
__global__ __launch_bounds__(256) void eri_diag_wave_kernel(DiagArgs args) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= args.npairs) return;  // wave-uniform
    const PairInfo pr = args.pairs[args.first + t];
    const PrimPair<double> *__restrict__ pp = args.prims + pr.pp_start;
    double I[NAB][NAB];
#pragma unroll
    for (int x = 0; x < NAB; ++x)
#pragma unroll
        for (int y = 0; y < NAB; ++y) I[x][y] = 0.0;
    const int KK = pr.K * pr.K;
    for (int e = lane; e < KK; e += 64) {
        const int i = e / pr.K, k = e - i * pr.K;
        accumulate_prim_quartet<LA, LB, LA, LB, 1, double>(pp[i], pp[k], args.bt, I);
    }
    const size_t N = (size_t)args.nbf;
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double v = wave_sum(I[a * NB + b][a * NB + b]);
            if (lane == 0) {
                args.D[(pr.fa + a) * N + (pr.fb + b)] = v;
                args.D[(pr.fb + b) * N + (pr.fa + a)] = v;
            }
        }
}

// q = sqrt(max_component (ij|ij)) of single primitive pairs (contraction coefficients
// included), the per-primitive Cauchy-Schwarz factor of the primitive screen.
template <int LA, int LB>
__global__ __launch_bounds__(256) void prim_diag_kernel(PrimDiagArgs args) {
    constexpr int NAB = ncart(LA) * ncart(LB);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= args.count) return;
    const int idx = args.prim_index[t];
    const PrimPair<double> pp = args.prims[idx];
    double I[NAB][NAB];
#pragma unroll
    for (int x = 0; x < NAB; ++x)
#pragma unroll
        for (int y = 0; y < NAB; ++y) I[x][y] = 0.0;
    accumulate_prim_quartet<LA, LB, LA, LB, 1, double>(pp, pp, args.bt, I);
    double m = 0.0;
#pragma unroll
    for (int x = 0; x < NAB; ++x) m = fmax(m, fabs(I[x][x]));
    args.q[idx] = (float)(sqrt(m) * (1.0 + 1e-6));  // rounded up: a bound stays a bound in float
}

// Number of primitive quartets the primitive screen keeps over a task list.
__global__ __launch_bounds__(256) void prim_count_kernel(PrimCountArgs args) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned long long c = 0, cs = 0;
    if (t < args.ntasks) {
        const Task task = args.tasks[t];
        // whole shell quartets below ket_tau are dropped by the kernels that apply it (FP32)
        bool kept = lane < task.ket_count;
        if (kept && args.ket_tau > 0.0f) kept = args.pairQ[task.bra] * args.pairQ[task.ket_begin + lane] >= args.ket_tau;
        if (kept) {
            cs = 1;
            const PairInfo bra = args.pairs[task.bra], ket = args.pairs[task.ket_begin + lane];
            for (int k = 0; k < ket.K; ++k) {
                const float qk = args.prim_q[ket.pp_start + k];
                for (int i = 0; i < bra.K; ++i) {
                    if (!(args.prim_q[bra.pp_start + i] * qk > args.prim_thr)) break;
                    ++c;
                }
            }
        }
    }
    c = wave_sum(c);
    cs = wave_sum(cs);
    if (lane == 0 && c) atomicAdd(args.out, c);
    if (lane == 0 && cs) atomicAdd(args.out_shell, cs);
}

// Blocked J/K digestion: one workgroup per block of tasks that share a bra pair (a, b). Every K
// contribution of such a task lands in row a or row b (K[a][c], K[a][d], K[b][c], K[b][d]), so the
// workgroup accumulates those rows in LDS with LDS atomics and flushes each nonzero entry once;
// J[a][b] is reduced over the workgroup; only J[c][d] is a per-lane global atomic.
template <int LA, int LB, int LC, int LD, int POLICY, bool REFSCREEN, typename Real, bool BATCH>
__global__ __launch_bounds__(256) void eri_block_kernel(JKArgs args_in) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    constexpr int L = LA + LB + LC + LD;
    constexpr int COLS = sizeof(Real) == 8 ? BOYS_PD_COLS64 : BOYS_PD_COLS32;
    constexpr int GN = sizeof(Real) == 8 ? BOYS_GRID_N : BOYS_PD32_N;  // rows of the Boys slab
    // the slab is staged in LDS for the FP32 kernels and FP64 (ss|ss); the other FP64 classes read
    // the Boys table from global memory (their 33 KB slab cost more occupancy than the L2 reads)
    constexpr bool SLAB = POLICY == 1 && block_slab_in_lds(L, sizeof(Real) == 4);
    // LDS: bra primitive pairs | their q | Boys slab | K rows a, b | J_ab partials of the 4 waves |
    // the block's tasks | its kept ket pairs
    extern __shared__ __align__(32) unsigned char lds_bytes[];
    JKArgs args = args_in;
    Block blk_;
    if constexpr (BATCH) {
        // batched launch: this block belongs to one molecule's record (record -1: a padding slot
        // of an XCD-affine schedule, no work)
        const int2 bb = args_in.bblocks[blockIdx.x];
        if (bb.x < 0) return;  // the same for every thread of the block
        const BatchRec r = args_in.batch[bb.x];
        if (!r.active) return;  // the same for every thread of the block
        // device-driven iteration: the molecule still iterates, in this kernel's precision
        if (r.state && (!r.state->active || r.state->phase != (sizeof(Real) == 4 ? 0 : 1))) return;
        args.tasks = r.tasks;
        args.pairs = r.pairs;
        args.prims = sizeof(Real) == 4 ? r.prims32 : r.prims;
        args.prim_q = r.prim_q;
        args.soa = sizeof(Real) == 4 ? r.soa32 : r.soa;
        if constexpr (sizeof(Real) == 4) args.pairQ = r.pairQ;  // the shell-quartet cut (ket_tau)
        args.soa_pairs = r.soa_pairs;
        args.soa_stride = r.soa_stride;
        args.D = r.D;
        args.P = r.P;
        args.J = r.J;
        args.K = r.K;
        args.Jk = r.Jk;
        args.npair = r.npair;
        args.nbf = r.nbf;
        blk_ = r.blocks[bb.y];
    } else {
        blk_ = args.blocks[blockIdx.x];
        if (args.block_qmax) {  // density screening: the whole block below the bound returns
            const float ub = args.pairQ[blk_.bra] * args.block_qmax[blockIdx.x] * (2.0f * *args.pm_max);
            if (!(ub >= args.bound_lo)) return;  // the same for every thread of the block
        }
    }
    const Block blk = blk_;
    const int n = args.nbf;  // LDS is sized for the largest basis of the launch
    // the class's Boys slab: static LDS ahead of the dynamic part (32-bit LDS addressing)
    __shared__ Real slice[SLAB ? GN * COLS : 1];  // [GN][COLS]
    // FP32: the slab's loads are issued here, into registers (7 per thread), and stored to LDS after
    // the ket screen, so their latency overlaps the task and screen loads
    constexpr bool SLAB_EARLY = SLAB && sizeof(Real) == 4;
    constexpr int SLAB_PER = SLAB_EARLY ? (GN * COLS + 255) / 256 : 1;
    const Real *slab_src = nullptr;
    if constexpr (SLAB) {
        if constexpr (sizeof(Real) == 8) slab_src = reinterpret_cast<const Real *>(args.bt.pd64);
        else slab_src = reinterpret_cast<const Real *>(args.bt.pd32);
        slab_src += (size_t)L * GN * COLS;
    }
    Real slab_reg[SLAB_PER];
    if constexpr (SLAB_EARLY) {
#pragma unroll
        for (int j = 0; j < SLAB_PER; ++j) {
            const int e = (int)threadIdx.x + 256 * j;
            slab_reg[j] = e < GN * COLS ? slab_src[e] : Real(0);
        }
    }
    double *lds = reinterpret_cast<double *>(lds_bytes);
    double *Ka = lds;                 // [NA][n]
    double *Kb = lds + NA * n;        // [NB][n]
    double *red = Kb + NB * n;        // [4][NAB]
    const int tile = args_in.tile;  // most tasks per block of this launch (blk.ntasks <= tile)
    Task *s_task = reinterpret_cast<Task *>(red + 4 * NAB);         // [tile]
    int *s_kidx = reinterpret_cast<int *>(s_task + tile);  // [tile * 64] kept ket pair indices
    unsigned long long *s_mask = reinterpret_cast<unsigned long long *>(s_kidx + tile * 64);  // [tile]
    // density screening: the shell-block maxima of the bra shells' rows (the lane bound reads them)
    float *s_pma = reinterpret_cast<float *>(s_mask + tile);  // [nshell]
    float *s_pmb = s_pma + args.nshell;                        // [nshell]
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const PairInfo bra = args.pairs[blk.bra];
    const size_t N = (size_t)n;
    const double *__restrict__ P = args.P;
    const int fa = bra.fa, fb = bra.fb;
    // the bra's density block, loaded as soon as the bra is known (used from the first ket on)
    Real pab[NAB];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) pab[a * NB + b] = (Real)P[(fa + a) * N + (fb + b)];
    if (threadIdx.x < blk.ntasks) {
        const Task tk = args.tasks[blk.first_task + threadIdx.x];
        s_task[threadIdx.x] = tk;
    }
    float pm_ab = 0.0f, q_ab = 0.0f;
    if (args.pair_shell) {
        const int2 ab = args.pair_shell[blk.bra];
        for (int e = threadIdx.x; e < args.nshell; e += blockDim.x) {
            s_pma[e] = args.Pm[(size_t)ab.x * args.nshell + e];
            s_pmb[e] = args.Pm[(size_t)ab.y * args.nshell + e];
        }
        pm_ab = args.Pm[(size_t)ab.x * args.nshell + ab.y];
        q_ab = args.pairQ[blk.bra];
    }
    __syncthreads();
    // the block's kept kets, packed densely in task order: wave w owns task chunks w, w + 4 (a
    // chunk is one task's 64 ket slots) and marks the kets it keeps (all of them, or those passing
    // the density-weighted screen of an incremental build)
    for (int t = wave; t < blk.ntasks; t += 4) {
        const Task tk = s_task[t];
        bool keep = lane < tk.ket_count;
        if (args.pair_shell && keep) {
            // Q_ab Q_cd max(2|P_ab|, 2|P_cd|, |P_ac|, |P_ad|, |P_bc|, |P_bd|) over shell blocks: the
            // ket's own block by pair (coalesced), the bra rows from LDS
            const int ket_idx = tk.ket_begin + lane;
            const int2 cd = args.pair_shell[ket_idx];
            float pm = 2.0f * fmaxf(pm_ab, args.Pm_pair[ket_idx]);
            pm = fmaxf(pm, fmaxf(s_pma[cd.x], s_pma[cd.y]));
            pm = fmaxf(pm, fmaxf(s_pmb[cd.x], s_pmb[cd.y]));
            const float bound = q_ab * args.pairQ[ket_idx] * pm;
            keep = bound >= args.bound_lo && bound < args.bound_hi;
        } else if (args.ket_tau > 0.0f && keep) {
            keep = args.pairQ[blk.bra] * args.pairQ[tk.ket_begin + lane] >= args.ket_tau;
        }
        const unsigned long long m = __ballot(keep);
        if (lane == 0) s_mask[t] = m;
    }
    __syncthreads();
    int kept = 0;
    for (int t = 0; t < blk.ntasks; ++t) kept += __popcll(s_mask[t]);
    if (args.kept_count && threadIdx.x == 0 && kept > 0) atomicAdd(args.kept_count, (unsigned long long)kept);
    if (kept == 0) return;  // the same for every thread of the block
    // every kept ket is written to its rank among the block's kept kets; the K rows, the bra
    // primitive pairs with their q and this class's Boys slab are staged
    for (int t = wave; t < blk.ntasks; t += 4) {
        const unsigned long long m = s_mask[t];
        if ((m >> lane) & 1ull) {
            int r = __popcll(m & ((1ull << lane) - 1ull));
            for (int u = 0; u < t; ++u) r += __popcll(s_mask[u]);
            s_kidx[r] = s_task[t].ket_begin + lane;
        }
    }
    for (int e = threadIdx.x; e < (NA + NB) * n; e += blockDim.x) lds[e] = 0.0;
    if constexpr (SLAB_EARLY) {
#pragma unroll
        for (int j = 0; j < SLAB_PER; ++j) {
            const int e = (int)threadIdx.x + 256 * j;
            if (e < GN * COLS) slice[e] = slab_reg[j];
        }
    } else if constexpr (SLAB) {
        for (int e = threadIdx.x; e < GN * COLS; e += blockDim.x) slice[e] = slab_src[e];
    }
    __syncthreads();
    double *__restrict__ J = args.J;
    double jab[NAB];
#pragma unroll
    for (int x = 0; x < NAB; ++x) jab[x] = 0.0;

    // groups of 64 kept kets, one lane each
    constexpr bool P_EARLY = NCD + NB * ND + NB * NC + NA * ND + NA * NC <= 16;
    for (int g = wave * 64; g < kept; g += 256) {
        const bool active = g + lane < kept;
        const int slot = active ? g + lane : g;
        // the ket pair's SoA start and bucket size from global memory (consecutive lanes hold
        // mostly consecutive pairs: coalesced, L2-resident), so LDS keeps 4 bytes per kept ket
        const int ket_idx = s_kidx[slot];
        const PairInfo ket = args.soa_pairs[ket_idx];
        const int ket_stride = args.soa_stride[ket_idx];
        const int fc = ket.fa, fd = ket.fb;
        // the density elements the digestion reads: for the small classes loaded before the
        // integrals (their latency hides behind the primitive loop), else where they are used
        Real p_cd[P_EARLY ? NCD : 1], p_bd[P_EARLY ? NB * ND : 1], p_bc[P_EARLY ? NB * NC : 1];
        Real p_ad[P_EARLY ? NA * ND : 1], p_ac[P_EARLY ? NA * NC : 1];
        if constexpr (P_EARLY) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < ND; ++d) p_cd[c * ND + d] = (Real)P[(fc + c) * N + (fd + d)];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int d = 0; d < ND; ++d) p_bd[b * ND + d] = (Real)P[(fb + b) * N + (fd + d)];
#pragma unroll
                for (int c = 0; c < NC; ++c) p_bc[b * NC + c] = (Real)P[(fb + b) * N + (fc + c)];
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
#pragma unroll
                for (int d = 0; d < ND; ++d) p_ad[a * ND + d] = (Real)P[(fa + a) * N + (fd + d)];
#pragma unroll
                for (int c = 0; c < NC; ++c) p_ac[a * NC + c] = (Real)P[(fa + a) * N + (fc + c)];
            }
        }
        auto Pcd = [&](int c, int d) -> Real {
            if constexpr (P_EARLY) return p_cd[c * ND + d];
            else return (Real)P[(fc + c) * N + (fd + d)];
        };
        auto Pbd = [&](int b, int d) -> Real {
            if constexpr (P_EARLY) return p_bd[b * ND + d];
            else return (Real)P[(fb + b) * N + (fd + d)];
        };
        auto Pbc = [&](int b, int c) -> Real {
            if constexpr (P_EARLY) return p_bc[b * NC + c];
            else return (Real)P[(fb + b) * N + (fc + c)];
        };
        auto Pad = [&](int a, int d) -> Real {
            if constexpr (P_EARLY) return p_ad[a * ND + d];
            else return (Real)P[(fa + a) * N + (fd + d)];
        };
        auto Pac = [&](int a, int c) -> Real {
            if constexpr (P_EARLY) return p_ac[a * NC + c];
            else return (Real)P[(fa + a) * N + (fc + c)];
        };
        Real I[NAB][NCD];
        lane_integrals_lds<LA, LB, LC, LD, POLICY, REFSCREEN, Real, SLAB>(args, bra, ket, ket_stride, active,
                                                                          (LdsPtr<Real>)slice, I);
        Real w = Real(1);
        if (fa == fb) w *= Real(0.5);
        if (fc == fd) w *= Real(0.5);
        if (blk.bra == ket_idx) w *= Real(0.5);
        // J[a][b] partial (lane-local; reduced after the loop)
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                Real s = Real(0);
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) s += I[a * NB + b][c * ND + d] * Pcd(c, d);
                jab[a * NB + b] += (double)(s * w);
            }
        if (!active) continue;
        // J[c][d]: pair-indexed accumulator (coalesced over the consecutive ket pairs of a wave)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                Real s = Real(0);
#pragma unroll
                for (int x = 0; x < NAB; ++x) s += I[x][c * ND + d] * pab[x];
                unsafeAtomicAdd(&args.Jk[(size_t)(c * ND + d) * args.npair + ket_idx], 2.0 * (double)(s * w));
            }
        // K rows a and b: LDS
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {  // K[a][c] += sum_bd v P[b][d]
                Real s = Real(0);
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int d = 0; d < ND; ++d) s += I[a * NB + b][c * ND + d] * Pbd(b, d);
                atomicAdd(&Ka[a * n + fc + c], (double)(s * w));
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) {  // K[a][d] += sum_bc v P[b][c]
                Real s = Real(0);
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int c = 0; c < NC; ++c) s += I[a * NB + b][c * ND + d] * Pbc(b, c);
                atomicAdd(&Ka[a * n + fd + d], (double)(s * w));
            }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {  // K[b][c] += sum_ad v P[a][d]
                Real s = Real(0);
#pragma unroll
                for (int a = 0; a < NA; ++a)
#pragma unroll
                    for (int d = 0; d < ND; ++d) s += I[a * NB + b][c * ND + d] * Pad(a, d);
                atomicAdd(&Kb[b * n + fc + c], (double)(s * w));
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) {  // K[b][d] += sum_ac v P[a][c]
                Real s = Real(0);
#pragma unroll
                for (int a = 0; a < NA; ++a)
#pragma unroll
                    for (int c = 0; c < NC; ++c) s += I[a * NB + b][c * ND + d] * Pac(a, c);
                atomicAdd(&Kb[b * n + fd + d], (double)(s * w));
            }
        }
    }
    // J[a][b]: wave reduction, then across the four waves
#pragma unroll
    for (int x = 0; x < NAB; ++x) {
        const double v = wave_sum(jab[x]);
        if (lane == 0) red[wave * NAB + x] = v;
    }
    __syncthreads();
    if (threadIdx.x < NAB) {
        const double v = red[threadIdx.x] + red[NAB + threadIdx.x] + red[2 * NAB + threadIdx.x] + red[3 * NAB + threadIdx.x];
        if (v != 0.0) unsafeAtomicAdd(&J[(fa + threadIdx.x / NB) * N + (fb + threadIdx.x % NB)], 2.0 * v);
    }
    // flush the K rows (nonzero entries only)
    double *__restrict__ K = args.K;
    for (int e = threadIdx.x; e < NA * n; e += blockDim.x) {
        const double v = Ka[e];
        if (v != 0.0) unsafeAtomicAdd(&K[(fa + e / n) * N + (e % n)], v);
    }
    for (int e = threadIdx.x; e < NB * n; e += blockDim.x) {
        const double v = Kb[e];
        if (v != 0.0) unsafeAtomicAdd(&K[(fb + e / n) * N + (e % n)], v);
    }
}

// ------------------------------------------------------------------ launchers

template <int LA, int LB, int LC, int LD, int POLICY, int MODE, bool REFSCREEN, typename Real>
static hipError_t launch_one(const JKArgs &a, hipStream_t s) {
    if (a.ntasks <= 0) return hipSuccess;
    dim3 grid((a.ntasks + 3) / 4), block(256);
    hipLaunchKernelGGL((eri_task_kernel<LA, LB, LC, LD, POLICY, MODE, REFSCREEN, Real>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD, int MODE>
static hipError_t dispatch_policy(int cls, const JKArgs &a, int policy, bool refscreen, bool fp32, hipStream_t s) {
    (void)cls;
    if (fp32) {
        if (policy != 1) return hipErrorInvalidValue;
        if constexpr (MODE == MODE_JK) {
            return refscreen ? launch_one<LA, LB, LC, LD, 1, MODE, true, float>(a, s)
                             : launch_one<LA, LB, LC, LD, 1, MODE, false, float>(a, s);
        } else {
            return hipErrorInvalidValue;
        }
    }
    if (policy == 0)
        return refscreen ? launch_one<LA, LB, LC, LD, 0, MODE, true, double>(a, s)
                         : launch_one<LA, LB, LC, LD, 0, MODE, false, double>(a, s);
    return refscreen ? launch_one<LA, LB, LC, LD, 1, MODE, true, double>(a, s)
                     : launch_one<LA, LB, LC, LD, 1, MODE, false, double>(a, s);
}

// LDS of one workgroup of the blocked digestion: K rows a and b, the J_ab reduction, the tile's
// task and kept-ket records, and the class's Boys slab
// (the dynamic part; the Boys slab is a static array of the kernel)
static size_t block_lds_dynamic(int cls, int nbf, int tile) {
    const int na = CLASS_BRA[cls] == 0 ? 1 : 3, nb = CLASS_BRA[cls] == 2 ? 3 : 1;
    return (size_t)(na + nb) * nbf * sizeof(double) + 4 * na * nb * sizeof(double) +
           (size_t)tile * (sizeof(Task) + 64 * sizeof(int) + sizeof(unsigned long long));
}
size_t block_lds_bytes(int cls, int nbf, int policy, bool fp32, int tile) {
    const size_t cols = fp32 ? BOYS_PD_COLS32 : BOYS_PD_COLS64, real = fp32 ? sizeof(float) : sizeof(double);
    const size_t rows = fp32 ? BOYS_PD32_N : BOYS_GRID_N;
    const bool slab = policy == 1 && block_slab_in_lds(CLASS_BRA[cls] + CLASS_KET[cls], fp32);
    return block_lds_dynamic(cls, nbf, tile) + (slab ? rows * cols * real : 0) + 32;  // + alignment of the dynamic part
}

template <int LA, int LB, int LC, int LD, int POLICY, bool REFSCREEN, typename Real, bool BATCH = false>
static hipError_t launch_block(const JKArgs &a, hipStream_t s) {
    if (a.nblocks <= 0) return hipSuccess;
    if (a.tile < 1 || a.tile > MAX_TILE) return hipErrorInvalidValue;
    constexpr int cls = (LA + LB == 0) ? 0 : (LA + LB == 1) ? (LC + LD == 0 ? 1 : 2) : (LC + LD == 0 ? 3 : LC + LD == 1 ? 4 : 5);
    const size_t lds = block_lds_dynamic(cls, a.nbf, a.tile) + (a.pair_shell ? 2 * (size_t)a.nshell * sizeof(float) : 0);
    hipLaunchKernelGGL((eri_block_kernel<LA, LB, LC, LD, POLICY, REFSCREEN, Real, BATCH>), dim3(a.nblocks), dim3(256),
                       lds, s, a);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
static hipError_t dispatch_block_policy(const JKArgs &a, int policy, bool refscreen, bool fp32, hipStream_t s) {
    if (fp32) {
        if (policy != 1) return hipErrorInvalidValue;
        if (a.batch) {  // mixed-form batches: accurate policy, Schwarz screen or none
            if (refscreen) return hipErrorInvalidValue;
            return launch_block<LA, LB, LC, LD, 1, false, float, true>(a, s);
        }
        return refscreen ? launch_block<LA, LB, LC, LD, 1, true, float>(a, s)
                         : launch_block<LA, LB, LC, LD, 1, false, float>(a, s);
    }
    if (a.batch) {  // a separate instantiation keeps the one-molecule kernels' registers as they are
        if (policy == 0)
            return refscreen ? launch_block<LA, LB, LC, LD, 0, true, double, true>(a, s)
                             : launch_block<LA, LB, LC, LD, 0, false, double, true>(a, s);
        return refscreen ? launch_block<LA, LB, LC, LD, 1, true, double, true>(a, s)
                         : launch_block<LA, LB, LC, LD, 1, false, double, true>(a, s);
    }
    if (policy == 0)
        return refscreen ? launch_block<LA, LB, LC, LD, 0, true, double>(a, s)
                         : launch_block<LA, LB, LC, LD, 0, false, double>(a, s);
    return refscreen ? launch_block<LA, LB, LC, LD, 1, true, double>(a, s)
                     : launch_block<LA, LB, LC, LD, 1, false, double>(a, s);
}

hipError_t launch_eri_blocks(int cls, const JKArgs &a, int policy, bool refscreen, bool fp32, hipStream_t s) {
    switch (cls) {
        case 0: return dispatch_block_policy<0, 0, 0, 0>(a, policy, refscreen, fp32, s);
        case 1: return dispatch_block_policy<1, 0, 0, 0>(a, policy, refscreen, fp32, s);
        case 2: return dispatch_block_policy<1, 0, 1, 0>(a, policy, refscreen, fp32, s);
        case 3: return dispatch_block_policy<1, 1, 0, 0>(a, policy, refscreen, fp32, s);
        case 4: return dispatch_block_policy<1, 1, 1, 0>(a, policy, refscreen, fp32, s);
        case 5: return dispatch_block_policy<1, 1, 1, 1>(a, policy, refscreen, fp32, s);
        default: return hipErrorInvalidValue;
    }
}

template <int MODE>
static hipError_t dispatch_class(int cls, const JKArgs &a, int policy, bool refscreen, bool fp32, hipStream_t s) {
    switch (cls) {
        case 0: return dispatch_policy<0, 0, 0, 0, MODE>(cls, a, policy, refscreen, fp32, s);
        case 1: return dispatch_policy<1, 0, 0, 0, MODE>(cls, a, policy, refscreen, fp32, s);
        case 2: return dispatch_policy<1, 0, 1, 0, MODE>(cls, a, policy, refscreen, fp32, s);
        case 3: return dispatch_policy<1, 1, 0, 0, MODE>(cls, a, policy, refscreen, fp32, s);
        case 4: return dispatch_policy<1, 1, 1, 0, MODE>(cls, a, policy, refscreen, fp32, s);
        case 5: return dispatch_policy<1, 1, 1, 1, MODE>(cls, a, policy, refscreen, fp32, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_eri_tasks(int cls, int mode, const JKArgs &a, int policy, bool refscreen, bool fp32,
                            hipStream_t s) {
    if (mode == MODE_DENSE) return dispatch_class<MODE_DENSE>(cls, a, policy, refscreen, fp32, s);
    return dispatch_class<MODE_JK>(cls, a, policy, refscreen, fp32, s);
}

hipError_t launch_prim_diag(int pair_class, const PrimDiagArgs &a, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    dim3 grid((a.count + 255) / 256), block(256);
    switch (pair_class) {
        case 0: hipLaunchKernelGGL((prim_diag_kernel<0, 0>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((prim_diag_kernel<1, 0>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((prim_diag_kernel<1, 1>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_prim_count(const PrimCountArgs &a, hipStream_t s) {
    if (a.ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(prim_count_kernel, dim3((a.ntasks + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_eri_diag(int pair_class, const DiagArgs &a, int policy, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    if (policy) {  // accurate: one wave per pair
        dim3 wgrid((a.npairs + 3) / 4), wblock(256);
        switch (pair_class) {
            case 0: hipLaunchKernelGGL((eri_diag_wave_kernel<0, 0>), wgrid, wblock, 0, s, a); break;
            case 1: hipLaunchKernelGGL((eri_diag_wave_kernel<1, 0>), wgrid, wblock, 0, s, a); break;
            case 2: hipLaunchKernelGGL((eri_diag_wave_kernel<1, 1>), wgrid, wblock, 0, s, a); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    dim3 grid((a.npairs + 255) / 256), block(256);  // reference-compatible: one thread per pair
    switch (pair_class) {
        case 0: hipLaunchKernelGGL((eri_diag_kernel<0, 0, 0>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((eri_diag_kernel<1, 0, 0>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((eri_diag_kernel<1, 1, 0>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fock
