// gradient.hip -- analytic RHF nuclear gradient (accurate integrals, FP64).
//
//   dE/dR = sum P dH - sum W dS + sum_{unique abcd} 4 w Gamma_abcd d(ab|cd) + dE_nuc
//   Gamma_abcd = P_ab P_cd - (P_ac P_bd + P_ad P_bc) / 4,   W = 2 sum_occ eps_i C_i C_i^T
//
// with w the unique-quartet weight of the Fock build ((a==b ? 1/2) (c==d ? 1/2) (ab==cd ? 1/2)).
// Energy expression: scf/src/lib.rs:223 at convergence; integrals: the reference's Cartesian
// Gaussians (A - r)^a exp(-alpha |r - A|^2) (basis/src/primitives.rs, contracted_gaussian.rs), whose
// centre derivative is
//   d/dA_x phi_a = a_x phi_{a - 1x} - 2 alpha phi_{a + 1x},
// so every derivative integral is two integrals of shifted angular momentum, evaluated with the
// McMurchie-Davidson E coefficients of one more unit of angular momentum on the differentiated
// centre and Hermite R of one more order. Translational invariance gives the last centre:
// d/dD = -(d/dA + d/dB + d/dC) for the ERIs, d/dB = -d/dA for S and T, d/dC = -(d/dA + d/dB) for
// the attraction to nucleus C.
#include "fock_device.hpp"
#include "kernels.hpp"

namespace fock {

// One primitive quartet: the nine derivatives (A, B, C; x, y, z) of every component integral,
// contracted with fac[ab][cd] (the pair-density weight of each component) into g[9].
template <int LA, int LB, int LC, int LD, typename Real>
__device__ __forceinline__ void accumulate_grad_quartet(const PrimPair<Real> &bra, const PrimPair<Real> &ket,
                                                        Real alpha_a, Real alpha_b, Real alpha_c,
                                                        const BoysTables &bt,
                                                        const double fac[ncart(LA) * ncart(LB)][ncart(LC) * ncart(LD)],
                                                        double g[9]) {
    constexpr int L1 = LA + LB + LC + LD + 1;
    constexpr int LAB = LA + LB, LCD = LC + LD;
    constexpr int NB = ncart(LB), ND = ncart(LD), NC = ncart(LC), NA = ncart(LA);
    const Real p = bra.p, q = ket.p;
    Real rs;
    if constexpr (sizeof(Real) == 8) rs = rsqrt(p + q);
    else rs = rsqrtf(p + q);
    const Real alpha = p * q * (rs * rs);
    const Real X = ket.Px - bra.Px, Y = ket.Py - bra.Py, Z = ket.Pz - bra.Pz;  // Q - P
    const Real T = alpha * (X * X + Y * Y + Z * Z);
    const Real pref = Real(34.98683665524972) * bra.Kc * ket.Kc * rs;  // 2 pi^{5/2}/(p q sqrt(p+q)) K K
    Real Fs[L1 + 1];
    if constexpr (sizeof(Real) == 8) boys_taylor64<L1>(T, bt.tab64, Fs);
    else boys_taylor32<L1>(T, bt.tab32, Fs);
    {
        const Real m2a = Real(-2) * alpha;
        Real f = m2a;
#pragma unroll
        for (int n = 1; n <= L1; ++n) {
            Fs[n] *= f;
            f *= m2a;
        }
    }
    Real R[L1 + 1][L1 + 1][L1 + 1];
    hermite_coulomb<L1, false, Real>(Fs, X, Y, Z, R);
    Real Eab[3][LA + 2][LB + 2][LAB + 3];
    Real Ecd[3][LC + 2][LD + 1][LCD + 2];
    hermite_1d<LA + 1, LB + 1, Real>(bra.xax, bra.xbx, bra.inv2p, Eab[0]);
    hermite_1d<LA + 1, LB + 1, Real>(bra.xay, bra.xby, bra.inv2p, Eab[1]);
    hermite_1d<LA + 1, LB + 1, Real>(bra.xaz, bra.xbz, bra.inv2p, Eab[2]);
    hermite_1d<LC + 1, LD, Real>(ket.xax, ket.xbx, ket.inv2p, Ecd[0]);
    hermite_1d<LC + 1, LD, Real>(ket.xay, ket.xby, ket.inv2p, Ecd[1]);
    hermite_1d<LC + 1, LD, Real>(ket.xaz, ket.xbz, ket.inv2p, Ecd[2]);

    static_for<0, NC>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        constexpr int cx = cart(LC, c, 0), cy = cart(LC, c, 1), cz = cart(LC, c, 2);
        static_for<0, ND>([&](auto d_) {
            constexpr int d = decltype(d_)::value;
            constexpr int dx = cart(LD, d, 0), dy = cart(LD, d, 1), dz = cart(LD, d, 2);
            // ket contractions over the bra Hermite simplex: W0 with the ket as it is (bra Hermite
            // order up to LAB + 1 for the shifted bra), WC[axis] with the ket's C derivative
            Real W0[LAB + 2][LAB + 2][LAB + 2];
            Real WC[3][LAB + 1][LAB + 1][LAB + 1];
            static_for<0, LAB + 2>([&](auto t_) {
                constexpr int t = decltype(t_)::value;
                static_for<0, LAB + 2 - t>([&](auto u_) {
                    constexpr int u = decltype(u_)::value;
                    static_for<0, LAB + 2 - t - u>([&](auto v_) {
                        constexpr int v = decltype(v_)::value;
                        Real acc = Real(0);
                        static_for<0, cx + dx + 1>([&](auto tt_) {
                            constexpr int tt = decltype(tt_)::value;
                            static_for<0, cy + dy + 1>([&](auto uu_) {
                                constexpr int uu = decltype(uu_)::value;
                                static_for<0, cz + dz + 1>([&](auto vv_) {
                                    constexpr int vv = decltype(vv_)::value;
                                    const Real e = Ecd[0][cx][dx][tt] * Ecd[1][cy][dy][uu] * Ecd[2][cz][dz][vv];
                                    if constexpr ((tt + uu + vv) & 1) acc -= e * R[t + tt][u + uu][v + vv];
                                    else acc += e * R[t + tt][u + uu][v + vv];
                                });
                            });
                        });
                        W0[t][u][v] = acc;
                        if constexpr (t + u + v <= LAB) {
                            // d/dC along axis k: E_k -> c_k E^{c-1} - 2 gamma E^{c+1}
                            static_for<0, 3>([&](auto k_) {
                                constexpr int k = decltype(k_)::value;
                                constexpr int ck = k == 0 ? cx : (k == 1 ? cy : cz);
                                constexpr int dk = k == 0 ? dx : (k == 1 ? dy : dz);
                                constexpr int nx = cx + dx + (k == 0), ny = cy + dy + (k == 1), nz = cz + dz + (k == 2);
                                Real accc = Real(0);
                                static_for<0, nx + 1>([&](auto tt_) {
                                    constexpr int tt = decltype(tt_)::value;
                                    static_for<0, ny + 1>([&](auto uu_) {
                                        constexpr int uu = decltype(uu_)::value;
                                        static_for<0, nz + 1>([&](auto vv_) {
                                            constexpr int vv = decltype(vv_)::value;
                                            Real ek = Real(0);
                                            constexpr int h = k == 0 ? tt : (k == 1 ? uu : vv);
                                            if constexpr (h <= ck + 1 + dk) ek = Real(-2) * alpha_c * Ecd[k][ck + 1][dk][h];
                                            if constexpr (ck >= 1 && h <= ck - 1 + dk) ek += Real(ck) * Ecd[k][ck - 1][dk][h];
                                            Real e = ek;
                                            if constexpr (k != 0) {
                                                if constexpr (tt <= cx + dx) e *= Ecd[0][cx][dx][tt];
                                                else e = Real(0);
                                            }
                                            if constexpr (k != 1) {
                                                if constexpr (uu <= cy + dy) e *= Ecd[1][cy][dy][uu];
                                                else e = Real(0);
                                            }
                                            if constexpr (k != 2) {
                                                if constexpr (vv <= cz + dz) e *= Ecd[2][cz][dz][vv];
                                                else e = Real(0);
                                            }
                                            if constexpr ((tt + uu + vv) & 1) accc -= e * R[t + tt][u + uu][v + vv];
                                            else accc += e * R[t + tt][u + uu][v + vv];
                                        });
                                    });
                                });
                                WC[k][t][u][v] = accc;
                            });
                        }
                    });
                });
            });
            static_for<0, NA>([&](auto a_) {
                constexpr int a = decltype(a_)::value;
                constexpr int ax = cart(LA, a, 0), ay = cart(LA, a, 1), az = cart(LA, a, 2);
                static_for<0, NB>([&](auto b_) {
                    constexpr int b = decltype(b_)::value;
                    constexpr int bx = cart(LB, b, 0), by = cart(LB, b, 1), bz = cart(LB, b, 2);
                    const double f = (double)pref * fac[a * NB + b][c * ND + d];
                    Real dA[3] = {Real(0), Real(0), Real(0)}, dB[3] = {Real(0), Real(0), Real(0)}, dC[3] = {Real(0), Real(0), Real(0)};
                    // d/dA, d/dB along axis k: the bra E of that axis shifted on a or b
                    static_for<0, 3>([&](auto k_) {
                        constexpr int k = decltype(k_)::value;
                        constexpr int ak = k == 0 ? ax : (k == 1 ? ay : az);
                        constexpr int bk = k == 0 ? bx : (k == 1 ? by : bz);
                        constexpr int mx = ax + bx + (k == 0), my = ay + by + (k == 1), mz = az + bz + (k == 2);
                        Real sa = Real(0), sb = Real(0);
                        static_for<0, mx + 1>([&](auto t_) {
                            constexpr int t = decltype(t_)::value;
                            static_for<0, my + 1>([&](auto u_) {
                                constexpr int u = decltype(u_)::value;
                                static_for<0, mz + 1>([&](auto v_) {
                                    constexpr int v = decltype(v_)::value;
                                    constexpr int h = k == 0 ? t : (k == 1 ? u : v);
                                    Real ea = Real(-2) * alpha_a * Eab[k][ak + 1][bk][h];
                                    Real eb = Real(-2) * alpha_b * Eab[k][ak][bk + 1][h];
                                    if constexpr (ak >= 1 && h <= ak - 1 + bk) ea += Real(ak) * Eab[k][ak - 1][bk][h];
                                    if constexpr (bk >= 1 && h <= ak + bk - 1) eb += Real(bk) * Eab[k][ak][bk - 1][h];
                                    Real rest = W0[t][u][v];
                                    if constexpr (k != 0) {
                                        if constexpr (t <= ax + bx) rest *= Eab[0][ax][bx][t];
                                        else rest = Real(0);
                                    }
                                    if constexpr (k != 1) {
                                        if constexpr (u <= ay + by) rest *= Eab[1][ay][by][u];
                                        else rest = Real(0);
                                    }
                                    if constexpr (k != 2) {
                                        if constexpr (v <= az + bz) rest *= Eab[2][az][bz][v];
                                        else rest = Real(0);
                                    }
                                    sa += ea * rest;
                                    sb += eb * rest;
                                });
                            });
                        });
                        dA[k] = sa;
                        dB[k] = sb;
                        Real sc = Real(0);
                        static_for<0, ax + bx + 1>([&](auto t_) {
                            constexpr int t = decltype(t_)::value;
                            static_for<0, ay + by + 1>([&](auto u_) {
                                constexpr int u = decltype(u_)::value;
                                static_for<0, az + bz + 1>([&](auto v_) {
                                    constexpr int v = decltype(v_)::value;
                                    sc += Eab[0][ax][bx][t] * Eab[1][ay][by][u] * Eab[2][az][bz][v] * WC[k][t][u][v];
                                });
                            });
                        });
                        dC[k] = sc;
                    });
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        g[k] += f * (double)dA[k];
                        g[3 + k] += f * (double)dB[k];
                        g[6 + k] += f * (double)dC[k];
                    }
                });
            });
        });
    });
}

// Two-electron part: one lane per unique shell quartet of the Fock build's task list; args.split
// consecutive waves share one task, wave s of them taking the primitive quartets s, s + split, ...
// (derivative integrals are linear in the primitive quartets).
template <int LA, int LB, int LC, int LD, typename Real>
__global__ __launch_bounds__(256) void grad_eri_kernel(GradEriArgs args) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NAB = NA * NB, NCD = NC * ND;
    extern __shared__ double sg[];  // [3 * natoms]
    for (int e = threadIdx.x; e < 3 * args.natoms; e += blockDim.x) sg[e] = 0.0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + wave, split = args.split;
    const int t = gw / split, part = gw - t * split;
    bool run = t < args.ntasks;
    Task task{};
    if (run) {
        task = args.tasks[t];
        if (args.pairQ) {
            // mixed form: the task's largest bound (its first ket) picks the precision pass
            const float bound = args.pairQ[task.bra] * args.pairQ[task.ket_begin];
            run = (sizeof(Real) == 8) == (bound >= args.tau64);
        }
    }
    if (run) {
        const PairInfo bra = args.pairs[task.bra];
        const int2 sab = args.pair_shell[task.bra];
        const int atA = args.shell_atom[sab.x], atB = args.shell_atom[sab.y];
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int atC = 0, atD = 0;
        if (lane < task.ket_count) {
            const int ket_idx = task.ket_begin + lane;
            const PairInfo ket = args.pairs[ket_idx];
            const int2 scd = args.pair_shell[ket_idx];
            atC = args.shell_atom[scd.x];
            atD = args.shell_atom[scd.y];
            const size_t N = (size_t)args.nbf;
            const double *__restrict__ P = args.P;
            double w = 4.0;
            if (sab.x == sab.y) w *= 0.5;
            if (scd.x == scd.y) w *= 0.5;
            if (task.bra == ket_idx) w *= 0.5;
            double fac[NAB][NCD];
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                        for (int d = 0; d < ND; ++d) {
                            const size_t ia = bra.fa + a, ib = bra.fb + b, ic = ket.fa + c, id = ket.fb + d;
                            const double gam = P[ia * N + ib] * P[ic * N + id] -
                                               0.25 * (P[ia * N + ic] * P[ib * N + id] + P[ia * N + id] * P[ib * N + ic]);
                            fac[a * NB + b][c * ND + d] = w * gam;
                        }
            const PrimPair<Real> *__restrict__ prims = reinterpret_cast<const PrimPair<Real> *>(
                sizeof(Real) == 8 ? (const void *)args.prims : (const void *)args.prims32);
            // primitive quartets ik = i * ket.K + k = part, part + split, ... (split = 1: all). The
            // classes of total angular momentum >= 3 compile to fewer registers as one flat loop
            // (measured on C8H18/6-31G: (pp|ps) 0.90 against 2.1 ms), the others as a walk that
            // loads each bra primitive once
            if constexpr (LA + LB + LC + LD >= 3) {
                const int nq = bra.K * ket.K;
                for (int ik = part; ik < nq; ik += split) {
                    const int i = ik / ket.K, k = ik - i * ket.K;
                    if (args.prim_q && args.prim_q[bra.pp_start + i] * args.prim_q[ket.pp_start + k] <= args.prim_thr)
                        continue;
                    const PrimPair<Real> bp = prims[bra.pp_start + i];
                    const double2 ab = args.prim_ab[bra.pp_start + i];
                    const PrimPair<Real> kp = prims[ket.pp_start + k];
                    const double2 cd = args.prim_ab[ket.pp_start + k];
                    accumulate_grad_quartet<LA, LB, LC, LD, Real>(bp, kp, (Real)ab.x, (Real)ab.y, (Real)cd.x,
                                                                  args.bt, fac, g);
                }
            } else {
                int i = ket.K > 0 ? part / ket.K : bra.K, k = ket.K > 0 ? part - i * ket.K : 0;
                while (i < bra.K) {
                    const PrimPair<Real> bp = prims[bra.pp_start + i];
                    const double2 ab = args.prim_ab[bra.pp_start + i];
                    const float qb = args.prim_q ? args.prim_q[bra.pp_start + i] : 1.0f;
                    for (; k < ket.K; k += split) {
                        if (args.prim_q && qb * args.prim_q[ket.pp_start + k] <= args.prim_thr) continue;
                        const PrimPair<Real> kp = prims[ket.pp_start + k];
                        const double2 cd = args.prim_ab[ket.pp_start + k];
                        accumulate_grad_quartet<LA, LB, LC, LD, Real>(bp, kp, (Real)ab.x, (Real)ab.y, (Real)cd.x,
                                                                      args.bt, fac, g);
                    }
                    k -= ket.K;
                    ++i;
                    while (k >= ket.K) {
                        k -= ket.K;
                        ++i;
                    }
                }
            }
        }
        // bra centres are the same for the whole wave: one reduction each
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double v = wave_sum(g[k]);
            if (lane == 0 && v != 0.0) atomicAdd(&sg[3 * (k < 3 ? atA : atB) + k % 3], v);
        }
        if (lane < task.ket_count) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicAdd(&sg[3 * atC + k], g[6 + k]);
                atomicAdd(&sg[3 * atD + k], -(g[k] + g[3 + k] + g[6 + k]));
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * args.natoms; e += blockDim.x)
        if (sg[e] != 0.0) unsafeAtomicAdd(&args.grad[e], sg[e]);
}

// One-electron part: one thread per (shell pair (a, b), term), all component pairs. Threads
// [0, npairs * natoms) take the attraction to nucleus C = t % natoms of pair t / natoms (d/dC =
// -(d/dA + d/dB)); the last npairs threads take overlap and kinetic of one pair each (d/dB = -d/dA).
// Neighbouring lanes share a pair (same primitive loop bounds) and a wave takes one kind of term.
template <int LA, int LB>
__global__ __launch_bounds__(128) void grad_one_kernel(GradOneArgs args) {
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr int LT = LA + LB + 1;
    extern __shared__ double sg[];  // [3 * natoms]
    for (int e = threadIdx.x; e < 3 * args.natoms; e += blockDim.x) sg[e] = 0.0;
    __syncthreads();
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nv = (long)args.npairs * args.natoms;
    if (t < nv + args.npairs) {
        const int pr = t < nv ? (int)(t / args.natoms) : (int)(t - nv);
        const int c = t < nv ? (int)(t % args.natoms) : -1;  // -1: overlap and kinetic
        const int sa = args.pair_a[pr], sb = args.pair_b[pr];
        const int fa = args.shell_f0[sa], fb = args.shell_f0[sb];
        const int atA = args.shell_atom[sa], atB = args.shell_atom[sb];
        const double A[3] = {args.centers[3 * fa], args.centers[3 * fa + 1], args.centers[3 * fa + 2]};
        const double B[3] = {args.centers[3 * fb], args.centers[3 * fb + 1], args.centers[3 * fb + 2]};
        const double dist[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        const size_t N = (size_t)args.nbf;
        const double wpair = sa == sb ? 1.0 : 2.0;  // P, W symmetric: (a, b) and (b, a)
        double Pab[NA][NB], Wab[NA][NB];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                Pab[i][j] = wpair * args.P[(size_t)(fa + i) * N + (fb + j)];
                Wab[i][j] = wpair * args.W[(size_t)(fa + i) * N + (fb + j)];
            }
        const int ka0 = args.prim_off[fa], ka1 = args.prim_off[fa + 1];
        const int kb0 = args.prim_off[fb], kb1 = args.prim_off[fb + 1];
        double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
        {
            double gCn[3] = {0, 0, 0};
            for (int ia = ka0; ia < ka1; ++ia) {
                const double a = args.exps[ia];
                for (int ib = kb0; ib < kb1; ++ib) {
                    const double b = args.exps[ib];
                    const double p = a + b, q = a * b / p, inv2p = 1.0 / (2.0 * p);
                    const double cc = args.coefs[ia] * args.coefs[ib];
                    double E[3][LA + 2][LB + 3][LA + LB + 4];
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        hermite_1d<LA + 1, LB + 2, double>(-(q * dist[x] / a), q * dist[x] / b, inv2p, E[x]);
                        const double ex = exp(-q * (dist[x] * dist[x]));
#pragma unroll
                        for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
                            for (int j = 0; j <= LB + 2; ++j)
#pragma unroll
                                for (int tt = 0; tt <= LA + LB + 3; ++tt) E[x][i][j][tt] *= ex;
                    }
                    if (c < 0) {
                        const double norm = sqrt((M_PI / p) * (M_PI / p) * (M_PI / p));
                        // overlap of Cartesian exponents (i, j) per axis, kinetic per the reference
                        auto S3 = [&](int ix, int iy, int iz, int jx, int jy, int jz) -> double {
                            return E[0][ix][jx][0] * E[1][iy][jy][0] * E[2][iz][jz][0] * norm;
                        };
                        auto T3 = [&](int ix, int iy, int iz, int jx, int jy, int jz) -> double {
                            double v = b * (double)(2 * (jx + jy + jz) + 3) * S3(ix, iy, iz, jx, jy, jz);
                            v -= 2.0 * b * b *
                                 (S3(ix, iy, iz, jx + 2, jy, jz) + S3(ix, iy, iz, jx, jy + 2, jz) + S3(ix, iy, iz, jx, jy, jz + 2));
                            if (jx >= 2) v -= 0.5 * (double)(jx * (jx - 1)) * S3(ix, iy, iz, jx - 2, jy, jz);
                            if (jy >= 2) v -= 0.5 * (double)(jy * (jy - 1)) * S3(ix, iy, iz, jx, jy - 2, jz);
                            if (jz >= 2) v -= 0.5 * (double)(jz * (jz - 1)) * S3(ix, iy, iz, jx, jy, jz - 2);
                            return v;
                        };
#pragma unroll
                        for (int i = 0; i < NA; ++i)
#pragma unroll
                            for (int j = 0; j < NB; ++j) {
                                const int ax = cart(LA, i, 0), ay = cart(LA, i, 1), az = cart(LA, i, 2);
                                const int bx = cart(LB, j, 0), by = cart(LB, j, 1), bz = cart(LB, j, 2);
#pragma unroll
                                for (int k = 0; k < 3; ++k) {
                                    const int ak = k == 0 ? ax : (k == 1 ? ay : az);
                                    const int up[3] = {ax + (k == 0), ay + (k == 1), az + (k == 2)};
                                    const int dn[3] = {ax - (k == 0), ay - (k == 1), az - (k == 2)};
                                    double dS = -2.0 * a * S3(up[0], up[1], up[2], bx, by, bz);
                                    double dT = -2.0 * a * T3(up[0], up[1], up[2], bx, by, bz);
                                    if (ak >= 1) {
                                        dS += (double)ak * S3(dn[0], dn[1], dn[2], bx, by, bz);
                                        dT += (double)ak * T3(dn[0], dn[1], dn[2], bx, by, bz);
                                    }
                                    const double gv = cc * (Pab[i][j] * dT - Wab[i][j] * dS);
                                    gA[k] += gv;
                                    gB[k] -= gv;
                                }
                            }
                    } else {
                        double Pc[3];
#pragma unroll
                        for (int x = 0; x < 3; ++x) Pc[x] = (a * A[x] + b * B[x]) / p;
                        const double X = args.nuc_pos[3 * c] - Pc[0], Y = args.nuc_pos[3 * c + 1] - Pc[1],
                                     Z = args.nuc_pos[3 * c + 2] - Pc[2];
                        double Fs[LT + 1];
                        boys_taylor64<LT>(p * (X * X + Y * Y + Z * Z), args.bt.tab64, Fs);
                        double f = -2.0 * p;
#pragma unroll
                        for (int n = 1; n <= LT; ++n) {
                            Fs[n] *= f;
                            f *= -2.0 * p;
                        }
                        double R[LT + 1][LT + 1][LT + 1];
                        hermite_coulomb<LT, false, double>(Fs, X, Y, Z, R);
                        const double wgt = -args.nuc_charge[c] * (2.0 * M_PI) / p * cc;
                        auto V3 = [&](int ix, int iy, int iz, int jx, int jy, int jz) -> double {
                            double v = 0.0;
                            for (int tt = 0; tt <= ix + jx; ++tt)
                                for (int uu = 0; uu <= iy + jy; ++uu)
                                    for (int vv = 0; vv <= iz + jz; ++vv)
                                        v += E[0][ix][jx][tt] * E[1][iy][jy][uu] * E[2][iz][jz][vv] * R[tt][uu][vv];
                            return v;
                        };
#pragma unroll
                        for (int i = 0; i < NA; ++i)
#pragma unroll
                            for (int j = 0; j < NB; ++j) {
                                const int ax = cart(LA, i, 0), ay = cart(LA, i, 1), az = cart(LA, i, 2);
                                const int bx = cart(LB, j, 0), by = cart(LB, j, 1), bz = cart(LB, j, 2);
#pragma unroll
                                for (int k = 0; k < 3; ++k) {
                                    const int ak = k == 0 ? ax : (k == 1 ? ay : az);
                                    const int bk = k == 0 ? bx : (k == 1 ? by : bz);
                                    double dVa = -2.0 * a * V3(ax + (k == 0), ay + (k == 1), az + (k == 2), bx, by, bz);
                                    double dVb = -2.0 * b * V3(ax, ay, az, bx + (k == 0), by + (k == 1), bz + (k == 2));
                                    if (ak >= 1) dVa += (double)ak * V3(ax - (k == 0), ay - (k == 1), az - (k == 2), bx, by, bz);
                                    if (bk >= 1) dVb += (double)bk * V3(ax, ay, az, bx - (k == 0), by - (k == 1), bz - (k == 2));
                                    const double pa = Pab[i][j] * wgt;
                                    gA[k] += pa * dVa;
                                    gB[k] += pa * dVb;
                                    gCn[k] -= pa * (dVa + dVb);
                                }
                            }
                    }
                }
            }
            if (c >= 0) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (gCn[k] != 0.0) atomicAdd(&sg[3 * c + k], gCn[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            atomicAdd(&sg[3 * atA + k], gA[k]);
            atomicAdd(&sg[3 * atB + k], gB[k]);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * args.natoms; e += blockDim.x)
        if (sg[e] != 0.0) unsafeAtomicAdd(&args.grad[e], sg[e]);
}

// W = 2 sum_{k < nocc} eps_k C[:, k] C[:, k]^T from column-major C (row-major W, symmetric).
__global__ void energy_weighted_density_kernel(int n, int nocc, const double *__restrict__ C,
                                               const double *__restrict__ eps, double *__restrict__ W) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    double s = 0.0;
    for (int k = 0; k < nocc; ++k) s += eps[k] * C[(size_t)k * n + i] * C[(size_t)k * n + j];
    W[e] = 2.0 * s;
}

hipError_t energy_weighted_density(int n, int nocc, const double *C, const double *eps, double *W, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(energy_weighted_density_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, nocc,
                       C, eps, W);
    return hipGetLastError();
}

template <typename Real>
static hipError_t launch_grad_eri_prec(int cls, const GradEriArgs &a_in, hipStream_t s) {
    // split tasks over waves until the launch holds about 16384 waves (C8H18/6-31G gradient, one
    // MI355X: 6.6, 5.6, 5.0 and 4.8 ms at 2048, 4096, 8192 and 16384)
    GradEriArgs a = a_in;
    a.split = a.ntasks >= 16384 ? 1 : (16384 + a.ntasks - 1) / a.ntasks;
    if (a.split > 64) a.split = 64;
    const size_t waves = (size_t)a.ntasks * a.split;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    const size_t lds = (size_t)3 * a.natoms * sizeof(double);
    switch (cls) {
        case 0: hipLaunchKernelGGL((grad_eri_kernel<0, 0, 0, 0, Real>), grid, block, lds, s, a); break;
        case 1: hipLaunchKernelGGL((grad_eri_kernel<1, 0, 0, 0, Real>), grid, block, lds, s, a); break;
        case 2: hipLaunchKernelGGL((grad_eri_kernel<1, 0, 1, 0, Real>), grid, block, lds, s, a); break;
        case 3: hipLaunchKernelGGL((grad_eri_kernel<1, 1, 0, 0, Real>), grid, block, lds, s, a); break;
        case 4: hipLaunchKernelGGL((grad_eri_kernel<1, 1, 1, 0, Real>), grid, block, lds, s, a); break;
        case 5: hipLaunchKernelGGL((grad_eri_kernel<1, 1, 1, 1, Real>), grid, block, lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_grad_eri(int cls, const GradEriArgs &a, hipStream_t s) {
    if (a.ntasks <= 0) return hipSuccess;
    hipError_t e = launch_grad_eri_prec<double>(cls, a, s);
    if (e == hipSuccess && a.pairQ) e = launch_grad_eri_prec<float>(cls, a, s);  // mixed: FP32 pass
    return e;
}

hipError_t launch_grad_one(int pair_class, const GradOneArgs &a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    const size_t threads = (size_t)a.npairs * ((size_t)a.natoms + 1);
    const dim3 grid((unsigned)((threads + 127) / 128)), block(128);
    const size_t lds = (size_t)3 * a.natoms * sizeof(double);
    switch (pair_class) {
        case 0: hipLaunchKernelGGL((grad_one_kernel<0, 0>), grid, block, lds, s, a); break;
        case 1: hipLaunchKernelGGL((grad_one_kernel<1, 0>), grid, block, lds, s, a); break;
        case 2: hipLaunchKernelGGL((grad_one_kernel<1, 1>), grid, block, lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fock
