// one_electron.hip -- overlap, kinetic and nuclear-attraction matrices on the GPU.
//
// One thread per shell pair computes every component pair. The arithmetic follows the
// reference's primitive formulas (basis/src/contracted_gaussian.rs:51-137):
//   S = E^{ax bx}_0 E^{ay by}_0 E^{az bz}_0 (pi/p)^{3/2}
//   T = b(2 lb + 3) S(a,b) - 2 b^2 [S(a,b+2x) + S(a,b+2y) + S(a,b+2z)]
//       - 1/2 [bx(bx-1) S(a,b-2x) + by(by-1) S(a,b-2y) + bz(bz-1) S(a,b-2z)]
//   V = sum_C -Z_C (2 pi / p) sum_tuv E_t E_u E_v R_{tuv}(p, C - P)
// with the reference's E (dist = B - A) and R (x branch per policy) conventions.
#include "fock_device.hpp"
#include "kernels.hpp"

namespace fock {

template <int LA, int LB, int POLICY>
__global__ __launch_bounds__(128) void one_electron_kernel(OneElArgs args, const int *__restrict__ shell_f0,
                                                           const int *__restrict__ pair_a,
                                                           const int *__restrict__ pair_b, int npairs) {
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr int LT = LA + LB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const int fa = shell_f0[pair_a[t]], fb = shell_f0[pair_b[t]];
    const double A[3] = {args.centers[3 * fa], args.centers[3 * fa + 1], args.centers[3 * fa + 2]};
    const double B[3] = {args.centers[3 * fb], args.centers[3 * fb + 1], args.centers[3 * fb + 2]};
    const double dist[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double S[NA][NB], Tk[NA][NB], V[NA][NB];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) S[i][j] = Tk[i][j] = V[i][j] = 0.0;
    const int ka0 = args.prim_off[fa], ka1 = args.prim_off[fa + 1];
    const int kb0 = args.prim_off[fb], kb1 = args.prim_off[fb + 1];
    BoysTables bt = args.bt;
    for (int ia = ka0; ia < ka1; ++ia) {
        const double a = args.exps[ia];
        for (int ib = kb0; ib < kb1; ++ib) {
            const double b = args.exps[ib];
            const double p = a + b;
            const double q = a * b / p;
            const double inv2p = 1.0 / (2.0 * p);
            // per-axis E (b angular up to LB + 2 for the kinetic shifts), exp factor included
            double E[3][LA + 1][LB + 3][LT + 3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const double xa = -(q * dist[x] / a), xb = q * dist[x] / b;
                hermite_1d<LA, LB + 2, double>(xa, xb, inv2p, E[x]);
                const double ex = exp(-q * (dist[x] * dist[x]));
#pragma unroll
                for (int i = 0; i <= LA; ++i)
#pragma unroll
                    for (int j = 0; j <= LB + 2; ++j)
#pragma unroll
                        for (int tt = 0; tt <= LT + 2; ++tt) E[x][i][j][tt] *= ex;
            }
            const double norm = sqrt((M_PI / p) * (M_PI / p) * (M_PI / p));
            double Pc[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) Pc[x] = (a * A[x] + b * B[x]) / p;
            // R_{tuv}(p, C - P) summed over nuclei with weights -Z_C 2 pi / p
            double Rsum[LT + 1][LT + 1][LT + 1];
#pragma unroll
            for (int x = 0; x <= LT; ++x)
#pragma unroll
                for (int y = 0; y <= LT; ++y)
#pragma unroll
                    for (int z = 0; z <= LT; ++z) Rsum[x][y][z] = 0.0;
            for (int c = 0; c < args.natoms; ++c) {
                const double X = args.nuc_pos[3 * c] - Pc[0], Y = args.nuc_pos[3 * c + 1] - Pc[1],
                             Z = args.nuc_pos[3 * c + 2] - Pc[2];
                const double T = p * (X * X + Y * Y + Z * Z);
                double Fs[LT + 1];
                if constexpr (POLICY == 0) boys_ref<LT>(T, bt.lut_ref, Fs);
                else boys_taylor64<LT>(T, bt.tab64, Fs);
                double f = -2.0 * p;
#pragma unroll
                for (int n = 1; n <= LT; ++n) {
                    Fs[n] *= f;
                    f *= -2.0 * p;
                }
                double R[LT + 1][LT + 1][LT + 1];
                hermite_coulomb<LT, POLICY == 0, double>(Fs, X, Y, Z, R);
                const double wgt = -args.nuc_charge[c] * (2.0 * M_PI) / p;
#pragma unroll
                for (int x = 0; x <= LT; ++x)
#pragma unroll
                    for (int y = 0; y <= LT - x; ++y)
#pragma unroll
                        for (int z = 0; z <= LT - x - y; ++z) Rsum[x][y][z] += wgt * R[x][y][z];
            }
            const double ca = args.coefs[ia], cb = args.coefs[ib];
            static_for<0, NA>([&](auto i_) {
                constexpr int i = decltype(i_)::value;
                constexpr int ax = cart(LA, i, 0), ay = cart(LA, i, 1), az = cart(LA, i, 2);
                static_for<0, NB>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    constexpr int bx = cart(LB, j, 0), by = cart(LB, j, 1), bz = cart(LB, j, 2);
                    const double cc = ca * cb;
                    const double s0 = E[0][ax][bx][0] * E[1][ay][by][0] * E[2][az][bz][0] * norm;
                    const double sx2 = E[0][ax][bx + 2][0] * E[1][ay][by][0] * E[2][az][bz][0] * norm;
                    const double sy2 = E[0][ax][bx][0] * E[1][ay][by + 2][0] * E[2][az][bz][0] * norm;
                    const double sz2 = E[0][ax][bx][0] * E[1][ay][by][0] * E[2][az][bz + 2][0] * norm;
                    double sxm = 0.0, sym = 0.0, szm = 0.0;
                    if constexpr (bx >= 2) sxm = E[0][ax][bx - 2][0] * E[1][ay][by][0] * E[2][az][bz][0] * norm;
                    if constexpr (by >= 2) sym = E[0][ax][bx][0] * E[1][ay][by - 2][0] * E[2][az][bz][0] * norm;
                    if constexpr (bz >= 2) szm = E[0][ax][bx][0] * E[1][ay][by][0] * E[2][az][bz - 2][0] * norm;
                    const double term0 = b * (double)(2 * (bx + by + bz) + 3) * s0;
                    const double term1 = -2.0 * (b * b) * (sx2 + sy2 + sz2);
                    const double term2 = -0.5 * ((double)(bx * (bx - 1)) * sxm + (double)(by * (by - 1)) * sym +
                                                 (double)(bz * (bz - 1)) * szm);
                    double v = 0.0;
                    static_for<0, ax + bx + 1>([&](auto t_) {
                        constexpr int tt = decltype(t_)::value;
                        static_for<0, ay + by + 1>([&](auto u_) {
                            constexpr int uu = decltype(u_)::value;
                            static_for<0, az + bz + 1>([&](auto v_) {
                                constexpr int vv = decltype(v_)::value;
                                v += E[0][ax][bx][tt] * E[1][ay][by][uu] * E[2][az][bz][vv] * Rsum[tt][uu][vv];
                            });
                        });
                    });
                    S[i][j] += cc * s0;
                    Tk[i][j] += cc * (term0 + term1 + term2);
                    V[i][j] += cc * v;
                });
            });
        }
    }
    const size_t N = (size_t)args.nbf;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const size_t r = fa + i, c = fb + j;
            if (args.S) { args.S[r * N + c] = S[i][j]; args.S[c * N + r] = S[i][j]; }
            if (args.T) { args.T[r * N + c] = Tk[i][j]; args.T[c * N + r] = Tk[i][j]; }
            if (args.V) { args.V[r * N + c] = V[i][j]; args.V[c * N + r] = V[i][j]; }
        }
}

// One launch per shell-pair class (0 ss, 1 ps, 2 pp); pair_a/pair_b list the shells of each
// pair with L(a) >= L(b), shell_f0 the first basis function of each shell.
hipError_t launch_one_electron_class(int pair_class, const OneElArgs &a, const int *shell_f0, const int *pair_a,
                                     const int *pair_b, int npairs, int policy, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    dim3 grid((npairs + 127) / 128), block(128);
    switch (pair_class * 2 + (policy ? 1 : 0)) {
        case 0: hipLaunchKernelGGL((one_electron_kernel<0, 0, 0>), grid, block, 0, s, a, shell_f0, pair_a, pair_b, npairs); break;
        case 1: hipLaunchKernelGGL((one_electron_kernel<0, 0, 1>), grid, block, 0, s, a, shell_f0, pair_a, pair_b, npairs); break;
        case 2: hipLaunchKernelGGL((one_electron_kernel<1, 0, 0>), grid, block, 0, s, a, shell_f0, pair_a, pair_b, npairs); break;
        case 3: hipLaunchKernelGGL((one_electron_kernel<1, 0, 1>), grid, block, 0, s, a, shell_f0, pair_a, pair_b, npairs); break;
        case 4: hipLaunchKernelGGL((one_electron_kernel<1, 1, 0>), grid, block, 0, s, a, shell_f0, pair_a, pair_b, npairs); break;
        case 5: hipLaunchKernelGGL((one_electron_kernel<1, 1, 1>), grid, block, 0, s, a, shell_f0, pair_a, pair_b, npairs); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fock
