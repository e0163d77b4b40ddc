// fock_device.hpp -- device building blocks shared by the integral kernels.
//
// McMurchie-Davidson integrals over Cartesian Gaussians in the reference's
// conventions: a basis function is (A - r)^{ijk} exp(-a |r - A|^2)
// (basis/src/contracted_gaussian.rs:320-331), Hermite coefficients are expanded
// with dist = B - A (basis/src/utils.rs:33-50, called from
// contracted_gaussian.rs:172-182) and the Hermite Coulomb integrals R_{tuv} are
// evaluated at Q - P, ket centre minus bra centre (contracted_gaussian.rs:294).
//
// Policies (compile-time template parameters, one kernel body):
//   BoysRef     the reference's table F_10(0.15k) with linear interpolation and
//               downward recursion (boys/src/lib.rs:58-81)
//   BoysTaylor  Taylor expansion about the nearest point of a 0.1-spaced grid of
//               accurate F_m(T_i), downward recursion, asymptotic form for T >= 45
//   RBug        the reference's x-branch R^n_{t,u,v} = X R^{n+1}_{t-1} + R^{n+1}_{t-2}
//               without the (t-1) factor (basis/src/utils.rs:98-105)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fock {

constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

// Compile-time loop: f(std::integral_constant<int, i>) for i in [Begin, End).
template <int Begin, int End, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (Begin < End) {
        f(std::integral_constant<int, Begin>{});
        static_for<Begin + 1, End>(f);
    }
}

// Cartesian exponent (axis 0 = x, 1 = y, 2 = z) of component c of a shell with angular
// momentum L, in the reference's order: for x in 0..=L, y in 0..=L, z = L - x - y >= 0
// (basis_set/src/lib.rs:87-99).
constexpr int cart(int L, int c, int axis) {
    int idx = 0;
    for (int x = 0; x <= L; ++x)
        for (int y = 0; y <= L; ++y) {
            int z = L - x - y;
            if (z < 0) continue;
            if (idx == c) return axis == 0 ? x : (axis == 1 ? y : z);
            ++idx;
        }
    return -1;
}

// ------------------------------------------------------------------ data layout

// One primitive pair (a, b) of a shell pair, precomputed on the host.
// Kc = c_a c_b exp(-mu X^2) exp(-mu Y^2) exp(-mu Z^2) / p carries the contraction
// coefficients, the Gaussian product prefactor (the t=0 Hermite coefficient of the
// reference is exp(-q dist^2) per axis, utils.rs:39-40) and 1/p of the ERI prefactor.
// xa = A - P and xb = B - P are the reference's recursion factors -(q dist / a) and
// +(q dist / b) (utils.rs:42-48).
template <typename Real>
struct alignas(sizeof(Real) * 4) PrimPair {
    Real p, Px, Py, Pz;
    Real Kc, xax, xay, xaz;
    Real xbx, xby, xbz, inv2p;
};

// Shell pair: primitive-pair range, first basis functions of both shells.
struct PairInfo {
    int pp_start, K, fa, fb;
};

// One wavefront task: bra pair, range of ket pairs (one per lane).
struct Task {
    int bra, ket_begin, ket_count, flags;
};

// One workgroup of the blocked digestion: tasks [first_task, first_task + ntasks) of one bra pair.
struct Block {
    int first_task, ntasks, bra, cost;
};

// ------------------------------------------------------------------ Boys

constexpr int BOYS_GRID_N = 451;        // T_i = 0.1 i, i = 0..450
constexpr double BOYS_GRID_STEP = 0.1;
constexpr double BOYS_TMAX = 45.0;      // asymptotic regime beyond (e^-T / F_n < 1e-14)
constexpr int BOYS_ORDERS = 20;         // tabulated orders m = 0..19 per grid point
constexpr int BOYS_TAYLOR64 = 8;        // terms (k = 0..7): |error| < 0.05^8/8! F_n ~ 1e-15 F_n
constexpr int BOYS_TAYLOR32 = 5;        // terms (k = 0..4): < 0.05^5/5! ~ 3e-9 relative

// Per-order Taylor tables with the coefficients already divided by k!: for a top order L,
// row i holds F_{L+k}(T_i) / k! for k < TERMS followed by exp(-T_i) (PD_COLS columns), so the
// expansion is a plain Horner chain of FMAs. One [BOYS_GRID_N][PD_COLS] slab per L <= PD_LMAX.
constexpr int BOYS_PD_LMAX = 4;
constexpr int BOYS_PD_COLS64 = BOYS_TAYLOR64 + 1;
// The FP32 slabs (staged in LDS by the FP32 block kernels) use a grid twice as coarse with one more
// term: T_i = 0.2 i, i = 0..225, |d| <= 0.1, error < 0.1^6/6! ~ 1.4e-9 relative, and 6.3 KB per
// slab instead of 10.8 KB (LDS is what bounds the FP32 kernels' occupancy)
constexpr int BOYS_PD32_N = 226;
constexpr double BOYS_PD32_STEP = 0.2;
constexpr int BOYS_PD32_TERMS = 6;
constexpr int BOYS_PD_COLS32 = BOYS_PD32_TERMS + 1;

struct BoysTables {
    const double *lut_ref;  // the reference's 200-entry F_10 table
    const double *tab64;    // [BOYS_GRID_N][BOYS_ORDERS]
    const float *tab32;     // [BOYS_GRID_N][BOYS_ORDERS]
    const double *pd64;     // [BOYS_PD_LMAX + 1][BOYS_GRID_N][BOYS_PD_COLS64]
    const float *pd32;      // [BOYS_PD_LMAX + 1][BOYS_PD32_N][BOYS_PD_COLS32]
};

__device__ __forceinline__ double rcp_exact(double x) { return 1.0 / x; }

// Pointers into the constant (scalar-load) and LDS address spaces. A load through a ConstPtr at a
// wave-uniform address is a scalar load whose value is an SGPR operand of the vector
// instructions; indexing through an LdsPtr keeps the address arithmetic in 32 bits.
template <typename Real>
using ConstPtr = const Real __attribute__((address_space(4))) *;
template <typename Real>
using LdsPtr = const Real __attribute__((address_space(3))) *;

// 1/sqrt(x) for the FP32 kernels, x a normal positive number (p + q, T >= BOYS_TMAX): the hardware
// reciprocal square root (1 ulp) without the rescaling that the library adds for subnormal x
__device__ __forceinline__ float rsqrt_normal(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_normal(double x) { return rsqrt(x); }

// (2n-1)!!, n >= 0 (boys/src/lib.rs:5-10 with the saturating 2n-1 of :67)
__host__ __device__ constexpr double dfact_odd(int n) {
    double p = 1.0;
    for (int k = 2 * n - 1; k > 1; k -= 2) p *= k;
    return p;
}

// powi with the square-and-multiply order of Rust's f64::powi
__device__ __forceinline__ double powi_ref(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

// F_0..F_L at T with the reference's algorithm (boys/src/lib.rs:58-81). One downward pass
// from order 10 yields every lower order with the same operations the reference performs in
// its separate calls boys(t, n), n = 0..L.
template <int L>
__device__ __forceinline__ void boys_ref(double T, const double *__restrict__ lut, double F[L + 1]) {
    static_assert(L <= 10, "the reference tabulates orders up to 10");
    if (fabs(T) < 1e-16) {
#pragma unroll
        for (int n = 0; n <= L; ++n) F[n] = 1.0 / (double)(2 * n + 1);
    } else if (fabs(T) >= 30.0) {
#pragma unroll
        for (int n = 0; n <= L; ++n)
            F[n] = dfact_odd(n) * (1.0 / (double)(1 << (n + 1))) * sqrt(M_PI * powi_ref(T, -2 * n - 1));
    } else {
        double index_f = 200.0 * T / 30.0;
        double lower = floor(index_f);
        double upper = ceil(index_f);
        double int_t = index_f - lower;
        // `as usize` saturates (negative or NaN -> 0); upper is clamped to N-1 (lib.rs:76-78)
        int lo = (int)fmin(fmax(lower, 0.0), 199.0);
        int up = (int)fmin(fmax(upper, 0.0), 199.0);
        double curr = int_t * lut[up] + (1.0 - int_t) * lut[lo];
        double e = exp(-T);
        double t2 = 2.0 * T;
#pragma unroll
        for (int k = 10; k > 0; --k) {
            if (k <= L) F[k] = curr;
            curr = (t2 * curr + e) * (1.0 / (double)(2 * k - 1));
        }
        F[0] = curr;
    }
}

// Accurate F_0..F_L (FP64): Taylor expansion about the nearest grid point.
// Low total order: every order from its own expansion; higher: the top order from the
// expansion and the rest by downward recursion with exp(-T) = exp(-T_i) exp(-dT).
template <int L>
__device__ __forceinline__ void boys_taylor64(double T, const double *__restrict__ tab, double F[L + 1]) {
    if (T < BOYS_TMAX) {
        int i = (int)fmax(T * 10.0 + 0.5, 0.0);  // T >= 0 for every integral argument
        double d = (double)i * BOYS_GRID_STEP - T;  // -dT
        const double *row = tab + i * BOYS_ORDERS;
        if constexpr (L <= 2) {
#pragma unroll
            for (int n = 0; n <= L; ++n) {
                double acc = row[n + BOYS_TAYLOR64 - 1];
#pragma unroll
                for (int k = BOYS_TAYLOR64 - 1; k > 0; --k) acc = fma(acc, d * (1.0 / k), row[n + k - 1]);
                F[n] = acc;
            }
        } else {
            double acc = row[L + BOYS_TAYLOR64 - 1];
#pragma unroll
            for (int k = BOYS_TAYLOR64 - 1; k > 0; --k) acc = fma(acc, d * (1.0 / k), row[L + k - 1]);
            F[L] = acc;
            // exp(-T) = exp(-T_i) * exp(d); the grid value exp(-T_i) sits in the last column
            double ed = 1.0 + d * (1.0 + d * (0.5 + d * (1.0 / 6 + d * (1.0 / 24 + d * (1.0 / 120 + d * (1.0 / 720 + d * (1.0 / 5040)))))));
            double e = row[BOYS_ORDERS - 1] * ed;
            double t2 = 2.0 * T;
#pragma unroll
            for (int n = L; n > 0; --n) F[n - 1] = fma(t2, F[n], e) * (1.0 / (2 * n - 1));
        }
    } else {
        double r = rsqrt(T);
        F[0] = 0.886226925452758 * r;  // sqrt(pi)/2 / sqrt(T) (erf(sqrt T) = 1 to 1e-21)
        double h = 0.5 * r * r;        // 1/(2T)
        if constexpr (L <= 4) {
            // upward recursion without the exp(-T) term: relative error below 1e-14 for the
            // orders of s/p quartets (L <= 4) at T >= 45
#pragma unroll
            for (int n = 1; n <= L; ++n) F[n] = F[n - 1] * ((2 * n - 1) * h);
        } else {
            // higher orders (derivative integrals): the exact upward recursion
            // F_n = ((2n - 1) F_{n-1} - exp(-T)) / (2T)
            const double e = exp(-T);
#pragma unroll
            for (int n = 1; n <= L; ++n) F[n] = ((2 * n - 1) * F[n - 1] - e) * h;
        }
    }
}

// Accurate F_0..F_L in FP32 (mixed precision); large T in FP64 for the asymptotic powers.
template <int L>
__device__ __forceinline__ void boys_taylor32(float T, const float *__restrict__ tab, float F[L + 1]) {
    if (T < (float)BOYS_TMAX) {
        int i = (int)fmaxf(T * 10.0f + 0.5f, 0.0f);
        float d = (float)i * 0.1f - T;
        const float *row = tab + i * BOYS_ORDERS;
        if constexpr (L <= 2) {
#pragma unroll
            for (int n = 0; n <= L; ++n) {
                float acc = row[n + BOYS_TAYLOR32 - 1];
#pragma unroll
                for (int k = BOYS_TAYLOR32 - 1; k > 0; --k) acc = fmaf(acc, d * (1.0f / k), row[n + k - 1]);
                F[n] = acc;
            }
        } else {
            float acc = row[L + BOYS_TAYLOR32 - 1];
#pragma unroll
            for (int k = BOYS_TAYLOR32 - 1; k > 0; --k) acc = fmaf(acc, d * (1.0f / k), row[L + k - 1]);
            F[L] = acc;
            float ed = 1.0f + d * (1.0f + d * (0.5f + d * (1.0f / 6 + d * (1.0f / 24))));
            float e = row[BOYS_ORDERS - 1] * ed;
            float t2 = 2.0f * T;
#pragma unroll
            for (int n = L; n > 0; --n) F[n - 1] = fmaf(t2, F[n], e) * (1.0f / (2 * n - 1));
        }
    } else {
        float r = rsqrtf(T);
        F[0] = 0.886226925452758f * r;
        float h = 0.5f * r * r;
#pragma unroll
        for (int n = 1; n <= L; ++n) F[n] = F[n - 1] * ((2 * n - 1) * h);
    }
}

// F_0..F_L from the order-L slab of the pre-divided table (staged in LDS): Horner in FMAs for
// F_L, exp(-T) = exp(-T_i) e^{T_i - T} with the 1/k! constants, then downward recursion.
template <int L, typename Real>
__device__ __forceinline__ void boys_taylor_lds(Real T, LdsPtr<Real> slab, Real F[L + 1]) {
    constexpr int TERMS = sizeof(Real) == 8 ? BOYS_TAYLOR64 : BOYS_PD32_TERMS;
    constexpr int COLS = TERMS + 1;
    if (T < Real(BOYS_TMAX)) {
        // T = alpha |P - Q|^2 >= 0, so the rounded index is >= 0
        int i;
        if constexpr (sizeof(Real) == 8) i = (int)(T * 10.0 + 0.5);
        else i = (int)(T * 5.0f + 0.5f);
        const Real d = (Real)i * Real(sizeof(Real) == 8 ? BOYS_GRID_STEP : BOYS_PD32_STEP) - T;
        const LdsPtr<Real> row = slab + __umul24(i, COLS);  // i < 2^24: the full-rate 24-bit multiply
        Real acc = row[TERMS - 1];
#pragma unroll
        for (int k = TERMS - 2; k >= 0; --k) acc = acc * d + row[k];
        F[L] = acc;
        if constexpr (L > 0) {
            // e^d = sum_k d^k / k!
            Real ed = Real(1.0 / 5040.0);
            if constexpr (TERMS == 8) {
                ed = ed * d + Real(1.0 / 720.0);
                ed = ed * d + Real(1.0 / 120.0);
                ed = ed * d + Real(1.0 / 24.0);
            } else {  // |d| <= 0.1: through d^5 / 5!
                ed = Real(1.0 / 120.0);
                ed = ed * d + Real(1.0 / 24.0);
            }
            ed = ed * d + Real(1.0 / 6.0);
            ed = ed * d + Real(0.5);
            ed = ed * d + Real(1);
            ed = ed * d + Real(1);
            const Real e = row[TERMS] * ed;
            const Real t2 = Real(2) * T;
#pragma unroll
            for (int n = L; n > 0; --n) F[n - 1] = (t2 * F[n] + e) * (Real(1) / Real(2 * n - 1));
        }
    } else {
        static_assert(L <= 4, "extend the grid or keep exp(-T) for higher orders");
        const Real r = rsqrt_normal(T);
        F[0] = Real(0.886226925452758) * r;
        const Real h = Real(0.5) * r * r;
#pragma unroll
        for (int n = 1; n <= L; ++n) F[n] = F[n - 1] * (Real(2 * n - 1) * h);
    }
}

// ------------------------------------------------------------------ Hermite

// 1-D Hermite coefficients E[i][j][t] (i <= LA, j <= LB) of one primitive pair along one
// axis, without the exp factor (carried in Kc). Recursion of basis/src/utils.rs:33-50:
// j > 0 lowers j with factor xb, j == 0 lowers i with factor xa.
template <int LA, int LB, typename Real>
__device__ __forceinline__ void hermite_1d(Real xa, Real xb, Real inv2p, Real E[LA + 1][LB + 1][LA + LB + 1]) {
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LA + LB; ++t) E[i][j][t] = Real(0);
    E[0][0][0] = Real(1);
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j) {
            if (i == 0 && j == 0) continue;
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                Real v;
                if (j > 0) {
                    v = xb * E[i][j - 1][t];
                    if (t + 1 <= i + j - 1) v += Real(t + 1) * E[i][j - 1][t + 1];
                    if (t > 0) v += inv2p * E[i][j - 1][t - 1];
                } else {
                    v = xa * E[i - 1][j][t];
                    if (t + 1 <= i + j - 1) v += Real(t + 1) * E[i - 1][j][t + 1];
                    if (t > 0) v += inv2p * E[i - 1][j][t - 1];
                }
                E[i][j][t] = v;
            }
        }
}

// Hermite Coulomb integrals R^{(0)}_{tuv}, t+u+v <= L, from Fs[n] = (-2 alpha)^n F_n(T),
// with the reference's branch order (basis/src/utils.rs:82-105): t > 0 -> x branch,
// else u > 0 -> y branch, else z branch. RBUG drops the (t-1) factor of the x branch.
template <int L, bool RBUG, typename Real>
__device__ __forceinline__ void hermite_coulomb(const Real Fs[L + 1], Real X, Real Y, Real Z,
                                                Real R[L + 1][L + 1][L + 1]) {
    R[0][0][0] = Fs[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
#pragma unroll
        for (int tot = L - n; tot >= 1; --tot) {
#pragma unroll
            for (int t = 0; t <= tot; ++t)
#pragma unroll
                for (int u = 0; u <= tot - t; ++u) {
                    const int v = tot - t - u;
                    Real val;
                    if (t > 0) {
                        val = X * R[t - 1][u][v];
                        if (t > 1) val += (RBUG ? Real(1) : Real(t - 1)) * R[t - 2][u][v];
                    } else if (u > 0) {
                        val = Y * R[t][u - 1][v];
                        if (u > 1) val += Real(u - 1) * R[t][u - 2][v];
                    } else {
                        val = Z * R[t][u][v - 1];
                        if (v > 1) val += Real(v - 1) * R[t][u][v - 2];
                    }
                    R[t][u][v] = val;
                }
        }
        R[0][0][0] = Fs[n];
    }
}

// ------------------------------------------------------------------ quartet body

// Accumulate one primitive quartet into the contracted integrals I[ab][cd] (bra components
// a-major, ket components c-major). POLICY: 0 = reference arithmetic (BoysRef + RBug),
// 1 = accurate. The bra and ket primitive pairs are passed by value.
template <int LA, int LB, int LC, int LD, int POLICY, typename Real, bool SLAB = false>
__device__ __forceinline__ void accumulate_prim_quartet(const PrimPair<Real> &bra, const PrimPair<Real> &ket,
                                                        const BoysTables &bt,
                                                        Real I[ncart(LA) * ncart(LB)][ncart(LC) * ncart(LD)],
                                                        LdsPtr<Real> lds_slice = nullptr) {
    constexpr int L = LA + LB + LC + LD;
    constexpr int LAB = LA + LB;
    constexpr int LCD = LC + LD;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    const Real p = bra.p, q = ket.p;
    const Real pq = p + q;
    const Real rs = rsqrt_normal(pq);
    const Real alpha = p * q * (rs * rs);
    // comp_diff = Q - P (contracted_gaussian.rs:294)
    const Real X = ket.Px - bra.Px, Y = ket.Py - bra.Py, Z = ket.Pz - bra.Pz;
    const Real T = alpha * (X * X + Y * Y + Z * Z);
    // 2 pi^{5/2} / (p q sqrt(p+q)) with 1/p, 1/q inside Kc; the accurate policy groups the ket
    // factor (the same for every bra primitive of the inner loop, so it is computed once per ket
    // primitive)
    Real pref;
    if constexpr (POLICY == 0) pref = Real(34.98683665524972) * bra.Kc * ket.Kc * rs;
    else pref = bra.Kc * (Real(34.98683665524972) * ket.Kc) * rs;
    Real Fs[L + 1];
    if constexpr (POLICY == 0) {
        double Fd[L + 1];
        boys_ref<L>((double)T, bt.lut_ref, Fd);
#pragma unroll
        for (int n = 0; n <= L; ++n) Fs[n] = (Real)Fd[n];
    } else if constexpr (SLAB) {
        boys_taylor_lds<L, Real>(T, lds_slice, Fs);
    } else if constexpr (sizeof(Real) == 8) {
        boys_taylor64<L>(T, bt.tab64, Fs);
    } else {
        boys_taylor32<L>(T, bt.tab32, Fs);
    }
    {
        const Real m2a = Real(-2) * alpha;
        Real f = m2a;
#pragma unroll
        for (int n = 1; n <= L; ++n) {
            Fs[n] *= f;
            f *= m2a;
        }
    }
    Real R[L + 1][L + 1][L + 1];
    hermite_coulomb<L, POLICY == 0, Real>(Fs, X, Y, Z, R);

    Real Eab[3][LA + 1][LB + 1][LAB + 1];
    Real Ecd[3][LC + 1][LD + 1][LCD + 1];
    if constexpr (LAB > 0) {
        hermite_1d<LA, LB, Real>(bra.xax, bra.xbx, bra.inv2p, Eab[0]);
        hermite_1d<LA, LB, Real>(bra.xay, bra.xby, bra.inv2p, Eab[1]);
        hermite_1d<LA, LB, Real>(bra.xaz, bra.xbz, bra.inv2p, Eab[2]);
    } else {
        Eab[0][0][0][0] = Eab[1][0][0][0] = Eab[2][0][0][0] = Real(1);
    }
    if constexpr (LCD > 0) {
        hermite_1d<LC, LD, Real>(ket.xax, ket.xbx, ket.inv2p, Ecd[0]);
        hermite_1d<LC, LD, Real>(ket.xay, ket.xby, ket.inv2p, Ecd[1]);
        hermite_1d<LC, LD, Real>(ket.xaz, ket.xbz, ket.inv2p, Ecd[2]);
    } else {
        Ecd[0][0][0][0] = Ecd[1][0][0][0] = Ecd[2][0][0][0] = Real(1);
    }

    static_for<0, NC>([&](auto c_) {
        constexpr int c = decltype(c_)::value;
        constexpr int cx = cart(LC, c, 0), cy = cart(LC, c, 1), cz = cart(LC, c, 2);
        static_for<0, ND>([&](auto d_) {
            constexpr int d = decltype(d_)::value;
            constexpr int dx = cart(LD, d, 0), dy = cart(LD, d, 1), dz = cart(LD, d, 2);
            // ket contraction Xk[t][u][v] = sum (-1)^{tt+uu+vv} E^cd_{tt,uu,vv} R_{t+tt,u+uu,v+vv}
            // over the bra Hermite simplex t+u+v <= LA+LB
            Real Xk[LAB + 1][LAB + 1][LAB + 1];
            static_for<0, LAB + 1>([&](auto t_) {
                constexpr int t = decltype(t_)::value;
                static_for<0, LAB + 1 - t>([&](auto u_) {
                    constexpr int u = decltype(u_)::value;
                    static_for<0, LAB + 1 - t - u>([&](auto v_) {
                        constexpr int v = decltype(v_)::value;
                        Real acc;  // the first term initialises it (no 0 + x)
                        static_for<0, cx + dx + 1>([&](auto tt_) {
                            constexpr int tt = decltype(tt_)::value;
                            static_for<0, cy + dy + 1>([&](auto uu_) {
                                constexpr int uu = decltype(uu_)::value;
                                static_for<0, cz + dz + 1>([&](auto vv_) {
                                    constexpr int vv = decltype(vv_)::value;
                                    Real e = Ecd[0][cx][dx][tt] * Ecd[1][cy][dy][uu] * Ecd[2][cz][dz][vv];
                                    if constexpr (tt + uu + vv == 0) acc = e * R[t][u][v];
                                    else if constexpr ((tt + uu + vv) & 1) acc -= e * R[t + tt][u + uu][v + vv];
                                    else acc += e * R[t + tt][u + uu][v + vv];
                                });
                            });
                        });
                        Xk[t][u][v] = acc;
                    });
                });
            });
            static_for<0, NA>([&](auto a_) {
                constexpr int a = decltype(a_)::value;
                constexpr int ax = cart(LA, a, 0), ay = cart(LA, a, 1), az = cart(LA, a, 2);
                static_for<0, NB>([&](auto b_) {
                    constexpr int b = decltype(b_)::value;
                    constexpr int bx = cart(LB, b, 0), by = cart(LB, b, 1), bz = cart(LB, b, 2);
                    Real acc;  // the first term initialises it (no 0 + x)
                    static_for<0, ax + bx + 1>([&](auto t_) {
                        constexpr int t = decltype(t_)::value;
                        static_for<0, ay + by + 1>([&](auto u_) {
                            constexpr int u = decltype(u_)::value;
                            static_for<0, az + bz + 1>([&](auto v_) {
                                constexpr int v = decltype(v_)::value;
                                const Real term = Eab[0][ax][bx][t] * Eab[1][ay][by][u] * Eab[2][az][bz][v] * Xk[t][u][v];
                                if constexpr (t + u + v == 0) acc = term;
                                else acc += term;
                            });
                        });
                    });
                    I[a * NB + b][c * ND + d] += pref * acc;
                });
            });
        });
    });
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace fock
