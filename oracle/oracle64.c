/*
 * oracle64.c -- accurate, unscreened FP64 integral oracle ("Accurate").
 *
 * TEST INFRASTRUCTURE ONLY (see refcpu.c): only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, as the checker.
 *
 * Written independently of the reference's recursive code: textbook
 * McMurchie-Davidson (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
 * Theory", ch. 9) with basis functions (r-A)^{ijk} exp(-a|r-A|^2), iterative Hermite
 * coefficients, the CORRECT R_{tuv} recursion, and a Boys function from its
 * convergent series / erf closed form. Results are converted to the reference's
 * basis-function convention at the end: the reference evaluates functions as
 * (A-r)^{ijk} (basis/src/contracted_gaussian.rs:320-331, dist = B - A in :210,270),
 * i.e. each function carries a factor (-1)^{i+j+k}, so every integral is multiplied
 * by (-1)^{sum of the four (two) angular momenta}.
 *
 * Parity status: pinned by tests/test_oracle.py -- Boys against mpmath; (ss|ss)
 * against its closed form; p-integrals against high-precision centre derivatives
 * of the s closed form; the Szabo-Ostlund H2/STO-3G known answers; 8-fold
 * symmetry and rotation invariance.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXL 2             /* per shell: s, p, d */
#define MAXLT (4 * MAXL)   /* total angular momentum of a quartet */

/* ---------------------------------------------------------------- Boys */

/* F_n(T) = int_0^1 x^{2n} exp(-T x^2) dx.
 * T <= 45: F_n(T) = exp(-T) sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))  (all terms positive).
 * T > 45 : F_0 = sqrt(pi/T)/2 * erf(sqrt T), then F_{m+1} = ((2m+1) F_m - exp(-T)) / (2T). */
double o64_boys(double T, int n) {
    if (T <= 45.0) {
        double term = 1.0 / (2.0 * n + 1.0);
        double sum = term;
        for (int k = 1; k < 2000; ++k) {
            term *= 2.0 * T / (2.0 * n + 2.0 * k + 1.0);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        return exp(-T) * sum;
    }
    double F = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
    double e = exp(-T);
    for (int m = 0; m < n; ++m) F = ((2.0 * m + 1.0) * F - e) / (2.0 * T);
    return F;
}

void o64_boys_array(const double *T, int nT, int n, double *out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nT; ++i) out[i] = o64_boys(T[i], n);
}

/* --------------------------------------------------------- basis/shells */

typedef struct {
    int n;
    const double *centers;
    const int32_t *angular;
    const int32_t *prim_off;
    const double *exps;
    const double *coefs;
} o64_basis;

typedef struct {
    int L, f0, nf, K;
    double A[3];
} o64_shell;

static int ncart(int L) { return (L + 1) * (L + 2) / 2; }

/* Group consecutive functions that share centre, exponents and total angular momentum
 * into shells of ncart(L) components. Returns number of shells or -1. */
static int build_shells(const o64_basis *B, o64_shell *sh) {
    int ns = 0, f = 0;
    while (f < B->n) {
        int L = B->angular[3 * f] + B->angular[3 * f + 1] + B->angular[3 * f + 2];
        int nf = ncart(L);
        if (L > MAXL || f + nf > B->n) return -1;
        int K = B->prim_off[f + 1] - B->prim_off[f];
        for (int g = f + 1; g < f + nf; ++g) {
            int Lg = B->angular[3 * g] + B->angular[3 * g + 1] + B->angular[3 * g + 2];
            if (Lg != L || B->prim_off[g + 1] - B->prim_off[g] != K) return -1;
            for (int x = 0; x < 3; ++x)
                if (B->centers[3 * g + x] != B->centers[3 * f + x]) return -1;
            for (int k = 0; k < K; ++k)
                if (B->exps[B->prim_off[g] + k] != B->exps[B->prim_off[f] + k]) return -1;
        }
        sh[ns].L = L;
        sh[ns].f0 = f;
        sh[ns].nf = nf;
        sh[ns].K = K;
        for (int x = 0; x < 3; ++x) sh[ns].A[x] = B->centers[3 * f + x];
        ++ns;
        f += nf;
    }
    return ns;
}

/* Hermite coefficients E[i][j][t] along one axis for exponents a (centre A), b (centre B). */
static void hermite_1d(int la, int lb, double a, double b, double Ax, double Bx, double E[MAXL + 3][MAXL + 3][2 * MAXL + 5]) {
    double p = a + b, mu = a * b / p;
    double Px = (a * Ax + b * Bx) / p;
    double XAB = Ax - Bx, XPA = Px - Ax, XPB = Px - Bx;
    double h = 1.0 / (2.0 * p);
    memset(E, 0, sizeof(double) * (MAXL + 3) * (MAXL + 3) * (2 * MAXL + 5));
    E[0][0][0] = exp(-mu * XAB * XAB);
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i > 0) {
                    v = XPA * E[i - 1][j][t] + (t + 1) * E[i - 1][j][t + 1];
                    if (t > 0) v += h * E[i - 1][j][t - 1];
                } else {
                    v = XPB * E[i][j - 1][t] + (t + 1) * E[i][j - 1][t + 1];
                    if (t > 0) v += h * E[i][j - 1][t - 1];
                }
                E[i][j][t] = v;
            }
        }
}

/* R_{tuv}(alpha, X) for t+u+v <= L, correct recursion (Helgaker 9.9.18-20). */
static void hermite_coulomb(int L, double alpha, const double X[3], double R[MAXLT + 1][MAXLT + 1][MAXLT + 1]) {
    static const int NMAX = MAXLT + 1;
    double Rn[MAXLT + 1][MAXLT + 1][MAXLT + 1][MAXLT + 1]; /* [n][t][u][v] */
    double T = alpha * (X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    double m2a = -2.0 * alpha, f = 1.0;
    (void)NMAX;
    for (int n = 0; n <= L; ++n) {
        Rn[n][0][0][0] = f * o64_boys(T, n);
        f *= m2a;
    }
    for (int n = L - 1; n >= 0; --n) {
        int top = L - n;
        for (int tot = 1; tot <= top; ++tot)
            for (int t = 0; t <= tot; ++t)
                for (int u = 0; u <= tot - t; ++u) {
                    int v = tot - t - u;
                    double val;
                    if (t > 0) {
                        val = X[0] * Rn[n + 1][t - 1][u][v];
                        if (t > 1) val += (t - 1) * Rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = X[1] * Rn[n + 1][t][u - 1][v];
                        if (u > 1) val += (u - 1) * Rn[n + 1][t][u - 2][v];
                    } else {
                        val = X[2] * Rn[n + 1][t][u][v - 1];
                        if (v > 1) val += (v - 1) * Rn[n + 1][t][u][v - 2];
                    }
                    Rn[n][t][u][v] = val;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = Rn[0][t][u][v];
}

typedef struct {
    double p, P[3], coef_pref;
    double Ex[MAXL + 3][MAXL + 3][2 * MAXL + 5], Ey[MAXL + 3][MAXL + 3][2 * MAXL + 5], Ez[MAXL + 3][MAXL + 3][2 * MAXL + 5];
} prim_pair;

static void make_pair(const o64_basis *B, const o64_shell *sa, const o64_shell *sb, int i, int j, prim_pair *pp) {
    double a = B->exps[B->prim_off[sa->f0] + i], b = B->exps[B->prim_off[sb->f0] + j];
    pp->p = a + b;
    for (int x = 0; x < 3; ++x) pp->P[x] = (a * sa->A[x] + b * sb->A[x]) / pp->p;
    hermite_1d(sa->L, sb->L, a, b, sa->A[0], sb->A[0], pp->Ex);
    hermite_1d(sa->L, sb->L, a, b, sa->A[1], sb->A[1], pp->Ey);
    hermite_1d(sa->L, sb->L, a, b, sa->A[2], sb->A[2], pp->Ez);
}

/* All component integrals of one shell quartet, out[fa][fb][fc][fd] (row-major, nf each). */
static void shell_quartet(const o64_basis *B, const o64_shell *sa, const o64_shell *sb, const o64_shell *sc,
                          const o64_shell *sd, double *out) {
    int na = sa->nf, nb = sb->nf, nc = sc->nf, nd = sd->nf;
    int L = sa->L + sb->L + sc->L + sd->L;
    memset(out, 0, sizeof(double) * na * nb * nc * nd);
    prim_pair bra, ket;
    double R[MAXLT + 1][MAXLT + 1][MAXLT + 1];
    const double two_pi_52 = 2.0 * pow(M_PI, 2.5);
    for (int i = 0; i < sa->K; ++i)
        for (int j = 0; j < sb->K; ++j) {
            make_pair(B, sa, sb, i, j, &bra);
            for (int k = 0; k < sc->K; ++k)
                for (int l = 0; l < sd->K; ++l) {
                    make_pair(B, sc, sd, k, l, &ket);
                    double p = bra.p, q = ket.p;
                    double alpha = p * q / (p + q);
                    double X[3] = {bra.P[0] - ket.P[0], bra.P[1] - ket.P[1], bra.P[2] - ket.P[2]};
                    hermite_coulomb(L, alpha, X, R);
                    double pref = two_pi_52 / (p * q * sqrt(p + q));
                    for (int fa = 0; fa < na; ++fa) {
                        const int *aa = B->angular + 3 * (sa->f0 + fa);
                        double ca = B->coefs[B->prim_off[sa->f0 + fa] + i];
                        for (int fb = 0; fb < nb; ++fb) {
                            const int *ab = B->angular + 3 * (sb->f0 + fb);
                            double cb = B->coefs[B->prim_off[sb->f0 + fb] + j];
                            for (int fc = 0; fc < nc; ++fc) {
                                const int *ac = B->angular + 3 * (sc->f0 + fc);
                                double cc = B->coefs[B->prim_off[sc->f0 + fc] + k];
                                for (int fd = 0; fd < nd; ++fd) {
                                    const int *ad = B->angular + 3 * (sd->f0 + fd);
                                    double cd = B->coefs[B->prim_off[sd->f0 + fd] + l];
                                    double val = 0.0;
                                    for (int t = 0; t <= aa[0] + ab[0]; ++t)
                                        for (int u = 0; u <= aa[1] + ab[1]; ++u)
                                            for (int v = 0; v <= aa[2] + ab[2]; ++v) {
                                                double eab = bra.Ex[aa[0]][ab[0]][t] * bra.Ey[aa[1]][ab[1]][u] *
                                                             bra.Ez[aa[2]][ab[2]][v];
                                                if (eab == 0.0) continue;
                                                double inner = 0.0;
                                                for (int tt = 0; tt <= ac[0] + ad[0]; ++tt)
                                                    for (int uu = 0; uu <= ac[1] + ad[1]; ++uu)
                                                        for (int vv = 0; vv <= ac[2] + ad[2]; ++vv) {
                                                            double ecd = ket.Ex[ac[0]][ad[0]][tt] *
                                                                         ket.Ey[ac[1]][ad[1]][uu] *
                                                                         ket.Ez[ac[2]][ad[2]][vv];
                                                            double sgn = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                                                            inner += sgn * ecd * R[t + tt][u + uu][v + vv];
                                                        }
                                                val += eab * inner;
                                            }
                                    out[((fa * nb + fb) * nc + fc) * nd + fd] += ca * cb * cc * cd * pref * val;
                                }
                            }
                        }
                    }
                }
        }
    if (L & 1)
        for (int x = 0; x < na * nb * nc * nd; ++x) out[x] = -out[x];
}

static inline size_t pair_index(size_t x, size_t y) { return x * (x + 1) / 2 + y; }
static inline size_t packed_slot(size_t x, size_t y, size_t z, size_t w) {
    if (x < y) { size_t t = x; x = y; y = t; }
    if (z < w) { size_t t = z; z = w; w = t; }
    size_t xy = pair_index(x, y), zw = pair_index(z, w);
    if (xy < zw) { size_t t = xy; xy = zw; zw = t; }
    return xy * (xy + 1) / 2 + zw;
}

/* Every unique (canonical) ERI into `packed` (layout as refcpu.c: slot XY*(XY+1)/2+ZW).
 * Returns 0, or -1 if the basis cannot be grouped into shells. */
int o64_eri_packed(const o64_basis *B, double *packed) {
    o64_shell *sh = (o64_shell *)malloc(sizeof(o64_shell) * (B->n + 1));
    int ns = build_shells(B, sh);
    if (ns < 0) { free(sh); return -1; }
    int nsp = ns * (ns + 1) / 2;
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 6 * 6 * 6 * 6);
#pragma omp for schedule(dynamic, 1)
        for (int ab = nsp - 1; ab >= 0; --ab) {
            int a = (int)((sqrt(8.0 * ab + 1.0) - 1.0) / 2.0);
            while (a * (a + 1) / 2 > ab) --a;
            while ((a + 1) * (a + 2) / 2 <= ab) ++a;
            int b = ab - a * (a + 1) / 2;
            for (int c = 0; c <= a; ++c)
                for (int d = 0; d <= c; ++d) {
                    int cd = c * (c + 1) / 2 + d;
                    if (cd > ab) continue;
                    shell_quartet(B, &sh[a], &sh[b], &sh[c], &sh[d], buf);
                    for (int fa = 0; fa < sh[a].nf; ++fa)
                        for (int fb = 0; fb < sh[b].nf; ++fb)
                            for (int fc = 0; fc < sh[c].nf; ++fc)
                                for (int fd = 0; fd < sh[d].nf; ++fd) {
                                    int ia = sh[a].f0 + fa, ib = sh[b].f0 + fb, ic = sh[c].f0 + fc, id = sh[d].f0 + fd;
                                    packed[packed_slot(ia, ib, ic, id)] =
                                        buf[((fa * sh[b].nf + fb) * sh[c].nf + fc) * sh[d].nf + fd];
                                }
                }
        }
        free(buf);
    }
    free(sh);
    return 0;
}

/* One shell quartet through the public interface (tests): out has nf_a*nf_b*nf_c*nf_d entries
 * for the shells starting at functions fa, fb, fc, fd. Returns -1 on a bad basis. */
int o64_shell_quartet(const o64_basis *B, int fa, int fb, int fc, int fd, double *out) {
    o64_shell *sh = (o64_shell *)malloc(sizeof(o64_shell) * (B->n + 1));
    int ns = build_shells(B, sh);
    if (ns < 0) { free(sh); return -1; }
    int ia = -1, ib = -1, ic = -1, id = -1;
    for (int s = 0; s < ns; ++s) {
        if (sh[s].f0 == fa) ia = s;
        if (sh[s].f0 == fb) ib = s;
        if (sh[s].f0 == fc) ic = s;
        if (sh[s].f0 == fd) id = s;
    }
    if (ia < 0 || ib < 0 || ic < 0 || id < 0) { free(sh); return -1; }
    shell_quartet(B, &sh[ia], &sh[ib], &sh[ic], &sh[id], out);
    free(sh);
    return 0;
}

/* J_ij = sum_kl (ij|kl) P_kl, K_ij = sum_kl (ik|jl) P_kl from packed integrals (row-major). */
void o64_jk_packed(int n, const double *packed, const double *P, double *J, double *K) {
    size_t N = (size_t)n;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double sj = 0.0, sk = 0.0;
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) {
                    double pkl = P[(size_t)k * N + l];
                    sj += packed[packed_slot(i, j, k, l)] * pkl;
                    sk += packed[packed_slot(i, k, j, l)] * pkl;
                }
            J[(size_t)i * N + j] = J[(size_t)j * N + i] = sj;
            K[(size_t)i * N + j] = K[(size_t)j * N + i] = sk;
        }
}

/* One-electron integrals (row-major, symmetric): overlap, kinetic, nuclear attraction. */
static double ov1d(int i, int j, const double E[MAXL + 3][MAXL + 3][2 * MAXL + 5]) { return E[i][j][0]; }

void o64_one_electron(const o64_basis *B, int natoms, const double *nuc, const double *charges, double *S,
                      double *T, double *V) {
    int n = B->n;
#pragma omp parallel for schedule(dynamic)
    for (int fa = 0; fa < n; ++fa)
        for (int fb = 0; fb <= fa; ++fb) {
            const int *la = B->angular + 3 * fa, *lb = B->angular + 3 * fb;
            const double *A = B->centers + 3 * fa, *Bc = B->centers + 3 * fb;
            double s = 0.0, t = 0.0, v = 0.0;
            for (int i = 0; i < B->prim_off[fa + 1] - B->prim_off[fa]; ++i)
                for (int j = 0; j < B->prim_off[fb + 1] - B->prim_off[fb]; ++j) {
                    double a = B->exps[B->prim_off[fa] + i], b = B->exps[B->prim_off[fb] + j];
                    double ca = B->coefs[B->prim_off[fa] + i], cb = B->coefs[B->prim_off[fb] + j];
                    double p = a + b;
                    double Ex[MAXL + 3][MAXL + 3][2 * MAXL + 5], Ey[MAXL + 3][MAXL + 3][2 * MAXL + 5],
                        Ez[MAXL + 3][MAXL + 3][2 * MAXL + 5];
                    /* angular up to l+2 for the kinetic operator */
                    hermite_1d(MAXL, MAXL + 2, a, b, A[0], Bc[0], Ex);
                    hermite_1d(MAXL, MAXL + 2, a, b, A[1], Bc[1], Ey);
                    hermite_1d(MAXL, MAXL + 2, a, b, A[2], Bc[2], Ez);
                    double norm = pow(M_PI / p, 1.5);
                    double sx = ov1d(la[0], lb[0], Ex), sy = ov1d(la[1], lb[1], Ey), sz = ov1d(la[2], lb[2], Ez);
                    s += ca * cb * norm * sx * sy * sz;
                    /* kinetic: -1/2 <a|nabla^2|b>, per axis T_ij = -2b^2 S_{i,j+2} + b(2j+1) S_ij - j(j-1)/2 S_{i,j-2} */
                    double tx = -2.0 * b * b * ov1d(la[0], lb[0] + 2, Ex) + b * (2 * lb[0] + 1) * sx -
                                (lb[0] >= 2 ? 0.5 * lb[0] * (lb[0] - 1) * ov1d(la[0], lb[0] - 2, Ex) : 0.0);
                    double ty = -2.0 * b * b * ov1d(la[1], lb[1] + 2, Ey) + b * (2 * lb[1] + 1) * sy -
                                (lb[1] >= 2 ? 0.5 * lb[1] * (lb[1] - 1) * ov1d(la[1], lb[1] - 2, Ey) : 0.0);
                    double tz = -2.0 * b * b * ov1d(la[2], lb[2] + 2, Ez) + b * (2 * lb[2] + 1) * sz -
                                (lb[2] >= 2 ? 0.5 * lb[2] * (lb[2] - 1) * ov1d(la[2], lb[2] - 2, Ez) : 0.0);
                    t += ca * cb * norm * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
                    double P[3];
                    for (int x = 0; x < 3; ++x) P[x] = (a * A[x] + b * Bc[x]) / p;
                    int L = la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2];
                    for (int at = 0; at < natoms; ++at) {
                        double X[3] = {P[0] - nuc[3 * at], P[1] - nuc[3 * at + 1], P[2] - nuc[3 * at + 2]};
                        double R[MAXLT + 1][MAXLT + 1][MAXLT + 1];
                        hermite_coulomb(L, p, X, R);
                        double sum = 0.0;
                        for (int tt = 0; tt <= la[0] + lb[0]; ++tt)
                            for (int uu = 0; uu <= la[1] + lb[1]; ++uu)
                                for (int vv = 0; vv <= la[2] + lb[2]; ++vv)
                                    sum += Ex[la[0]][lb[0]][tt] * Ey[la[1]][lb[1]][uu] * Ez[la[2]][lb[2]][vv] *
                                           R[tt][uu][vv];
                        v += -charges[at] * ca * cb * 2.0 * M_PI / p * sum;
                    }
                }
            int sgn = ((la[0] + la[1] + la[2] + lb[0] + lb[1] + lb[2]) & 1) ? -1 : 1;
            S[(size_t)fa * n + fb] = S[(size_t)fb * n + fa] = sgn * s;
            T[(size_t)fa * n + fb] = T[(size_t)fb * n + fa] = sgn * t;
            V[(size_t)fa * n + fb] = V[(size_t)fb * n + fa] = sgn * v;
        }
}
