// dense.hip -- the dense FP64 steps of the SCF iteration (scf/src/lib.rs:139-223).
//
// GEMM on the FP64 matrix cores: v_mfma_f64_16x16x4f64 through
// __builtin_amdgcn_mfma_f64_16x16x4f64. A wavefront owns a 16x16 tile of C; per k-step of
// 4 every lane feeds one element of A (row lane%16, k lane/16) and one of B (k lane/16,
// column lane%16) and holds 4 accumulators C[4*r + lane/16][lane%16], r = 0..3.
// Used for P = 2 C_occ C_occ^T, X^T F X, X C', F P S and the Loewdin transform.
// The remaining helpers are elementwise or single-workgroup deterministic reductions.
#include "kernels.hpp"

namespace fock {

typedef double double4_t __attribute__((ext_vector_type(4)));

// C = alpha op(A) op(B) + beta C on the FP64 matrix cores (v_mfma_f64_16x16x4): one workgroup per
// 16 x 16 tile of C, the K range split over its NW waves (wave w takes the k blocks 4w, 4w + 4 NW,
// ...), the partial tiles summed through LDS. Small tiles and the split keep the SCF's n <= 600
// products spread over the CUs (n = 264: 289 workgroups) and each wave's chain of dependent
// load -> MFMA steps short: with eight waves and eight k blocks of operands requested before their
// MFMAs, a wave of an n = 264 product waits for memory once.
// The product tile op(A) op(B) of rows row0.., columns col0.. in wave 0's accumulators (element
// [row0 + 4 r + lane / 16][col0 + lane % 16] in acc[r]); the other waves return zeros.
template <int NW>
__device__ __forceinline__ double4_t gemm_f64_tile_acc(bool transA, bool transB, int M, int N, int K,
                                                       const double *__restrict__ A, int lda,
                                                       const double *__restrict__ B, int ldb, int row0, int col0) {
    __shared__ double4_t part[NW - 1][64];  // partial tiles of waves 1..NW-1
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int ar = row0 + li, bc = col0 + li;
    double4_t acc = {0.0, 0.0, 0.0, 0.0};
    auto load_a = [&](int kk) -> double {
        return (ar < M && kk < K) ? (transA ? A[(size_t)kk * lda + ar] : A[(size_t)ar * lda + kk]) : 0.0;
    };
    auto load_b = [&](int kk) -> double {
        return (bc < N && kk < K) ? (transB ? B[(size_t)bc * ldb + kk] : B[(size_t)kk * ldb + bc]) : 0.0;
    };
    // DEPTH k blocks of operands are requested before their MFMAs (blocks past K load zeros)
    constexpr int DEPTH = 8;
    for (int k0 = 4 * wave; k0 < K; k0 += 4 * NW * DEPTH) {
        double a[DEPTH], b[DEPTH];
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) {
            a[u] = load_a(k0 + 4 * NW * u + lk);
            b[u] = load_b(k0 + 4 * NW * u + lk);
        }
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    if (wave > 0) part[wave - 1][lane] = acc;
    __syncthreads();
    if (wave != 0) return double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int w = 0; w < NW - 1; ++w) acc += part[w][lane];
    return acc;
}

template <int NW>
__device__ __forceinline__ void gemm_f64_tile(bool transA, bool transB, int M, int N, int K, double alpha,
                                              const double *__restrict__ A, int lda, const double *__restrict__ B,
                                              int ldb, double beta, double *__restrict__ C, int ldc, int row0,
                                              int col0) {
    const double4_t acc = gemm_f64_tile_acc<NW>(transA, transB, M, N, K, A, lda, B, ldb, row0, col0);
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = row0 + 4 * r + lk, j = col0 + li;
        if (i < M && j < N) {
            double v = alpha * acc[r];
            if (beta != 0.0) v += beta * C[(size_t)i * ldc + j];
            C[(size_t)i * ldc + j] = v;
        }
    }
}

// C = alpha op(A) op(B) + beta C on the FP64 matrix cores (v_mfma_f64_16x16x4): one workgroup per
// 16 x 16 tile of C, the K range split over its four waves (wave w takes the k blocks 4w, 4w + 16,
// ...), the partial tiles summed through LDS. Small tiles and the split keep the SCF's n <= 600
// products spread over the CUs (n = 264: 289 workgroups) and each wave's chain of dependent
// load -> MFMA steps a quarter as long.
constexpr int GEMM_NW = 8;  // waves per tile of the single products
__global__ __launch_bounds__(64 * GEMM_NW) void gemm_f64_mfma_kernel(bool transA, bool transB, int M, int N, int K,
                                                                    double alpha, const double *__restrict__ A, int lda,
                                                                    const double *__restrict__ B, int ldb, double beta,
                                                                    double *__restrict__ C, int ldc) {
    gemm_f64_tile<GEMM_NW>(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, blockIdx.y * 16,
                           blockIdx.x * 16);
}

// the same over a batch: blockIdx.z is the record; tiles outside a record's C return at once
__global__ __launch_bounds__(256) void gemm_f64_batch_kernel(bool transA, bool transB, double alpha, double beta,
                                                             const GemmRec *__restrict__ recs) {
    const GemmRec r = recs[blockIdx.z];
    const int row0 = blockIdx.y * 16, col0 = blockIdx.x * 16;
    if (row0 >= r.m || col0 >= r.n) return;  // the same for the whole workgroup
    if (r.flag && *r.flag == 0) return;      // the same for the whole workgroup
    gemm_f64_tile<4>(transA, transB, r.m, r.n, r.k, alpha, r.A, r.lda, r.B, r.ldb, beta, r.C, r.ldc, row0, col0);
}

hipError_t gemm_f64_batch(bool transA, bool transB, double alpha, double beta, const GemmRec *recs, int count,
                          int max_m, int max_n, hipStream_t s) {
    if (count <= 0 || max_m <= 0 || max_n <= 0) return hipSuccess;
    dim3 grid((max_n + 15) / 16, (max_m + 15) / 16, count), block(256);
    hipLaunchKernelGGL(gemm_f64_batch_kernel, grid, block, 0, s, transA, transB, alpha, beta, recs);
    return hipGetLastError();
}

hipError_t gemm_f64(bool transA, bool transB, int m, int n, int k, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *C, int ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    dim3 grid((n + 15) / 16, (m + 15) / 16), block(64 * GEMM_NW);
    hipLaunchKernelGGL(gemm_f64_mfma_kernel, grid, block, 0, s, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta,
                       C, ldc);
    return hipGetLastError();
}

__global__ void symmetrize_add_kernel(int n, const double *__restrict__ x, double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    out[e] = x[e] + x[(int64_t)j * n + i];
}

hipError_t symmetrize_add(int n, const double *x, double *out, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(symmetrize_add_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, x, out);
    return hipGetLastError();
}

// G = (Jb + Jb^T) - scaleG (Kb + Kb^T); F = H + G (H may be null: F not written)
__global__ void fock_from_jk_kernel(int n, const double *__restrict__ H, const double *__restrict__ Jb,
                                    const double *__restrict__ Kb, double *__restrict__ G, double *__restrict__ F,
                                    double scaleG) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    const int64_t t = (int64_t)j * n + i;
    const double g = (Jb[e] + Jb[t]) - scaleG * (Kb[e] + Kb[t]);
    G[e] = g;
    if (F) F[e] = H[e] + g;
}

hipError_t fock_from_jk(int n, const double *H, const double *Jb, const double *Kb, double *G, double *F,
                        double scaleG, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(fock_from_jk_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, H, Jb, Kb, G, F,
                       scaleG);
    return hipGetLastError();
}

// max of non-negative doubles through their bit patterns (monotone for x >= 0; NaN wins)
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

__global__ void antisym_maxabs_kernel(int n, const double *__restrict__ M, double *__restrict__ e,
                                      double *__restrict__ maxabs) {
    __shared__ double red[256];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (idx < (int64_t)n * n) {
        const int i = (int)(idx / n), j = (int)(idx % n);
        const double d = M[idx] - M[(int64_t)j * n + i];
        if (e) e[idx] = d;
        v = fabs(d);
        if (d != d) v = d;  // keep NaN
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const double o = red[threadIdx.x + off];
            const double m = red[threadIdx.x];
            red[threadIdx.x] = (o != o || o > m) ? o : m;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(maxabs, red[0]);
}

hipError_t antisym_maxabs(int n, const double *M, double *e, double *maxabs, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(antisym_maxabs_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, M, e, maxabs);
    return hipGetLastError();
}

// one workgroup per dot product: deterministic reduction order
// Sum over the workgroup of NT threads (a power of two) in a fixed order; thread 0 holds it.
template <int NT>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[NT];
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    return red[0];
}

// sum_i f(i) over i < len by NT threads, four independent loads per thread in flight (the SCF's
// reductions over n^2 elements are latency-bound otherwise); fixed order
template <int NT, typename Fn>
__device__ __forceinline__ double block_reduce(int64_t len, Fn f) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 3 * NT < len; i += 4 * NT) {
        a0 += f(i);
        a1 += f(i + NT);
        a2 += f(i + 2 * NT);
        a3 += f(i + 3 * NT);
    }
    for (; i < len; i += NT) a0 += f(i);
    return block_sum<NT>((a0 + a1) + (a2 + a3));
}

constexpr int REDUCE_NT = 1024;

__global__ __launch_bounds__(REDUCE_NT) void frob_dots_kernel(int64_t len, const double *__restrict__ a,
                                                              const double *__restrict__ base, int64_t stride,
                                                              double *__restrict__ dots) {
    const double *b = base + (int64_t)blockIdx.x * stride;
    const double v = block_reduce<REDUCE_NT>(len, [&](int64_t i) { return a[i] * b[i]; });
    if (threadIdx.x == 0) dots[blockIdx.x] = v;
}

hipError_t frob_dots(int64_t len, const double *a, const double *base, int64_t stride, int count, double *dots,
                     hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(frob_dots_kernel, dim3(count), dim3(REDUCE_NT), 0, s, len, a, base, stride, dots);
    return hipGetLastError();
}

__global__ void lincomb_kernel(int64_t len, const double *__restrict__ base, int64_t stride,
                               const int *__restrict__ slots, const double *__restrict__ c, int count,
                               double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= len) return;
    double acc = c[0] * base[(int64_t)slots[0] * stride + x];
    for (int i = 1; i < count; ++i) acc = acc + c[i] * base[(int64_t)slots[i] * stride + x];
    out[x] = acc;
}

hipError_t lincomb(int64_t len, const double *base, int64_t stride, const int *slots, const double *c, int count,
                   double *out, hipStream_t s) {
    hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, base, stride, slots,
                       c, count, out);
    return hipGetLastError();
}

hipError_t density_from_coeffs(int n, int nocc, const double *C, double *P, hipStream_t s) {
    // P = 2 C[:, :nocc] C[:, :nocc]^T (scf/src/lib.rs:214-216)
    if (nocc <= 0) return fill_f64((int64_t)n * n, 0.0, P, s);
    return gemm_f64(false, true, n, n, nocc, 2.0, C, n, C, n, 0.0, P, n, s);
}

template <int NT>
__device__ __forceinline__ void energy_trace_kernel_body(int n, const double *__restrict__ P,
                                                         const double *__restrict__ H, const double *__restrict__ G,
                                                         double *__restrict__ out) {
    const double v = block_reduce<NT>((int64_t)n * n, [&](int64_t e) { return P[e] * (2.0 * H[e] + G[e]); });
    if (threadIdx.x == 0) out[0] = 0.5 * v;
}

// E_elec split over ENERGY_PARTS workgroups (each a contiguous range of the n^2 products):
// parts[b] = 1/2 sum over range b of P (2H + G); the caller adds the parts in order
__global__ __launch_bounds__(256) void energy_parts_kernel(int n, const double *__restrict__ P,
                                                           const double *__restrict__ H,
                                                           const double *__restrict__ G, double *__restrict__ parts) {
    const int64_t len = (int64_t)n * n;
    const int64_t e0 = len * blockIdx.x / ENERGY_PARTS, e1 = len * (blockIdx.x + 1) / ENERGY_PARTS;
    const double v = block_reduce<256>(e1 - e0, [&](int64_t i) {
        const int64_t e = e0 + i;
        return P[e] * (2.0 * H[e] + G[e]);
    });
    if (threadIdx.x == 0) parts[blockIdx.x] = 0.5 * v;
}
hipError_t energy_trace(int n, const double *P, const double *H, const double *G, double *parts, hipStream_t s) {
    hipLaunchKernelGGL(energy_parts_kernel, dim3(ENERGY_PARTS), dim3(256), 0, s, n, P, H, G, parts);
    return hipGetLastError();
}

__global__ void sub_kernel(int64_t len, const double *__restrict__ a, const double *__restrict__ b,
                           double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = a[x] - b[x];
}
hipError_t sub_f64(int64_t len, const double *a, const double *b, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sub_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, a, b, out);
    return hipGetLastError();
}

__global__ void mix_kernel(int64_t len, double lam, const double *__restrict__ a, const double *__restrict__ b,
                           double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = lam * a[x] + (1.0 - lam) * b[x];
}
hipError_t mix_f64(int64_t len, double lam, const double *a, const double *b, double *out, hipStream_t s) {
    if (len <= 0) return hipSuccess;
    hipLaunchKernelGGL(mix_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, lam, a, b, out);
    return hipGetLastError();
}

__global__ void fill_kernel(int64_t len, double v, double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = v;
}
hipError_t fill_f64(int64_t len, double v, double *out, hipStream_t s) {
    if (len <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, v, out);
    return hipGetLastError();
}

__global__ void copy_kernel(int64_t len, const double *__restrict__ in, double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = in[x];
}
hipError_t copy_f64(int64_t len, const double *in, double *out, hipStream_t s) {
    if (len <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, in, out);
    return hipGetLastError();
}

// scf/src/lib.rs:151-153: 0.875 * S_ij * (H_ii + H_jj), diagonal included
__global__ void gwh_kernel(int n, const double *__restrict__ S, const double *__restrict__ H,
                           double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    out[e] = 0.875 * S[e] * (H[(int64_t)i * n + i] + H[(int64_t)j * n + j]);
}
hipError_t gwh_guess(int n, const double *S, const double *H, double *out, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(gwh_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, S, H, out);
    return hipGetLastError();
}

// out[i][j] = U[i][j] / sqrt(Dm[j][j])  (scf/src/lib.rs:143-144: diag(1/sqrt((U^T S U)_jj)))
__global__ void scale_columns_kernel(int n, const double *__restrict__ U, const double *__restrict__ Dm, int ldd,
                                     double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int j = (int)(e % n);
    out[e] = U[e] * (1.0 / sqrt(Dm[(int64_t)j * ldd + j]));
}
hipError_t scale_columns(int n, const double *U, const double *Dm, int ldd, double *out, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(scale_columns_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, U, Dm, ldd, out);
    return hipGetLastError();
}

// out[sa][sb] = max |P_ij| over the block of shells sa, sb (density-weighted screening)
__global__ void shell_block_maxabs_kernel(int nshell, const int *__restrict__ f0, const int *__restrict__ nf, int n,
                                          const double *__restrict__ P, float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nshell * nshell) return;
    const int sa = (int)(e / nshell), sb = (int)(e % nshell);
    double m = 0.0;
    for (int i = 0; i < nf[sa]; ++i)
        for (int j = 0; j < nf[sb]; ++j) m = fmax(m, fabs(P[(int64_t)(f0[sa] + i) * n + f0[sb] + j]));
    out[e] = (float)m;
}
hipError_t shell_block_maxabs(int nshell, const int *shell_f0, const int *shell_nf, int n, const double *P,
                              float *out, hipStream_t s) {
    const int64_t len = (int64_t)nshell * nshell;
    hipLaunchKernelGGL(shell_block_maxabs_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, nshell,
                       shell_f0, shell_nf, n, P, out);
    return hipGetLastError();
}

__global__ void pair_block_max_kernel(int npair, const int2 *__restrict__ pair_shell, const float *__restrict__ Pm,
                                      int nshell, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npair) return;
    const int2 ab = pair_shell[j];
    out[j] = Pm[(int64_t)ab.x * nshell + ab.y];
}
hipError_t pair_block_max(int npair, const int2 *pair_shell, const float *Pm, int nshell, float *out, hipStream_t s) {
    if (npair <= 0) return hipSuccess;
    hipLaunchKernelGGL(pair_block_max_kernel, dim3((npair + 255) / 256), dim3(256), 0, s, npair, pair_shell, Pm, nshell,
                       out);
    return hipGetLastError();
}

__global__ __launch_bounds__(1024) void max_f32_kernel(int64_t len, const float *__restrict__ x, float *__restrict__ out) {
    __shared__ float red[1024];
    float m = 0.0f;
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x) {
        const float v = x[i];
        m = (v != v || v > m) ? v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const float o = red[threadIdx.x + off], c = red[threadIdx.x];
            red[threadIdx.x] = (o != o || o > c) ? o : c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}
hipError_t max_f32(int64_t len, const float *x, float *out, hipStream_t s) {
    hipLaunchKernelGGL(max_f32_kernel, dim3(1), dim3(1024), 0, s, len, x, out);
    return hipGetLastError();
}

__global__ void transpose_kernel(int n, const double *__restrict__ in, double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    out[(int64_t)j * n + i] = in[e];
}
hipError_t transpose_f64(int n, const double *in, double *out, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, in, out);
    return hipGetLastError();
}

__global__ void add_kernel(int64_t len, const double *__restrict__ a, const double *__restrict__ b,
                           double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = a[x] + b[x];
}
hipError_t add_f64(int64_t len, const double *a, const double *b, double *out, hipStream_t s) {
    hipLaunchKernelGGL(add_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, a, b, out);
    return hipGetLastError();
}

__global__ void accumulate_g_kernel(int n, const double *__restrict__ Jb, const double *__restrict__ Kb,
                                    double *__restrict__ G) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    const int64_t t = (int64_t)j * n + i;
    G[e] += (Jb[e] + Jb[t]) - 0.5 * (Kb[e] + Kb[t]);
}
hipError_t accumulate_g(int n, const double *Jb, const double *Kb, double *Gacc, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(accumulate_g_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, Jb, Kb, Gacc);
    return hipGetLastError();
}

__global__ void scatter_jk_kernel(int npair, const int2 *__restrict__ pair_shell, const int *__restrict__ f0,
                                  const int *__restrict__ nf, double *__restrict__ Jk, double *__restrict__ J, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npair) return;
    const int2 ab = pair_shell[j];
    const int fa = f0[ab.x], fb = f0[ab.y], na = nf[ab.x], nb = nf[ab.y];
    for (int c = 0; c < na; ++c)
        for (int d = 0; d < nb; ++d) {
            double &v = Jk[(size_t)(c * nb + d) * npair + j];
            J[(size_t)(fa + c) * n + (fb + d)] += v;
            v = 0.0;
        }
}
// Gershgorin bounds of the spectrum of a symmetric matrix: out = {min_i (A_ii - r_i), max_i (A_ii + r_i)},
// r_i = sum_{j != i} |A_ij|; one workgroup.
// (one wave per row: the 64 lanes read the row coalesced; 16 waves)
__global__ __launch_bounds__(1024) void gershgorin_kernel(int n, const double *__restrict__ A, double *__restrict__ out) {
    __shared__ double lo[16], hi[16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double l = HUGE_VAL, h = -HUGE_VAL;
    for (int i = wave; i < n; i += 16) {
        const double *row = A + (size_t)i * n;
        double r = 0.0;
        for (int j = lane; j < n; j += 64)
            if (j != i) r += fabs(row[j]);
        r = wave_sum(r);
        const double d = row[i];
        l = fmin(l, d - r);
        h = fmax(h, d + r);
    }
    if (lane == 0) {
        lo[wave] = l;
        hi[wave] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = lo[0], b = hi[0];
        for (int w = 1; w < 16; ++w) {
            a = fmin(a, lo[w]);
            b = fmax(b, hi[w]);
        }
        out[0] = a;
        out[1] = b;
    }
}
hipError_t gershgorin_bounds(int n, const double *A, double *out, hipStream_t s) {
    hipLaunchKernelGGL(gershgorin_kernel, dim3(1), dim3(1024), 0, s, n, A, out);
    return hipGetLastError();
}

// Coupled Newton-Schulz iteration for S^{-1/2} (Higham, Functions of Matrices, 2008, sec. 6.3):
// Y0 = S / c, Z0 = I; T = (3 I - Z Y) / 2, Y <- Y T, Z <- T Z; then Z -> (S / c)^{-1/2}.
// Y0 = S / bound[1] (a Gershgorin upper bound of the spectrum) and Z0 = I
__global__ void ns_init_kernel(int n, const double *__restrict__ S, const double *__restrict__ bound,
                               double *__restrict__ Y, double *__restrict__ Z) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    Y[e] = S[e] / bound[1];
    Z[e] = i == j ? 1.0 : 0.0;
}
hipError_t ns_init(int n, const double *S, const double *bound, double *Y, double *Z, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(ns_init_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, S, bound, Y, Z);
    return hipGetLastError();
}
// T = (3 I - ZY) / 2 and *res = max(*res, max |I - ZY|) (res starts at 0)
__global__ __launch_bounds__(256) void ns_update_kernel(int n, const double *__restrict__ ZY, double *__restrict__ T,
                                                        double *__restrict__ res) {
    __shared__ double red[256];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (e < (int64_t)n * n) {
        const int i = (int)(e / n), j = (int)(e % n);
        const double d = (i == j ? 1.0 : 0.0) - ZY[e];
        T[e] = 0.5 * ((i == j ? 2.0 : 0.0) + d);
        r = fabs(d);
        if (d != d) r = d;
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const double o = red[threadIdx.x + off], m = red[threadIdx.x];
            red[threadIdx.x] = (o != o || o > m) ? o : m;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(res, red[0]);
}
hipError_t ns_update(int n, const double *ZY, double *T, double *res, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(ns_update_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, ZY, T, res);
    return hipGetLastError();
}
// X = (Z + Z^T) / (2 sqrt(bound[1]))
__global__ void ns_finish_kernel(int n, const double *__restrict__ Z, const double *__restrict__ bound,
                                 double *__restrict__ X) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    X[e] = (Z[e] + Z[(int64_t)j * n + i]) * (0.5 / sqrt(bound[1]));
}
hipError_t ns_finish(int n, const double *Z, const double *bound, double *X, hipStream_t s) {
    const int64_t len = (int64_t)n * n;
    hipLaunchKernelGGL(ns_finish_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, Z, bound, X);
    return hipGetLastError();
}

// sum over the workgroup of NT threads of (t, f), returned to every thread (fixed order)
template <int NT>
__device__ __forceinline__ void block_sum2(double &t, double &f) {
    __shared__ double rt[NT], rf[NT];
    rt[threadIdx.x] = t;
    rf[threadIdx.x] = f;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rt[threadIdx.x] += rt[threadIdx.x + s];
            rf[threadIdx.x] += rf[threadIdx.x + s];
        }
        __syncthreads();
    }
    t = rt[0];
    f = rf[0];
}

// Gershgorin interval of every row of a symmetric matrix: rowb[i] = {A_ii - r_i, A_ii + r_i},
// r_i = sum_{j != i} |A_ij|; one wave per row, four rows per workgroup (the rows are read at once
// by the whole device instead of one after another by one workgroup)
__global__ __launch_bounds__(256) void gershgorin_rows_kernel(int n, const double *__restrict__ A,
                                                              double *__restrict__ rowb) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;  // uniform per wave
    const double *row = A + (size_t)i * n;
    double r = 0.0;
    for (int j = lane; j < n; j += 64)
        if (j != i) r += fabs(row[j]);
    r = wave_sum(r);
    if (lane == 0) {
        const double d = row[i];
        rowb[2 * i] = d - r;
        rowb[2 * i + 1] = d + r;
    }
}
hipError_t gershgorin_rows(int n, const double *A, double *rowb, hipStream_t s) {
    hipLaunchKernelGGL(gershgorin_rows_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, A, rowb);
    return hipGetLastError();
}

// {lo, hi} = {min_i rowb[i].lo, max_i rowb[i].hi} over the n row intervals, to every thread of a
// workgroup of 256
__device__ __forceinline__ void rows_interval(int n, const double *__restrict__ rowb, double &lo, double &hi) {
    __shared__ double sl[4], sh[4];
    double l = HUGE_VAL, h = -HUGE_VAL;
    for (int i = threadIdx.x; i < n; i += 256) {
        l = fmin(l, rowb[2 * i]);
        h = fmax(h, rowb[2 * i + 1]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l = fmin(l, __shfl_xor(l, o));
        h = fmax(h, __shfl_xor(h, o));
    }
    if ((threadIdx.x & 63) == 0) {
        sl[threadIdx.x >> 6] = l;
        sh[threadIdx.x >> 6] = h;
    }
    __syncthreads();
    lo = fmin(fmin(sl[0], sl[1]), fmin(sl[2], sl[3]));
    hi = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// {lo, hi} from the row intervals (gershgorin_rows; workgroup 0 writes them to bounds), then
// X = (hi I - A) / (hi - lo); part[block] = (tr, sum of squares) of the block's elements
__global__ __launch_bounds__(256) void tc2_init_kernel(int n, const double *__restrict__ A,
                                                       const double *__restrict__ rowb, double *__restrict__ bounds,
                                                       double *__restrict__ X, double *__restrict__ part) {
    double lo, hi;
    rows_interval(n, rowb, lo, hi);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        bounds[0] = lo;
        bounds[1] = hi;
    }
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double t = 0.0, f = 0.0;
    if (e < (int64_t)n * n) {
        const int i = (int)(e / n), j = (int)(e % n);
        const double v = ((i == j ? hi : 0.0) - A[e]) / (hi - lo);
        X[e] = v;
        t = i == j ? v : 0.0;
        f = v * v;
    }
    block_sum2<256>(t, f);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = t;
        part[2 * blockIdx.x + 1] = f;
    }
}
int tc2_init_parts(int n) { return (int)(((int64_t)n * n + 255) / 256); }
hipError_t tc2_init(int n, const double *A, const double *rowb, double *bounds, double *X, double *part,
                    hipStream_t s) {
    hipLaunchKernelGGL(tc2_init_kernel, dim3((unsigned)tc2_init_parts(n)), dim3(256), 0, s, n, A, rowb, bounds, X,
                       part);
    return hipGetLastError();
}

// One trace-correcting purification step in one launch: the traces of X come from the previous
// launch's per-workgroup partials (tr X, and tr X^2 = sum of squares for symmetric X), every
// workgroup takes the same choice (X^2, or 2X - X^2 when its trace is closer to nocc), multiplies
// its 16 x 16 tile of X X on the matrix cores, writes the chosen update into Xout and its own
// partials of the new matrix. hist (workgroup (0, 0)) = {tr X, tr X^2, grow} of the input.
__global__ __launch_bounds__(64 * GEMM_NW) void tc2_step_kernel(int n, const double *__restrict__ X,
                                                                double *__restrict__ Xout,
                                                                const double *__restrict__ part_in, int npart_in,
                                                                double *__restrict__ part_out, double nocc,
                                                                double *__restrict__ hist) {
    __shared__ int s_grow;
    // the partials are requested before the product's operands (one memory wait for both); the
    // choice is needed only by the update
    double t = 0.0, f = 0.0;
    for (int p = threadIdx.x; p < npart_in; p += blockDim.x) {
        t += part_in[2 * p];
        f += part_in[2 * p + 1];
    }
    const int row0 = blockIdx.y * 16, col0 = blockIdx.x * 16;
    const double4_t acc = gemm_f64_tile_acc<GEMM_NW>(false, false, n, n, n, X, n, X, n, row0, col0);
    block_sum2<64 * GEMM_NW>(t, f);
    if (threadIdx.x == 0) {
        const int grow = fabs(2.0 * t - f - nocc) < fabs(f - nocc) ? 1 : 0;
        s_grow = grow;
        if (blockIdx.x == 0 && blockIdx.y == 0) {
            hist[0] = t;
            hist[1] = f;
            hist[2] = grow;
        }
    }
    __syncthreads();
    const bool grow = s_grow != 0;
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    double tt = 0.0, ff = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = row0 + 4 * r + lk, j = col0 + li;
        if (i < n && j < n) {
            const double x2 = acc[r];
            const double v = grow ? 2.0 * X[(size_t)i * n + j] - x2 : x2;
            Xout[(size_t)i * n + j] = v;
            if (i == j) tt += v;
            ff += v * v;
        }
    }
    tt = wave_sum(tt);
    ff = wave_sum(ff);
    if (lane == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        part_out[2 * b] = tt;
        part_out[2 * b + 1] = ff;
    }
}
int tc2_step_parts(int n) { return ((n + 15) / 16) * ((n + 15) / 16); }
hipError_t tc2_step(int n, const double *X, double *Xout, const double *part_in, int npart_in, double *part_out,
                    double nocc, double *hist, hipStream_t s) {
    const int nt = (n + 15) / 16;
    hipLaunchKernelGGL(tc2_step_kernel, dim3(nt, nt), dim3(64 * GEMM_NW), 0, s, n, X, Xout, part_in, npart_in,
                       part_out, nocc, hist);
    return hipGetLastError();
}


// Batched symmetric eigensolver, one workgroup per matrix, the matrix in LDS. One-sided (Hestenes)
// Jacobi on B = A + c I, with c from the Gershgorin bound so that every eigenvalue of B is at
// least 1: the columns of U = B V converge to lambda_B,k v_k, so the eigenvalues of A are
// ||u_k|| - c and its eigenvectors u_k / ||u_k|| (one n x n array, no separate V). Each step
// rotates the n/2 disjoint column pairs of a round of the round-robin (circle) order, four lanes
// per pair (n <= 128: all pairs of a round at once in 256 threads); a sweep
// without a rotation ends the iteration (at most JACOBI_SWEEPS). Output as a dsyevd with vectors:
// W ascending, the eigenvector of the k-th smallest eigenvalue at A[k * n + i].
constexpr int JACOBI_SWEEPS = 40;
constexpr double JACOBI_TOL = 1e-14;  // rotate while |u_p . u_q| > tol ||u_p|| ||u_q||
__global__ __launch_bounds__(256) void jacobi_eig_kernel(const EigRec *__restrict__ recs) {
    const EigRec r = recs[blockIdx.x];
    const int n = r.n;
    if (n <= 0 || n > JACOBI_MAX_N) return;  // the same for the whole workgroup
    if (r.flag && *r.flag == 0) return;      // the same for the whole workgroup
    extern __shared__ double U[];  // [n][n]: column k at U[k * n]
    __shared__ double s_red[4];
    __shared__ double s_norm[JACOBI_MAX_N];
    __shared__ int s_rank[JACOBI_MAX_N];
    __shared__ int s_rot;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double *__restrict__ A = r.A;
    // Gershgorin lower bound of the symmetrised matrix
    double lo = HUGE_VAL;
    for (int i = tid; i < n; i += blockDim.x) {
        double rs = 0.0;
        for (int j = 0; j < n; ++j)
            if (j != i) rs += fabs(0.5 * (A[i * n + j] + A[j * n + i]));
        lo = fmin(lo, A[i * n + i] - rs);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lo = fmin(lo, __shfl_xor(lo, off, 64));
    if (lane == 0) s_red[wave] = lo;
    __syncthreads();
    const double c = fmax(0.0, 1.0 - fmin(fmin(s_red[0], s_red[1]), fmin(s_red[2], s_red[3])));
    for (int e = tid; e < n * n; e += blockDim.x) {
        const int k = e / n, i = e - k * n;
        U[e] = 0.5 * (A[i * n + k] + A[k * n + i]) + (i == k ? c : 0.0);
    }
    const int m = n + (n & 1);  // players of the round robin; index n is a bye when n is odd
    for (int sweep = 0; sweep < JACOBI_SWEEPS; ++sweep) {
        if (tid == 0) s_rot = 0;
        __syncthreads();  // U loaded / previous sweep done; s_rot cleared
        bool rotated = false;
        for (int round = 0; round < m - 1; ++round) {
            // circle method: position 0 fixed, positions 1..m-1 rotated by `round`; position j meets
            // position m - 1 - j. Four lanes per pair (one quad), all pairs of the round at once.
            const int j = tid >> 2, sub = tid & 3;
            if (j < m / 2) {
                const int a = j == 0 ? 0 : 1 + (j - 1 + round) % (m - 1);
                const int b = 1 + (m - 2 - j + round) % (m - 1);
                if (a < n && b < n) {  // not a bye (the same for the four lanes of the quad)
                    const int p = min(a, b), q = max(a, b);
                    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll 4
                    for (int i = sub; i < n; i += 4) {
                        const double up = U[p * n + i], uq = U[q * n + i];
                        al += up * up;
                        be += uq * uq;
                        ga += up * uq;
                    }
#pragma unroll
                    for (int off = 1; off < 4; off <<= 1) {
                        al += __shfl_xor(al, off, 4);
                        be += __shfl_xor(be, off, 4);
                        ga += __shfl_xor(ga, off, 4);
                    }
                    if (fabs(ga) > JACOBI_TOL * sqrt(al * be)) {  // the same for the quad
                        const double zeta = (be - al) / (2.0 * ga);
                        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll 4
                        for (int i = sub; i < n; i += 4) {
                            const double up = U[p * n + i], uq = U[q * n + i];
                            U[p * n + i] = cs * up - sn * uq;
                            U[q * n + i] = sn * up + cs * uq;
                        }
                        rotated = true;
                    }
                }
            }
            __syncthreads();  // the next round's pairs share columns with this round's
        }
        if (rotated) s_rot = 1;  // any lane; every writer stores the same value
        __syncthreads();
        const bool again = s_rot != 0;
        __syncthreads();  // every thread has read s_rot before the next sweep clears it
        if (!again) break;
    }
    // eigenvalues ||u_k|| - c, ranks (ties by index), normalised vectors in rank order
    for (int k = wave; k < n; k += 4) {
        double s2 = 0.0;
        for (int i = lane; i < n; i += 64) s2 += U[k * n + i] * U[k * n + i];
        s2 = wave_sum(s2);
        if (lane == 0) s_norm[k] = sqrt(s2);
    }
    __syncthreads();
    for (int k = tid; k < n; k += blockDim.x) {
        const double lk = s_norm[k];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double lj = s_norm[j];
            rank += (lj < lk || (lj == lk && j < k)) ? 1 : 0;
        }
        s_rank[k] = rank;
        r.W[rank] = lk - c;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += blockDim.x) {
        const int k = e / n, i = e - k * n;
        r.A[s_rank[k] * n + i] = U[e] / s_norm[k];
    }
}
hipError_t jacobi_eig_batch(const EigRec *recs, int count, int max_n, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (max_n > JACOBI_MAX_N) return hipErrorInvalidValue;
    hipLaunchKernelGGL(jacobi_eig_kernel, dim3(count), dim3(256), (size_t)max_n * max_n * sizeof(double), s, recs);
    return hipGetLastError();
}

__global__ void scatter_jk_batch_kernel(const ScatterRec *__restrict__ recs) {
    const ScatterRec r = recs[blockIdx.y];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= r.npair) return;
    if (r.flag && *r.flag == 0) return;
    const int2 ab = r.pair_shell[j];
    const int fa = r.shell_f0[ab.x], fb = r.shell_f0[ab.y], na = r.shell_nf[ab.x], nb = r.shell_nf[ab.y];
    for (int c = 0; c < na; ++c)
        for (int d = 0; d < nb; ++d) {
            double &v = r.Jk[(size_t)(c * nb + d) * r.npair + j];
            r.J[(size_t)(fa + c) * r.n + (fb + d)] += v;
            v = 0.0;
        }
}
hipError_t scatter_jk_batch(int nrec, const ScatterRec *recs, int max_npair, hipStream_t s) {
    if (nrec <= 0 || max_npair <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_jk_batch_kernel, dim3((max_npair + 255) / 256, nrec), dim3(256), 0, s, recs);
    return hipGetLastError();
}
hipError_t scatter_jk(int npair, const int2 *pair_shell, const int *shell_f0, const int *shell_nf, double *Jk,
                      double *J, int n, hipStream_t s) {
    if (npair <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_jk_kernel, dim3((npair + 255) / 256), dim3(256), 0, s, npair, pair_shell, shell_f0,
                       shell_nf, Jk, J, n);
    return hipGetLastError();
}

__global__ void level_shift_kernel(int64_t len, const double *__restrict__ F, const double *__restrict__ Sm,
                                   const double *__restrict__ SPS, double sigma, double *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = F[x] + sigma * (Sm[x] - 0.5 * SPS[x]);
}
hipError_t level_shift_fock(int64_t len, const double *F, const double *Sm, const double *SPS, double sigma,
                            double *out, hipStream_t s) {
    hipLaunchKernelGGL(level_shift_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, F, Sm, SPS,
                       sigma, out);
    return hipGetLastError();
}

// DIIS error of every molecule (blockIdx.y): e = FPS - (FPS)^T with max |e| (NaN kept), and the
// copy of F into its ring slot
__global__ __launch_bounds__(256) void diis_error_batch_kernel(const DiisRec *__restrict__ recs) {
    __shared__ double red[256];
    const DiisRec r = recs[blockIdx.y];
    const int64_t len = (int64_t)r.n * r.n;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)blockIdx.x * blockDim.x >= len) return;  // the same for the whole workgroup
    double v = 0.0;
    if (idx < len) {
        const int i = (int)(idx / r.n), j = (int)(idx % r.n);
        const double d = r.FPS[idx] - r.FPS[(int64_t)j * r.n + i];
        r.e[idx] = d;
        r.f_slot[idx] = r.F[idx];
        v = fabs(d);
        if (d != d) v = d;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const double o = red[threadIdx.x + off];
            const double m = red[threadIdx.x];
            red[threadIdx.x] = (o != o || o > m) ? o : m;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(r.dots + 16, red[0]);
}
// <e, errs[s]> for the 16 ring slots of every molecule: one workgroup per (slot, molecule), a fixed
// reduction order
__global__ __launch_bounds__(256) void diis_dots_batch_kernel(const DiisRec *__restrict__ recs) {
    __shared__ double red[256];
    const DiisRec r = recs[blockIdx.y];
    const int64_t len = (int64_t)r.n * r.n;
    const double *b = r.errs + (int64_t)blockIdx.x * len;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x) acc += r.e[i] * b[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) r.dots[blockIdx.x] = red[0];
}
hipError_t diis_error_batch(const DiisRec *recs, int count, int max_n, hipStream_t s) {
    if (count <= 0 || max_n <= 0) return hipSuccess;
    const int64_t len = (int64_t)max_n * max_n;
    hipLaunchKernelGGL(diis_error_batch_kernel, dim3((unsigned)((len + 255) / 256), count), dim3(256), 0, s, recs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(diis_dots_batch_kernel, dim3(16, count), dim3(256), 0, s, recs);
    return hipGetLastError();
}

__global__ void lincomb_batch_kernel(const CombRec *__restrict__ recs) {
    const CombRec r = recs[blockIdx.y];
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= r.len) return;
    double acc = r.c[0] * r.base[(int64_t)r.slots[0] * r.len + x];
    for (int i = 1; i < r.count; ++i) acc = acc + r.c[i] * r.base[(int64_t)r.slots[i] * r.len + x];
    r.out[x] = acc;
}
hipError_t lincomb_batch(const CombRec *recs, int count, int64_t max_len, hipStream_t s) {
    if (count <= 0 || max_len <= 0) return hipSuccess;
    hipLaunchKernelGGL(lincomb_batch_kernel, dim3((unsigned)((max_len + 255) / 256), count), dim3(256), 0, s, recs);
    return hipGetLastError();
}

__global__ void level_shift_batch_kernel(const ShiftRec *__restrict__ recs, double sigma) {
    const ShiftRec r = recs[blockIdx.y];
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < r.len) r.F[x] = r.F[x] + sigma * (r.S[x] - 0.5 * r.SPS[x]);
}
hipError_t level_shift_batch(const ShiftRec *recs, int count, int64_t max_len, double sigma, hipStream_t s) {
    if (count <= 0 || max_len <= 0) return hipSuccess;
    hipLaunchKernelGGL(level_shift_batch_kernel, dim3((unsigned)((max_len + 255) / 256), count), dim3(256), 0, s, recs,
                       sigma);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void energy_trace_batch_kernel(const TraceRec *__restrict__ recs) {
    const TraceRec r = recs[blockIdx.x];
    energy_trace_kernel_body<256>(r.n, r.P, r.H, r.G, r.out);
}
hipError_t energy_trace_batch(const TraceRec *recs, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(energy_trace_batch_kernel, dim3(count), dim3(256), 0, s, recs);
    return hipGetLastError();
}

}  // namespace fock
