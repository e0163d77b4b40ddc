// boys_kernel.hip -- batched Boys function F_n(T); replaces boys::boys (boys/src/lib.rs:58-81).
//
// ACCURATE: each workgroup stages the slice of the Taylor table it needs (orders
// n .. n+7 at the 451 grid points T_i = 0.1 i) in LDS once, then evaluates
// F_n(T) = sum_{k<8} F_{n+k}(T_i) (T_i - T)^k / k! about the nearest grid point
// for T < 45, and upward recursion from sqrt(pi/T)/2 beyond.
// REF_COMPAT: the reference's algorithm (interpolated F_10 table, downward recursion,
// asymptotic form for T >= 30) with its 200-entry table in LDS.
#include "fock_device.hpp"
#include "kernels.hpp"

namespace fock {

constexpr int BOYS_SLICE = BOYS_TAYLOR64;  // orders n .. n+7

__global__ __launch_bounds__(256) void boys_accurate_kernel(const double *__restrict__ T, int64_t nT, int n,
                                                            const double *__restrict__ tab, double *__restrict__ out) {
    __shared__ double slice[BOYS_GRID_N * BOYS_SLICE];
    for (int idx = threadIdx.x; idx < BOYS_GRID_N * BOYS_SLICE; idx += blockDim.x) {
        const int i = idx / BOYS_SLICE, k = idx % BOYS_SLICE;
        slice[idx] = tab[i * BOYS_ORDERS + n + k];
    }
    __syncthreads();
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nT; e += (int64_t)gridDim.x * blockDim.x) {
        const double t = T[e];
        double r;
        if (t < BOYS_TMAX) {
            const int i = (int)fmax(t * 10.0 + 0.5, 0.0);  // host rejects t < 0
            const double d = (double)i * BOYS_GRID_STEP - t;
            const double *row = slice + i * BOYS_SLICE;
            double acc = row[BOYS_SLICE - 1];
#pragma unroll
            for (int k = BOYS_SLICE - 1; k > 0; --k) acc = fma(acc, d * (1.0 / k), row[k - 1]);
            r = acc;
        } else {
            // F_0 = sqrt(pi/T)/2 (erfc(sqrt 45) < 1e-20), then the exact upward recursion
            // F_m = ((2m-1) F_{m-1} - exp(-T)) / (2T), stable for T > n
            const double rs = rsqrt(t);
            const double e = exp(-t);
            const double h = 0.5 * rs * rs;
            r = 0.886226925452758 * rs;
            for (int m = 1; m <= n; ++m) r = ((2 * m - 1) * r - e) * h;
        }
        out[e] = r;
    }
}

__global__ __launch_bounds__(256) void boys_ref_kernel(const double *__restrict__ T, int64_t nT, int n,
                                                       const double *__restrict__ lut, double *__restrict__ out) {
    __shared__ double lds[200];
    for (int i = threadIdx.x; i < 200; i += blockDim.x) lds[i] = lut[i];
    __syncthreads();
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nT; e += (int64_t)gridDim.x * blockDim.x) {
        double F[11];
        boys_ref<10>(T[e], lds, F);
        // the reference's asymptotic branch is evaluated per order; boys_ref<10> computes
        // every order with the same expression, so F[n] is exactly boys(t, n)
        double r = F[0];
#pragma unroll
        for (int m = 1; m <= 10; ++m)
            if (m == n) r = F[m];
        out[e] = r;
    }
}

hipError_t launch_boys(const double *T, int64_t nT, int n, int policy, const BoysTables &bt, double *out,
                       hipStream_t s) {
    if (nT <= 0) return hipSuccess;
    int64_t blocks = (nT + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (policy == 0) {
        hipLaunchKernelGGL(boys_ref_kernel, dim3((unsigned)blocks), dim3(256), 0, s, T, nT, n, bt.lut_ref, out);
    } else {
        hipLaunchKernelGGL(boys_accurate_kernel, dim3((unsigned)blocks), dim3(256), 0, s, T, nT, n, bt.tab64, out);
    }
    return hipGetLastError();
}

}  // namespace fock
