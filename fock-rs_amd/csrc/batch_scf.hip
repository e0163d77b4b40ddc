// batch_scf.hip -- the state-driven steps of the batched SCF iteration.
//
// Every step of one SCF iteration (scf/src/lib.rs:176-257) is one launch over all molecules of the
// batch, and every decision the host used to take between launches -- which molecules still
// iterate, the DIIS ring and its coefficients (scf/src/utils.rs:102-138), whether the level shift
// applies, the exit rule (scf/src/lib.rs:224-249), the precision phase of the mixed form -- is taken
// on the device and kept in each molecule's BatchState. The launch sequence is then the same every
// iteration, so the host captures it once as a HIP graph, replays it once per iteration and reads
// one counter (molecules still active) per iteration.
#include "kernels.hpp"

namespace fock {

namespace {

__device__ __forceinline__ bool normal_f64(double x) {
    const double a = fabs(x);
    return a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e308;  // false for NaN, inf, 0, subnormal
}

__device__ __forceinline__ void atomic_max_nonneg64(double *addr, double v) {
    // non-negative doubles order as their bit patterns (NaN wins)
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

__global__ void batch_begin_kernel(const BatchMolRec *__restrict__ recs, int *active_count) {
    const BatchMolRec r = recs[blockIdx.y];
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *active_count = 0;
    if (!r.st->active) return;  // the same for the whole workgroup
    const int64_t len = (int64_t)r.n * r.n;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) r.st->maxabs = 0.0;
    if (x < len) {
        r.Jb[x] = 0.0;
        r.Kb[x] = 0.0;
    }
}

__global__ void batch_fock_from_jk_kernel(const BatchMolRec *__restrict__ recs) {
    const BatchMolRec r = recs[blockIdx.y];
    if (!r.st->active) return;
    const int n = r.n;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    const int64_t t = (int64_t)j * n + i;
    const double g = (r.Jb[e] + r.Jb[t]) - 0.5 * (r.Kb[e] + r.Kb[t]);
    r.G[e] = g;
    r.F[e] = r.H[e] + g;
}

// e = FPS - (FPS)^T into ring slot next_slot, F into the same Fock slot, max |e| (NaN kept)
__global__ __launch_bounds__(256) void batch_diis_error_kernel(const BatchMolRec *__restrict__ recs) {
    __shared__ double red[256];
    const BatchMolRec r = recs[blockIdx.y];
    if (!r.st->active) return;  // the same for the whole workgroup
    const int n = r.n;
    const int64_t len = (int64_t)n * n;
    if ((int64_t)blockIdx.x * blockDim.x >= len) return;  // the same for the whole workgroup
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = r.st->next_slot;
    double v = 0.0;
    if (idx < len) {
        const int i = (int)(idx / n), j = (int)(idx % n);
        const double d = r.FPS[idx] - r.FPS[(int64_t)j * n + i];
        r.errs[(int64_t)slot * len + idx] = d;
        r.focks[(int64_t)slot * len + idx] = r.F[idx];
        v = fabs(d);
        if (d != d) v = d;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const double o = red[threadIdx.x + off], m = red[threadIdx.x];
            red[threadIdx.x] = (o != o || o > m) ? o : m;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg64(&r.st->maxabs, red[0]);
}

// <e_new, e_s> for the 16 ring slots: one workgroup per (slot, molecule), fixed reduction order
__global__ __launch_bounds__(256) void batch_diis_dots_kernel(const BatchMolRec *__restrict__ recs) {
    __shared__ double red[256];
    const BatchMolRec r = recs[blockIdx.y];
    if (!r.st->active) return;  // the same for the whole workgroup
    const int64_t len = (int64_t)r.n * r.n;
    const double *e = r.errs + (int64_t)r.st->next_slot * len;
    const double *b = r.errs + (int64_t)blockIdx.x * len;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < len; i += blockDim.x) acc += e[i] * b[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) r.st->dots[blockIdx.x] = red[0];
}

// One workgroup of 64 lanes per molecule. Lane 0 keeps the ring (window diis_window, the
// reference's 16), the workgroup solves the bordered system B c = (0, ..., 0, -1) by Gauss-Jordan
// elimination with partial pivoting in LDS (scf/src/utils.rs:102-138); a zero pivot or a non-finite
// coefficient falls back to the newest Fock matrix, as the reference does when try_inverse fails.
__global__ __launch_bounds__(64) void batch_diis_update_kernel(const BatchMolRec *__restrict__ recs, BatchParams p) {
    __shared__ double A[17][18];
    __shared__ int s_m, s_solve, s_piv;
    BatchState *st = recs[blockIdx.x].st;
    if (!st->active) return;  // the same for the whole workgroup
    const int t = threadIdx.x;
    if (t == 0) {
        const double err = st->maxabs;
        st->err = err;
        s_solve = 0;
        s_m = 0;
        if (!normal_f64(err) && !(err < p.eps)) {  // NaN or inf: the reference's None
            st->status = 2;                        // FOCK_DIVERGED
            st->iter_done = st->iter;
            st->active = 0;
        } else {
            const int slot = st->next_slot;
            st->next_slot = (slot + 1) % 16;
            int m = st->ring_len;
            if (m == p.diis_window) {  // drop the oldest pair
                for (int i = 0; i + 1 < m; ++i) {
                    st->ring[i] = st->ring[i + 1];
                    for (int j = 0; j + 1 < m; ++j) st->B[i * 16 + j] = st->B[(i + 1) * 16 + j + 1];
                }
                --m;
            }
            st->ring[m] = slot;
            for (int i = 0; i <= m; ++i) {
                const double d = st->dots[st->ring[i]];
                st->B[m * 16 + i] = d;
                st->B[i * 16 + m] = d;
            }
            ++m;
            st->ring_len = m;
            st->ncoef = 1;
            st->cslot[0] = slot;
            st->coef[0] = 1.0;
            const bool exit_plain = p.accurate && st->phase == 1 && err < p.eps;
            s_m = m;
            s_solve = (m >= p.diis_start && !exit_plain) ? 1 : 0;
            st->shift = (p.level_shift > 0.0 && err > p.level_shift_until) ? 1 : 0;
        }
    }
    __syncthreads();
    if (!s_solve) return;  // the same for the whole workgroup
    const int m = s_m, N = m + 1;
    for (int e = t; e < N * (N + 1); e += blockDim.x) {
        const int i = e / (N + 1), j = e % (N + 1);
        double v;
        if (j == N) v = (i == m) ? -1.0 : 0.0;
        else if (i < m && j < m) v = st->B[i * 16 + j];
        else if (i == m && j == m) v = 0.0;
        else v = -1.0;
        A[i][j] = v;
    }
    __syncthreads();
    bool ok = true;
    for (int col = 0; col < N; ++col) {
        if (t == 0) {
            int piv = col;
            for (int r = col + 1; r < N; ++r)
                if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
            s_piv = A[piv][col] == 0.0 ? -1 : piv;
        }
        __syncthreads();
        const int piv = s_piv;
        if (piv < 0) {
            ok = false;
            break;  // the same for the whole workgroup
        }
        if (piv != col)
            for (int k = t; k <= N; k += blockDim.x) {
                const double x = A[piv][k];
                A[piv][k] = A[col][k];
                A[col][k] = x;
            }
        __syncthreads();
        if (t < N && t != col) {
            const double f = A[t][col] / A[col][col];
            if (f != 0.0)
                for (int k = col; k <= N; ++k) A[t][k] -= f * A[col][k];
        }
        __syncthreads();
    }
    if (t == 0 && ok) {
        double c[16];
        bool finite = true;
        for (int i = 0; i < m; ++i) {
            c[i] = A[i][N] / A[i][i];
            finite = finite && (c[i] - c[i] == 0.0);
        }
        if (finite) {
            st->ncoef = m;
            for (int i = 0; i < m; ++i) {
                st->cslot[i] = st->ring[i];
                st->coef[i] = c[i];
            }
        }
    }
}

__global__ void batch_lincomb_kernel(const BatchMolRec *__restrict__ recs) {
    const BatchMolRec r = recs[blockIdx.y];
    const BatchState *st = r.st;
    if (!st->active) return;
    const int64_t len = (int64_t)r.n * r.n;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= len) return;
    double acc = st->coef[0] * r.focks[(int64_t)st->cslot[0] * len + x];
    for (int i = 1; i < st->ncoef; ++i) acc = acc + st->coef[i] * r.focks[(int64_t)st->cslot[i] * len + x];
    r.Fx[x] = acc;
}

__global__ void batch_level_shift_kernel(const BatchMolRec *__restrict__ recs, double sigma) {
    const BatchMolRec r = recs[blockIdx.y];
    if (!r.st->active || !r.st->shift) return;
    const int64_t len = (int64_t)r.n * r.n;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) r.Fx[x] = r.Fx[x] + sigma * (r.S[x] - 0.5 * r.FPS[x]);
}

__global__ void batch_copy_basis_kernel(const BatchMolRec *__restrict__ recs) {
    const BatchMolRec r = recs[blockIdx.y];
    if (!r.st->active) return;
    const int64_t len = (int64_t)r.n * r.n;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) r.E[x] = r.U[x];
}

__global__ __launch_bounds__(256) void batch_energy_kernel(const BatchMolRec *__restrict__ recs) {
    __shared__ double red[256];
    const BatchMolRec r = recs[blockIdx.x];
    if (!r.st->active) return;  // the same for the whole workgroup
    const int64_t len = (int64_t)r.n * r.n;
    double acc = 0.0;
    for (int64_t e = threadIdx.x; e < len; e += blockDim.x) acc += r.P[e] * (2.0 * r.H[e] + r.G[e]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) r.st->e_elec = 0.5 * red[0];
}

// The exit rule (scf/src/lib.rs:224-249) for every molecule still active after this iteration: the
// result returns when err < eps or at iteration max_iters (a build of the mixed form's FP32 phase
// is never returned: it hands over to FP64 builds near convergence, on stagnation, or three
// iterations before the limit).
__global__ void batch_exit_kernel(const BatchMolRec *__restrict__ recs, int count, BatchParams p, int *active_count) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= count) return;
    BatchState *st = recs[m].st;
    if (!st->active) return;
    const double err = st->err;
    const int iter = st->iter;
    if (p.mixed && st->phase == 0) {
        if (err < st->best_err * 0.5) {
            st->best_err = err;
            st->stall = 0;
        } else if (err < p.mixed_stall) {
            ++st->stall;
        }
        if (err < fmax(10.0 * p.eps, p.mixed_switch) || st->stall >= 3 || iter + 3 >= p.max_iters) st->phase = 1;
    } else if (err < p.eps || iter >= p.max_iters) {
        st->status = err < p.eps ? 0 : 1;  // FOCK_OK / FOCK_NOT_CONVERGED
        st->iter_done = iter;
        st->active = 0;
    }
    st->iter = iter + 1;
    if (st->active) atomicAdd(active_count, 1);
}

inline dim3 elems(int max_n, int count) {
    const int64_t len = (int64_t)max_n * max_n;
    return dim3((unsigned)((len + 255) / 256), (unsigned)count);
}

}  // namespace

hipError_t batch_begin(const BatchMolRec *recs, int count, int max_n, int *active_count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_begin_kernel, elems(max_n, count), dim3(256), 0, s, recs, active_count);
    return hipGetLastError();
}
hipError_t batch_fock_from_jk(const BatchMolRec *recs, int count, int max_n, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_fock_from_jk_kernel, elems(max_n, count), dim3(256), 0, s, recs);
    return hipGetLastError();
}
hipError_t batch_diis_error(const BatchMolRec *recs, int count, int max_n, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_diis_error_kernel, elems(max_n, count), dim3(256), 0, s, recs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_diis_dots_kernel, dim3(16, count), dim3(256), 0, s, recs);
    return hipGetLastError();
}
hipError_t batch_diis_update(const BatchMolRec *recs, int count, BatchParams p, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_diis_update_kernel, dim3(count), dim3(64), 0, s, recs, p);
    return hipGetLastError();
}
hipError_t batch_lincomb(const BatchMolRec *recs, int count, int max_n, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_lincomb_kernel, elems(max_n, count), dim3(256), 0, s, recs);
    return hipGetLastError();
}
hipError_t batch_level_shift(const BatchMolRec *recs, int count, int max_n, double sigma, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_level_shift_kernel, elems(max_n, count), dim3(256), 0, s, recs, sigma);
    return hipGetLastError();
}
hipError_t batch_copy_basis(const BatchMolRec *recs, int count, int max_n, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_copy_basis_kernel, elems(max_n, count), dim3(256), 0, s, recs);
    return hipGetLastError();
}
hipError_t batch_energy(const BatchMolRec *recs, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_energy_kernel, dim3(count), dim3(256), 0, s, recs);
    return hipGetLastError();
}
hipError_t batch_exit(const BatchMolRec *recs, int count, BatchParams p, int *active_count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_exit_kernel, dim3((count + 63) / 64), dim3(64), 0, s, recs, count, p, active_count);
    return hipGetLastError();
}

}  // namespace fock
