// compact.hip -- density-weighted screening and deterministic compaction of quartet tasks.
//
// Incremental Fock builds contract G(dP) for dP = P_k - P_{k-1}; most quartets then contribute
// nothing. Every (bra pair, ket pair) lane of the static Schwarz task list is tested against
//   Q_ab Q_cd max(2|dP_ab|, 2|dP_cd|, |dP_ac|, |dP_ad|, |dP_bc|, |dP_bd|)
// (shell-block maxima of |dP|; Q = sqrt of the largest diagonal integral of a pair) and
// classified: dropped (< tau), FP32 (tau <= bound < tau64) or FP64 (>= tau64).
// Survivors of each class are compacted with exclusive scans into dense tasks of up to 64 kets
// that share one bra pair and one primitive-count bucket, so every wavefront stays full and
// uniform. Scans and scatters are order preserving: the lists are identical run to run.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace fock {

__global__ void screen_flags_kernel(ScreenArgs a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (int64_t)a.ntasks * 64) return;
    const int t = (int)(s >> 6), l = (int)(s & 63);
    const Task task = a.tasks[t];
    int f32 = 0, f64 = 0;
    if (l < task.ket_count) {
        const int ket = task.ket_begin + l;
        const int2 ab = a.pair_shell[task.bra], cd = a.pair_shell[ket];
        const float *Pm = a.Pm;
        const int ns = a.nshell;
        float pm = 2.0f * fmaxf(Pm[ab.x * ns + ab.y], Pm[cd.x * ns + cd.y]);
        pm = fmaxf(pm, fmaxf(Pm[ab.x * ns + cd.x], Pm[ab.x * ns + cd.y]));
        pm = fmaxf(pm, fmaxf(Pm[ab.y * ns + cd.x], Pm[ab.y * ns + cd.y]));
        const float bound = a.pairQ[task.bra] * a.pairQ[ket] * pm;
        if (bound >= a.tau64) f64 = 1;
        else if (bound >= a.tau) f32 = 1;
    }
    a.flag32[s] = f32;
    a.flag64[s] = f64;
}

__global__ void fill_int_kernel(int64_t len, int v, int *__restrict__ out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < len) out[x] = v;
}

// survivors -> compacted ket indices; marks the survivors that open a new task
__global__ void scatter_kernel(const Task *__restrict__ tasks, int ntasks, const int *__restrict__ gfirst,
                               const int *__restrict__ glast, const int *__restrict__ flag,
                               const int *__restrict__ pos, int *__restrict__ kets, int *__restrict__ head,
                               int *__restrict__ kbra, int *__restrict__ kend, int *__restrict__ kcost) {
    const int64_t S = (int64_t)ntasks * 64;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !flag[s]) return;
    const int t = (int)(s >> 6), l = (int)(s & 63);
    const Task task = tasks[t];
    const int i = pos[s];
    const int64_t total = (int64_t)pos[S - 1] + flag[S - 1];
    const int64_t g0 = (int64_t)gfirst[t] * 64, ge = ((int64_t)glast[t] + 1) * 64;
    const int gstart = pos[g0];
    const int gend = ge < S ? pos[ge] : (int)total;
    kets[i] = task.ket_begin + l;
    head[i] = ((i - gstart) % 64) == 0 ? 1 : 0;
    kbra[i] = task.bra;
    kend[i] = gend;
    kcost[i] = task.flags;
}

__global__ void make_tasks_kernel(int64_t S, const int *__restrict__ flag, const int *__restrict__ pos,
                                  const int *__restrict__ head, const int *__restrict__ tid,
                                  const int *__restrict__ kbra, const int *__restrict__ kend,
                                  const int *__restrict__ kcost, Task *__restrict__ out, int *__restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)pos[S - 1] + flag[S - 1];
    if (i == 0) *count = total == 0 ? 0 : tid[total - 1] + head[total - 1];
    if (i >= total || !head[i]) return;
    const int cnt = kend[i] - (int)i;
    out[tid[i]] = Task{kbra[i], (int)i, cnt < 64 ? cnt : 64, kcost[i]};
}

hipError_t screen_flags(const ScreenArgs &a, hipStream_t s) {
    const int64_t S = (int64_t)a.ntasks * 64;
    if (S <= 0) return hipSuccess;
    hipLaunchKernelGGL(screen_flags_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

size_t compact_temp_bytes(int64_t S) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int *)nullptr, (int *)nullptr, (int)S);
    return bytes;
}

hipError_t compact_tasks(const Task *tasks, int ntasks, const int *gfirst, const int *glast, const int *flag,
                         const CompactBuffers &b, hipStream_t s) {
    const int64_t S = (int64_t)ntasks * 64;
    if (S <= 0) {
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, (int64_t)1, 0, b.count);
        return hipGetLastError();
    }
    const dim3 grid((unsigned)((S + 255) / 256)), block(256);
    size_t bytes = b.tmp_bytes;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(b.tmp, bytes, flag, b.pos, (int)S, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fill_int_kernel, grid, block, 0, s, S, 0, b.head);
    hipLaunchKernelGGL(scatter_kernel, grid, block, 0, s, tasks, ntasks, gfirst, glast, flag, b.pos, b.kets, b.head,
                       b.kbra, b.kend, b.kcost);
    bytes = b.tmp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(b.tmp, bytes, b.head, b.tid, (int)S, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(make_tasks_kernel, grid, block, 0, s, S, flag, b.pos, b.head, b.tid, b.kbra, b.kend, b.kcost,
                       b.tasks, b.count);
    return hipGetLastError();
}

}  // namespace fock
