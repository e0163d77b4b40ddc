// mo_grid.hip -- molecular orbitals evaluated at many points: the renderers' loop.
//
// psi_k(r) = sum_f C_fk phi_f(r) over the basis functions f whose coefficient has |C_fk| > 0.05
// (MolecularOrbitals::new, scf/src/molecular_orbitals.rs:69-90; MolecularOrbital::evaluate, :37-43),
// phi_f(r) = sum_p coef_p d_x^i d_y^j d_z^k exp(-alpha_p |d|^2) with d = A_f - r
// (ContractedGaussian::evaluate, basis/src/contracted_gaussian.rs:320-331).
//
// One thread per point, FP64 arithmetic. The host groups each orbital's kept functions by shell
// (shared centre and exponents): a thread computes |d|^2 and the K exponentials of a shell once and
// every kept function of the shell as poly_f(d) * sum_p w_fp e_p, with w_fp = C_fk coef_p folded in
// on the host. Primitives with alpha |d|^2 > MO_EXP_CUT are skipped (exp below 1e-52: under FP32's
// smallest subnormal, and 1e-50 of any orbital value that matters in FP64). The orbital's records
// are uniform over the workgroup (orbital = blockIdx.y), so they are read as broadcast loads.
#include <mutex>

#include "kernels.hpp"

namespace fock {

namespace {

constexpr double MO_EXP_CUT = 120.0;

// Records of the allocation-free entry point (fock_mo_grid_device_noalloc): one fixed array per
// device, held until the process ends. Host code never passes `scratch` itself to a kernel: its
// per-device address comes from hipGetSymbolAddress (mo_scratch_address).
static __device__ __align__(256) unsigned char scratch[MO_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void *g_scratch_ptr[MO_SCRATCH_MAX_DEVICES] = {};

__device__ inline double ipow(double x, int e) {
    double r = 1.0;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

template <bool GRID, typename Out>
__global__ __launch_bounds__(256) void mo_eval_kernel(MoArgs a) {
    const int o = blockIdx.y;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.npoints) return;
    double rx, ry, rz;
    if constexpr (GRID) {
        const int64_t nxy = (int64_t)a.dims[0] * a.dims[1];
        const int z = (int)(v / nxy);
        const int64_t rem = v - (int64_t)z * nxy;
        const int y = (int)(rem / a.dims[0]);
        const int x = (int)(rem - (int64_t)y * a.dims[0]);
        rx = a.origin[0] + x * a.step[0];
        ry = a.origin[1] + y * a.step[1];
        rz = a.origin[2] + z * a.step[2];
    } else {
        rx = a.points[3 * v];
        ry = a.points[3 * v + 1];
        rz = a.points[3 * v + 2];
    }
    double acc = 0.0;
    const int s1 = a.shell_off[o + 1];
    for (int s = a.shell_off[o]; s < s1; ++s) {
        const MoShellRec sh = a.shells[s];
        const double dx = sh.A[0] - rx, dy = sh.A[1] - ry, dz = sh.A[2] - rz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (sh.amin * d2 > MO_EXP_CUT) continue;  // every primitive of the shell is negligible
        double e[MO_MAXK];
        const double *alpha = a.dbl + sh.alpha;
#pragma unroll
        for (int p = 0; p < MO_MAXK; ++p) {
            e[p] = 0.0;
            if (p < sh.K) {
                const double t = alpha[p] * d2;
                if (t <= MO_EXP_CUT) e[p] = exp(-t);
            }
        }
        for (int f = sh.fun; f < sh.fun + sh.nf; ++f) {
            const MoFunRec fr = a.funs[f];
            const double *w = a.dbl + fr.w;
            double rad = 0.0;
#pragma unroll
            for (int p = 0; p < MO_MAXK; ++p)
                if (p < sh.K) rad = fma(w[p], e[p], rad);
            acc = fma(ipow(dx, fr.ang[0]) * ipow(dy, fr.ang[1]) * ipow(dz, fr.ang[2]), rad, acc);
        }
    }
    static_cast<Out *>(a.out)[(int64_t)o * a.out_stride + v] = (Out)acc;
}

}  // namespace

hipError_t mo_scratch_address(int device, void **out) {
    if (device < 0 || device >= MO_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_mo_eval(const MoArgs &a, bool grid, bool f32, hipStream_t s) {
    if (a.npoints <= 0 || a.norb <= 0) return hipSuccess;
    const dim3 g((unsigned)((a.npoints + 255) / 256), (unsigned)a.norb), b(256);
    if (grid) {
        if (f32) hipLaunchKernelGGL((mo_eval_kernel<true, float>), g, b, 0, s, a);
        else hipLaunchKernelGGL((mo_eval_kernel<true, double>), g, b, 0, s, a);
    } else {
        if (f32) hipLaunchKernelGGL((mo_eval_kernel<false, float>), g, b, 0, s, a);
        else hipLaunchKernelGGL((mo_eval_kernel<false, double>), g, b, 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace fock
