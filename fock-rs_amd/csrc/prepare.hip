// prepare.hip -- device side of a basis preparation: the primitive-pair arrays are reordered,
// converted and laid out on the GPU from the one copy the host uploads, instead of being rebuilt
// on the host and uploaded again in every form.
//
//   permute_prims_kernel  primitive pairs of every shell pair in descending screening factor q
//                         (the host computes the stable order from q; src[e] is the old index of
//                         the pair's e-th primitive pair), with their exponent pairs
//   prims_to_f32_kernel   the FP32 copy the mixed form's bra side reads
//   build_soa_kernel      the ket-side bucket structure-of-arrays copy: field f of primitive k of
//                         a bucket's i-th pair at soa[start + (k * PRIM_FIELDS + f) * nb], start =
//                         bucket base + i, nb = the bucket's pair count (FP64 and, when asked, FP32)
#include "kernels.hpp"

namespace fock {

__global__ void permute_prims_kernel(int64_t n, const int *__restrict__ src, const PrimPair<double> *__restrict__ in,
                                     const double2 *__restrict__ ab_in, PrimPair<double> *__restrict__ out,
                                     double2 *__restrict__ ab_out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = src[e];
    out[e] = in[s];
    ab_out[e] = ab_in[s];
}

__global__ void prims_to_f32_kernel(int64_t n, const PrimPair<double> *__restrict__ in, PrimPair<float> *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const PrimPair<double> d = in[e];
    out[e] = PrimPair<float>{(float)d.p,   (float)d.Px,  (float)d.Py,  (float)d.Pz,  (float)d.Kc,  (float)d.xax,
                             (float)d.xay, (float)d.xaz, (float)d.xbx, (float)d.xby, (float)d.xbz, (float)d.inv2p};
}

// one thread per primitive pair e: the 13 fields of e into the column of its shell pair j
// (prim_pair[e], sorted pair index) as primitive k = prim_k[e] of that pair
__global__ void build_soa_kernel(int64_t n, const int *__restrict__ prim_pair, const int *__restrict__ prim_k,
                                 const PairInfo *__restrict__ sinfo, const int *__restrict__ stride,
                                 const PrimPair<double> *__restrict__ prims, const float *__restrict__ q,
                                 double *__restrict__ soa64, float *__restrict__ soa32) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int j = prim_pair[e], k = prim_k[e];
    const PrimPair<double> pp = prims[e];
    const double f[PRIM_FIELDS] = {pp.p,   pp.Px,  pp.Py,  pp.Pz,  pp.Kc,    pp.xax,
                                   pp.xay, pp.xaz, pp.xbx, pp.xby, pp.xbz, pp.inv2p,
                                   q ? (double)q[e] : 0.0};
    const int64_t nb = stride[j];
    const int64_t base = (int64_t)sinfo[j].pp_start + (int64_t)k * PRIM_FIELDS * nb;
#pragma unroll
    for (int x = 0; x < PRIM_FIELDS; ++x) {
        soa64[base + x * nb] = f[x];
        if (soa32) soa32[base + x * nb] = (float)f[x];
    }
}

hipError_t permute_prims(int64_t n, const int *src, const PrimPair<double> *in, const double2 *ab_in,
                         PrimPair<double> *out, double2 *ab_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(permute_prims_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, in, ab_in, out,
                       ab_out);
    return hipGetLastError();
}

hipError_t prims_to_f32(int64_t n, const PrimPair<double> *in, PrimPair<float> *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(prims_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, in, out);
    return hipGetLastError();
}

hipError_t build_soa(int64_t n, const int *prim_pair, const int *prim_k, const PairInfo *sinfo, const int *stride,
                     const PrimPair<double> *prims, const float *q, double *soa64, float *soa32, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(build_soa_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, prim_pair, prim_k,
                       sinfo, stride, prims, q, soa64, soa32);
    return hipGetLastError();
}

}  // namespace fock
