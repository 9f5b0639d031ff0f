// outer.hip -- device kernels of the morphodynamic outer loop (Morfo70::Run,
// src/Morfo70.cpp:131-218): the per-step maxima of the changes of E, E_r, U_c, V_c, Zs and
// h that the reference prints (:184-199), reduced on the device.
#include <hip/hip_runtime.h>

#include "outer.h"

namespace m70 {

constexpr int OD_BLOCK = 256;
constexpr int OD_FIELDS = 6;

// out[q] = max over cells of |prev_q - cur_q| (NaN differences ignored, as std::max(err, d)
// keeps err when d is NaN), then prev_q = cur_q. Maxima of non-negative doubles are
// combined as their bit patterns (the order is the same). cur_q may be NULL: field q
// unchanged (its difference stays 0).
__global__ void __launch_bounds__(OD_BLOCK) k_outer_diff(OuterDiffArgs a)
{
    __shared__ double smax[OD_FIELDS][OD_BLOCK];
    const int t = threadIdx.x;
    double m[OD_FIELDS];
#pragma unroll
    for (int q = 0; q < OD_FIELDS; ++q) m[q] = 0.0;
    for (long i = (long)blockIdx.x * OD_BLOCK + t; i < a.n; i += (long)gridDim.x * OD_BLOCK) {
#pragma unroll
        for (int q = 0; q < OD_FIELDS; ++q) {
            if (!a.cur[q]) continue;
            const double c = a.cur[q][i];
            const double d = fabs(a.prev[q][i] - c);
            m[q] = (m[q] < d) ? d : m[q];
            a.prev[q][i] = c;
        }
    }
#pragma unroll
    for (int q = 0; q < OD_FIELDS; ++q) smax[q][t] = m[q];
    __syncthreads();
    for (int o = OD_BLOCK / 2; o > 0; o >>= 1) {
        if (t < o) {
#pragma unroll
            for (int q = 0; q < OD_FIELDS; ++q)
                smax[q][t] = (smax[q][t] < smax[q][t + o]) ? smax[q][t + o] : smax[q][t];
        }
        __syncthreads();
    }
    if (t < OD_FIELDS) atomicMax(&a.out[t], (unsigned long long)__double_as_longlong(smax[t][0]));
}

__global__ void k_outer_zero(unsigned long long *out, int n)
{
    if ((int)threadIdx.x < n) out[threadIdx.x] = 0ull;
}

hipError_t launch_outer_diff(hipStream_t st, const OuterDiffArgs &a)
{
    hipLaunchKernelGGL(k_outer_zero, dim3(1), dim3(64), 0, st, a.out, OD_FIELDS);
    long blocks = (a.n + OD_BLOCK - 1) / OD_BLOCK;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_outer_diff, dim3((unsigned)blocks), dim3(OD_BLOCK), 0, st, a);
    return hipGetLastError();
}

}  // namespace m70
