// outer.h -- device kernels of the morphodynamic outer loop (Morfo70::Run).
#pragma once
#include <hip/hip_runtime.h>

namespace m70 {

// The six fields whose largest change per outer step the reference prints
// (src/Morfo70.cpp:184-199): E, E_r, U_c, V_c, Zs, h.
struct OuterDiffArgs {
    const double *cur[6];
    double *prev[6];
    long n;
    unsigned long long *out;   // [6] bit patterns of the maxima
};

// Zeroes out[0..5], then out[q] = max |prev_q - cur_q| and prev_q = cur_q.
hipError_t launch_outer_diff(hipStream_t st, const OuterDiffArgs &a);

}  // namespace m70
