// adjoint_kernels.h -- launchers of the reverse Hydro ADI step (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "adjoint_body.h"

namespace m70 {

// Zeroes A and the accumulated intermediate adjoints of W (UN, UcN, gD).
hipError_t launch_adj_zero(hipStream_t st, const AdjFields &A, const AdjWork &W, long F, long Fp);
// Copies the ten fields of a state (F = nx*ny, Fp = (nx+1)*ny).
hipError_t launch_copy_state(hipStream_t st, const State &dst, const State &src, long F, long Fp);

// Reverse of solveQy (Y5..Y1): reads Ain (adjoint of the state after the step), adds the
// y half-step's contributions to A (adjoint of the state before it), W and R.
hipError_t launch_adj_y(hipStream_t st, const Consts &k, const Metrics &m, const State &S0, const State &S1,
                        const RunFields &rf, const StepTape &T, const AdjFields &Ain, const AdjFields &A,
                        const AdjWork &W, const AdjRun &R, int tile);
// Reverse of solveQx (X5..X1); f0 is the step's offshore flux offset.
hipError_t launch_adj_x(hipStream_t st, const Consts &k, const Metrics &m, const State &S0, const State &S1,
                        const RunFields &rf, const StepTape &T, const AdjFields &Ain, const AdjFields &A,
                        const AdjWork &W, const AdjRun &R, double f0, int tile);
// Reverse of the per-Run pre-loop: per-Run field adjoints R -> Dbar (depth at Run start)
// and params_out[0..2] += adjoints of (cD0, nu0, M).
hipError_t launch_adj_preloop(hipStream_t st, const Consts &k, const Metrics &m, const double *D0, const Forcing &fw,
                              const AdjRun &R, double *Dbar, double *params_out);

}  // namespace m70
