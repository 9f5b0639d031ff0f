// adjoint.hip -- kernels of the reverse (adjoint) Hydro ADI step for gfx950.
//
// One thread per cell, blocks of 256 threads along j, one grid row per i, as the forward
// stencil kernels. The per-cell bodies are in adjoint_body.h; the taped forward step and
// the checkpoint schedule are driven from m70h_api.hip.
#include "adjoint_body.h"
#include "adjoint_kernels.h"
#include "hydro_kernels.h"

namespace m70 {

#define ADJ_CELL(nrows)                                  \
    const int j = blockIdx.x * blockDim.x + threadIdx.x; \
    const int i = blockIdx.y;                            \
    if (j >= k.ny || i >= (nrows)) return;

static dim3 cell_grid(const Consts &k, int rows) { return dim3((unsigned)((k.ny + 255) / 256), (unsigned)rows); }

__global__ void __launch_bounds__(256) k_adj_qy_tail_a(Consts k, State S1, AdjFields Ain, AdjFields A, AdjWork W)
{
    ADJ_CELL(k.nx)
    adj_qy_tail_a(k, S1, Ain, A, W, i, j);
}
__global__ void __launch_bounds__(256) k_adj_qy_tail_b(Consts k, AdjFields Ain, AdjWork W)
{
    ADJ_CELL(k.nx - 1)
    adj_qy_tail_b(k, Ain, W, i, j);
}
__global__ void __launch_bounds__(256) k_adj_qyx(Consts k, Metrics m, State S1, RunFields rf, StepTape T, AdjWork W,
                                                 AdjRun R)
{
    ADJ_CELL(k.nx - 1)
    adj_qyx(k, m, S1, rf, T, W, R, i, j);
}
__global__ void __launch_bounds__(256) k_adj_qyy(Consts k, Metrics m, State S0, State S1, RunFields rf, StepTape T,
                                                 AdjFields A, AdjWork W, AdjRun R)
{
    ADJ_CELL(k.nx - 1)
    adj_qyy(k, m, S0, S1, rf, T, A, W, R, i, j);
}
__global__ void __launch_bounds__(256) k_adj_qx_tail_a(Consts k, Metrics m, State S1, StepTape T, AdjFields Ain,
                                                       AdjFields A, AdjWork W)
{
    ADJ_CELL(k.nx)
    adj_qx_tail_a(k, m, S1, T, Ain, A, W, i, j);
}
__global__ void __launch_bounds__(256) k_adj_qx_tail_b(Consts k, Metrics m, AdjFields Ain, AdjFields A, AdjWork W)
{
    ADJ_CELL(k.nx - 1)
    adj_qx_tail_b(k, m, Ain, A, W, i, j);
}
__global__ void __launch_bounds__(256) k_adj_qxy(Consts k, State S0, State S1, RunFields rf, StepTape T, AdjFields A,
                                                 AdjWork W, AdjRun R)
{
    ADJ_CELL(k.nx - 1)
    adj_qxy(k, S0, S1, rf, T, A, W, R, i, j);
}
// The Qx predictor's reverse by runs of ADJ_RUN rows of a column (adj_qxx_rows): workgroups
// of 64 x 4 threads, a wavefront being 64 consecutive columns of one run.
__global__ void __launch_bounds__(256) k_adj_qxx_rows(Consts k, Metrics m, State S0, RunFields rf, StepTape T,
                                                      AdjFields A, AdjWork W, AdjRun R, double f0)
{
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int run = __builtin_amdgcn_readfirstlane((int)blockIdx.y * 4 + (int)threadIdx.y);
    const int n = k.nx - 1;
    const int k0 = run * ADJ_RUN;
    if (j >= k.ny || k0 >= n) return;
    const int k1 = k0 + ADJ_RUN < n ? k0 + ADJ_RUN : n;
    adj_qxx_rows(k, m, S0, rf, T, A, W, R, f0, k0, k1, j);
}
__global__ void __launch_bounds__(256) k_adj_run_fields(Consts k, Metrics m, AdjRun R)
{
    ADJ_CELL(k.nx)
    adj_run_fields(k, m, R, i, j);
}

// Depth adjoint of the drag/viscosity models plus a block reduction of the parameter
// adjoints (cD0, nu0, M) into out[0..2].
__global__ void __launch_bounds__(256) k_adj_drag_visc(Consts k, const double *__restrict__ D, Forcing fw, AdjRun R,
                                                       double *Dbar, double *out)
{
    __shared__ double red[3][256];
    const long n = (long)k.nx * k.ny;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    double c = 0.0, nu = 0.0, mm = 0.0;
    if (q < n) adj_drag_visc(k, D, fw, R, Dbar, q, c, nu, mm);
    red[0][threadIdx.x] = c;
    red[1][threadIdx.x] = nu;
    red[2][threadIdx.x] = mm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
            red[2][threadIdx.x] += red[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc(&out[0], red[0][0]);
        acc(&out[1], red[1][0]);
        acc(&out[2], red[2][0]);
    }
}

// Zeroes the ten state adjoints and the accumulators of the intermediates (one launch
// instead of thirteen fills).
__global__ void __launch_bounds__(256) k_adj_zero(AdjFields A, double *UN, double *UcN, double *gD, long F, long Fp)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < F) {
        A.Zs[q] = 0; A.D[q] = 0; A.Uc[q] = 0; A.Vc[q] = 0; A.dQx[q] = 0; A.dQy[q] = 0;
        UcN[q] = 0; gD[q] = 0;
    }
    if (q < Fp) {
        A.Qx[q] = 0; A.Qy[q] = 0; A.U[q] = 0; A.V[q] = 0;
        UN[q] = 0;
    }
}

// Copies the ten state fields (one launch instead of ten copies).
__global__ void __launch_bounds__(256) k_copy_state(State dst, State src, long F, long Fp)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < F) {
        dst.Zs[q] = src.Zs[q]; dst.D[q] = src.D[q]; dst.Uc[q] = src.Uc[q]; dst.Vc[q] = src.Vc[q];
        dst.dQx[q] = src.dQx[q]; dst.dQy[q] = src.dQy[q];
    }
    if (q < Fp) {
        dst.Qx[q] = src.Qx[q]; dst.Qy[q] = src.Qy[q]; dst.U[q] = src.U[q]; dst.V[q] = src.V[q];
    }
}

hipError_t launch_adj_zero(hipStream_t st, const AdjFields &A, const AdjWork &W, long F, long Fp)
{
    hipLaunchKernelGGL(k_adj_zero, dim3((unsigned)((Fp + 255) / 256)), dim3(256), 0, st, A, W.UN, W.UcN, W.gD, F, Fp);
    return hipGetLastError();
}

hipError_t launch_copy_state(hipStream_t st, const State &dst, const State &src, long F, long Fp)
{
    hipLaunchKernelGGL(k_copy_state, dim3((unsigned)((Fp + 255) / 256)), dim3(256), 0, st, dst, src, F, Fp);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------
hipError_t launch_adj_y(hipStream_t st, const Consts &k, const Metrics &m, const State &S0, const State &S1,
                        const RunFields &rf, const StepTape &T, const AdjFields &Ain, const AdjFields &A,
                        const AdjWork &W, const AdjRun &R, int tile)
{
    const int nx = k.nx, ny = k.ny;
    hipError_t e;
    hipLaunchKernelGGL(k_adj_qy_tail_a, cell_grid(k, nx), dim3(256), 0, st, k, S1, Ain, A, W);
    hipLaunchKernelGGL(k_adj_qy_tail_b, cell_grid(k, nx - 1), dim3(256), 0, st, k, Ain, W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // Y4: A4^T lam = rhs (columns)
    if ((e = launch_transposed_cols(st, T.a4, T.b4, T.c4, W.rhs, W.lam, nx - 1, ny, ny, tile)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_qyx, cell_grid(k, nx - 1), dim3(256), 0, st, k, m, S1, rf, T, W, R);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // Y2: A3^T lam = rhs (periodic rows)
    if ((e = launch_transposed_rows(st, T.a3, T.b3, T.c3, W.rhs, ny, W.lam, ny, nx - 1, ny, tile)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_adj_qyy, cell_grid(k, nx - 1), dim3(256), 0, st, k, m, S0, S1, rf, T, A, W, R);
    return hipGetLastError();
}

hipError_t launch_adj_x(hipStream_t st, const Consts &k, const Metrics &m, const State &S0, const State &S1,
                        const RunFields &rf, const StepTape &T, const AdjFields &Ain, const AdjFields &A,
                        const AdjWork &W, const AdjRun &R, double f0, int tile)
{
    const int nx = k.nx, ny = k.ny;
    hipError_t e;
    hipLaunchKernelGGL(k_adj_qx_tail_a, cell_grid(k, nx), dim3(256), 0, st, k, m, S1, T, Ain, A, W);
    hipLaunchKernelGGL(k_adj_qx_tail_b, cell_grid(k, nx - 1), dim3(256), 0, st, k, m, Ain, A, W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // X4: A2^T lam = rhs (periodic rows)
    if ((e = launch_transposed_rows(st, T.a2, T.b2, T.c2, W.rhs, ny, W.lam, ny, nx - 1, ny, tile)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_adj_qxy, cell_grid(k, nx - 1), dim3(256), 0, st, k, S0, S1, rf, T, A, W, R);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // X2: A1^T lam = rhs (columns)
    if ((e = launch_transposed_cols(st, T.a1, T.b1, T.c1, W.rhs, W.lam, nx - 1, ny, ny, tile)) != hipSuccess) return e;
    const int runs = (nx - 1 + ADJ_RUN - 1) / ADJ_RUN;
    hipLaunchKernelGGL(k_adj_qxx_rows, dim3((unsigned)((ny + 63) / 64), (unsigned)((runs + 3) / 4)), dim3(64, 4), 0, st, k,
                       m, S0, rf, T, A, W, R, f0);
    return hipGetLastError();
}

hipError_t launch_adj_preloop(hipStream_t st, const Consts &k, const Metrics &m, const double *D0, const Forcing &fw,
                              const AdjRun &R, double *Dbar, double *params_out)
{
    hipLaunchKernelGGL(k_adj_run_fields, cell_grid(k, k.nx), dim3(256), 0, st, k, m, R);
    const long n = (long)k.nx * k.ny;
    hipLaunchKernelGGL(k_adj_drag_visc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, k, D0, fw, R, Dbar,
                       params_out);
    return hipGetLastError();
}

}  // namespace m70
