// adjoint_batch.hip -- cell stages of the grouped adjoint (adjoint_batch.h): the taped
// forward step's assemblies and tails and the reverse step's per-cell bodies for every
// scenario of a batch in one launch. The bodies are those of the single-scenario kernels
// (assembly.h, tails.h, adjoint_body.h); a block covers 256 consecutive columns of one
// grid row of one scenario, so a wavefront is 64 consecutive columns of one row, as the
// reverse bodies' lane shuffles require.
#include "adjoint_batch.h"
#include "assembly.h"
#include "tails.h"

namespace m70 {

template <int OP>
__global__ void __launch_bounds__(256) k_adj_cells_grouped(cAdjScen *D, const int *__restrict__ start,
                                                           int nscen)
{
    const int b = blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const AdjScen &d = *(const AdjScen *)(D + sid);
    const Consts &k = d.k;
    const int ny = k.ny;
    const int gx = (ny + 255) / 256;
    const int lb = b - start[sid];
    const int i = lb / gx;
    const int j = (lb - i * gx) * 256 + (int)threadIdx.x;
    if (j >= ny || i >= agc_rows(OP, k.nx)) return;
    const long r = (long)i * ny;
    const long F = (long)k.nx * ny, Fp = F + ny;
    if constexpr (OP == AGC_COPY) {
        const long q = r + j;
        if (q < F) {
            d.dst.Zs[q] = d.src.Zs[q]; d.dst.D[q] = d.src.D[q]; d.dst.Uc[q] = d.src.Uc[q]; d.dst.Vc[q] = d.src.Vc[q];
            d.dst.dQx[q] = d.src.dQx[q]; d.dst.dQy[q] = d.src.dQy[q];
        }
        if (q < Fp) {
            d.dst.Qx[q] = d.src.Qx[q]; d.dst.Qy[q] = d.src.Qy[q]; d.dst.U[q] = d.src.U[q]; d.dst.V[q] = d.src.V[q];
        }
    } else if constexpr (OP == AGC_ZERO) {
        const long q = r + j;
        const AdjFields &A = d.Aout;
        if (q < F) {
            A.Zs[q] = 0; A.D[q] = 0; A.Uc[q] = 0; A.Vc[q] = 0; A.dQx[q] = 0; A.dQy[q] = 0;
            d.W.UcN[q] = 0; d.W.gD[q] = 0;
        }
        if (q < Fp) {
            A.Qx[q] = 0; A.Qy[q] = 0; A.U[q] = 0; A.V[q] = 0;
            d.W.UN[q] = 0;
        }
    } else if constexpr (OP == AGC_QXX_ASM || OP == AGC_QXY_ASM || OP == AGC_QX_TAIL) {
        // x half-step of the taped step: the tape holds the mask, drag and guess depth
        Work w = d.w;
        w.gD = d.T.gD;
        w.w = d.T.wx;
        w.drag = d.T.dragx;
        w.f = d.T.f;
        if constexpr (OP == AGC_QXX_ASM) {
            w.a = d.T.a1; w.b = d.T.b1; w.c = d.T.c1;
            asm_qxx(k, d.m, d.s, d.rf, w, d.f0, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
        } else if constexpr (OP == AGC_QXY_ASM) {
            w.a = d.T.a2; w.b = d.T.b2; w.c = d.T.c2;
            State sx = d.s;
            sx.Qx = d.T.qxs - ny;   // the predictor's solution: unknown i is row i+1
            asm_qxy(k, sx, d.rf, w, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
        } else {
            w.w = d.T.wy;           // the tail writes the Qy predictor's mask
            qx_tail_body(k, d.m, d.s, w, d.f0, i, j);
        }
    } else if constexpr (OP == AGC_QYY_ASM || OP == AGC_QYX_ASM) {
        Work w = d.w;
        w.gD = d.T.gD;
        w.w = d.T.wy;
        w.drag = d.T.dragy;
        w.f = d.T.f;
        if constexpr (OP == AGC_QYY_ASM) {
            w.a = d.T.a3; w.b = d.T.b3; w.c = d.T.c3;
            asm_qyy(k, d.m, d.s, d.rf, w, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
        } else {
            w.a = d.T.a4; w.b = d.T.b4; w.c = d.T.c4;
            State sy = d.s;
            sy.Qy = d.T.qys - ny;
            asm_qyx(k, d.m, sy, d.rf, w, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
        }
    } else if constexpr (OP == AGC_QY_TAIL) {
        qy_tail_body(k, d.s, i, j);
    } else if constexpr (OP == AGC_QY_TAIL_A) {
        adj_qy_tail_a(k, d.s, d.Ain, d.Aout, d.W, i, j);
    } else if constexpr (OP == AGC_QY_TAIL_B) {
        adj_qy_tail_b(k, d.Ain, d.W, i, j);
    } else if constexpr (OP == AGC_QYX) {
        adj_qyx(k, d.m, d.s, d.rf, d.T, d.W, d.R, i, j);
    } else if constexpr (OP == AGC_QYY) {
        adj_qyy(k, d.m, d.S0, d.s, d.rf, d.T, d.Aout, d.W, d.R, i, j);
    } else if constexpr (OP == AGC_QX_TAIL_A) {
        adj_qx_tail_a(k, d.m, d.s, d.T, d.Ain, d.Aout, d.W, i, j);
    } else if constexpr (OP == AGC_QX_TAIL_B) {
        adj_qx_tail_b(k, d.m, d.Ain, d.Aout, d.W, i, j);
    } else if constexpr (OP == AGC_QXY) {
        adj_qxy(k, d.S0, d.s, d.rf, d.T, d.Aout, d.W, d.R, i, j);

    }
}

// The Qx predictor's reverse of every scenario by runs of ADJ_RUN rows of a column, as the
// single-scenario k_adj_qxx_rows: workgroups of 64 x 4 threads, each scenario's workgroups
// tiling ceil(ny/64) columns by ceil(runs/4) run groups.
__global__ void __launch_bounds__(256) k_adj_qxx_rows_grouped(cAdjScen *D, const int *__restrict__ start, int nscen)
{
    const int b = blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const AdjScen &d = *(const AdjScen *)(D + sid);
    const int ny = d.k.ny, n = d.k.nx - 1;
    const int gx = (ny + 63) / 64;
    const int lb = b - start[sid];
    const int by = lb / gx, bx = lb - by * gx;
    const int j = bx * 64 + (int)threadIdx.x;
    const int run = __builtin_amdgcn_readfirstlane(by * 4 + (int)threadIdx.y);
    const int k0 = run * ADJ_RUN;
    if (j >= ny || k0 >= n) return;
    const int k1 = k0 + ADJ_RUN < n ? k0 + ADJ_RUN : n;
    adj_qxx_rows(d.k, d.m, d.S0, d.rf, d.T, d.Aout, d.W, d.R, d.f0, k0, k1, j);
}

int adj_qxx_rows_blocks(int nx, int ny)
{
    const int runs = (nx - 1 + ADJ_RUN - 1) / ADJ_RUN;
    return ((ny + 63) / 64) * ((runs + 3) / 4);
}

hipError_t launch_adj_qxx_rows_grouped(hipStream_t st, const AdjScen *D, const int *start, int nscen, int blocks)
{
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_qxx_rows_grouped, dim3((unsigned)blocks), dim3(64, 4), 0, st, (cAdjScen *)D, start, nscen);
    return hipGetLastError();
}

template <int OP>
static void launch_one(hipStream_t st, const AdjScen *D, const int *start, int nscen, int blocks)
{
    hipLaunchKernelGGL(k_adj_cells_grouped<OP>, dim3((unsigned)blocks), dim3(256), 0, st, (cAdjScen *)D, start, nscen);
}

hipError_t launch_adj_cells_grouped(hipStream_t st, int op, const AdjScen *D, const int *start, int nscen, int blocks)
{
    if (blocks <= 0) return hipSuccess;
    switch (op) {
    case AGC_COPY: launch_one<AGC_COPY>(st, D, start, nscen, blocks); break;
    case AGC_ZERO: launch_one<AGC_ZERO>(st, D, start, nscen, blocks); break;
    case AGC_QXX_ASM: launch_one<AGC_QXX_ASM>(st, D, start, nscen, blocks); break;
    case AGC_QXY_ASM: launch_one<AGC_QXY_ASM>(st, D, start, nscen, blocks); break;
    case AGC_QX_TAIL: launch_one<AGC_QX_TAIL>(st, D, start, nscen, blocks); break;
    case AGC_QYY_ASM: launch_one<AGC_QYY_ASM>(st, D, start, nscen, blocks); break;
    case AGC_QYX_ASM: launch_one<AGC_QYX_ASM>(st, D, start, nscen, blocks); break;
    case AGC_QY_TAIL: launch_one<AGC_QY_TAIL>(st, D, start, nscen, blocks); break;
    case AGC_QY_TAIL_A: launch_one<AGC_QY_TAIL_A>(st, D, start, nscen, blocks); break;
    case AGC_QY_TAIL_B: launch_one<AGC_QY_TAIL_B>(st, D, start, nscen, blocks); break;
    case AGC_QYX: launch_one<AGC_QYX>(st, D, start, nscen, blocks); break;
    case AGC_QYY: launch_one<AGC_QYY>(st, D, start, nscen, blocks); break;
    case AGC_QX_TAIL_A: launch_one<AGC_QX_TAIL_A>(st, D, start, nscen, blocks); break;
    case AGC_QX_TAIL_B: launch_one<AGC_QX_TAIL_B>(st, D, start, nscen, blocks); break;
    case AGC_QXY: launch_one<AGC_QXY>(st, D, start, nscen, blocks); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace m70
