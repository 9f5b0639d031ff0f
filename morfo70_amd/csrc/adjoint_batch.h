// adjoint_batch.h -- grouped (multi-scenario) launches of the adjoint of Run: the taped
// forward step and the reverse step of many scenarios (uneven grids) in one launch per
// stage. Block b belongs to scenario scen_of(start, nscen, b); each scenario's blocks tile
// its own grid exactly as the single-scenario kernels do (256 threads along j per grid
// row; one row line per block; 1024-thread workgroups of column lines), so a scenario's
// results are those of its own launches.
#pragma once
#include <hip/hip_runtime.h>

#include "adjoint_body.h"

namespace m70 {

// Everything one scenario's stages read, from a device-side table.
struct AdjScen {
    Consts k;
    Metrics m;
    RunFields rf;
    State s;           // the scenario's current state
    State S0;          // state before the step being reversed
    State dst, src;    // state copy
    StepTape T;        // intermediates of the step being taped or reversed
    Work w;            // the handle's work arrays (the tape replaces gD, drag and the masks)
    AdjFields Ain;     // adjoint of the state after the step (input)
    AdjFields Aout;    // adjoint of the state before it (accumulated)
    AdjWork W;
    AdjRun R;
    double f0;         // offshore flux offset of the step
    int S_cols;        // segments per column line (partition solver, 8 unknowns each)
    int L_cols;        // column lines per 1024-thread workgroup
    int S_rows;        // segments per row line (LDS row solver)
    int pad;
};
// The kernels read the table through the constant address space (see cScenDesc).
typedef __attribute__((address_space(4))) const AdjScen cAdjScen;

// Cell stages. Rows: COPY and ZERO nx+1 (the F+ extent), the two tails and the tails'
// first reverses nx, the others nx-1.
enum AdjGroupedCell {
    AGC_COPY, AGC_ZERO,
    AGC_QXX_ASM, AGC_QXY_ASM, AGC_QX_TAIL, AGC_QYY_ASM, AGC_QYX_ASM, AGC_QY_TAIL,
    AGC_QY_TAIL_A, AGC_QY_TAIL_B, AGC_QYX, AGC_QYY, AGC_QX_TAIL_A, AGC_QX_TAIL_B, AGC_QXY,
    AGC_COUNT
};
// Grid rows a cell stage covers.
__host__ __device__ inline int agc_rows(int op, int nx)
{
    if (op == AGC_COPY || op == AGC_ZERO) return nx + 1;
    if (op == AGC_QX_TAIL || op == AGC_QY_TAIL || op == AGC_QY_TAIL_A || op == AGC_QX_TAIL_A) return nx;
    return nx - 1;
}

hipError_t launch_adj_cells_grouped(hipStream_t st, int op, const AdjScen *D, const int *start, int nscen, int blocks);
// The Qx predictor's reverse (the last reverse stage) by runs of rows of a column;
// adj_qxx_rows_blocks gives one scenario's workgroups.
int adj_qxx_rows_blocks(int nx, int ny);
hipError_t launch_adj_qxx_rows_grouped(hipStream_t st, const AdjScen *D, const int *start, int nscen, int blocks);
// Column line solves of every scenario (n = nx-1 unknowns, ny lines), 1024-thread
// workgroups: sys 1 (the Qx predictor, A1) or 4 (the Qy projector, A4); trans 0 solves
// the taped system (right-hand side T.f, solution onto the tape or the state), trans 1
// its transpose (W.rhs -> W.lam).
hipError_t launch_adj_cols_grouped(hipStream_t st, int sys, int trans, const AdjScen *D, const int *start, int nscen,
                                   int blocks);
// Periodic row line solves (n = ny, nx-1 lines), RB-thread LDS row workgroups: sys 2 (the
// Qx projector, A2) or 3 (the Qy predictor, A3), trans as above.
hipError_t launch_adj_rows_grouped(hipStream_t st, int rb, int sys, int trans, const AdjScen *D, const int *start,
                                   int nscen, int blocks);
// Segment counts the grouped solves use (0: the line is too long for them).
int adj_cols_segments(int n);
int adj_rows_segments(int n, int rb);

}  // namespace m70
