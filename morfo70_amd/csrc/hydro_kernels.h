// hydro_kernels.h -- host launchers of the Hydro ADI kernels (internal to libm70h.so).
#pragma once
#include <hip/hip_runtime.h>
#include "hydro_device.h"

namespace m70 {

hipError_t launch_drag_viscosity(hipStream_t st, const Consts &k, const double *D, const Forcing &fw,
                                 const RunFields &rf);
hipError_t launch_run_fields(hipStream_t st, const Consts &k, const Metrics &m, const Forcing &fw,
                             const RunFields &rf);
hipError_t launch_qxx_assemble(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                               const RunFields &rf, const Work &w, double f0);
hipError_t launch_qxy_assemble(hipStream_t st, const Consts &k, const State &s, const RunFields &rf,
                               const Work &w);
// write_uc = false skips the U_c store: the step's own readers form U_c from U, so only
// the last step of a Run needs it in memory. rev (here and in the FUSED sweeps below):
// take the rows last to first; results do not depend on it. stream (here and in the FUSED
// sweeps below): the streaming cache policy (stream_policy); results do not depend on it.
hipError_t launch_qx_tail(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const Work &w,
                          double f0, bool write_uc = true, int rev = 0, bool stream = false);
hipError_t launch_qyy_assemble(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                               const RunFields &rf, const Work &w);
hipError_t launch_qyx_assemble(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                               const RunFields &rf, const Work &w);
// write_vc = false skips the V_c store, for steps whose successors in the Run form V_c from V.
hipError_t launch_qy_tail(hipStream_t st, const Consts &k, const State &s, bool write_vc = true, int rev = 0,
                          bool stream = false);
hipError_t launch_nonfinite(hipStream_t st, const double *x, long n, int *flag);
// *flag = 0 by a one-thread kernel (capturable in a graph, unlike a memset node).
hipError_t launch_flag_clear(hipStream_t st, int *flag);
hipError_t launch_init_zs(hipStream_t st, double *Zs, const double *D, const double *Z, const double *hb, long n);

hipError_t launch_thomas_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                              double *Y, double *V, int n, int m, long ld);
int periodic_rows_padded_lines(int nL);
hipError_t launch_periodic_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                long ld, double *Y, long ldy, double *Sr, double *Sv, double *Sk, int nL, int n);

bool partition_supported(int n);
hipError_t launch_partition_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                 double *Y, int n, int m, long ld, int tile);
hipError_t launch_partition_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                 long ld, double *Y, long ldy, int nL, int n, bool periodic, int tile);

// Transposed solves of the adjoint (A^T x = F for the systems A stored in A, B, C).
hipError_t launch_transposed_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                  double *Y, int n, int m, long ld, int tile);
hipError_t launch_transposed_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                  long ld, double *Y, long ldy, int nL, int n, int tile);

int cols_segments(int n);
enum ColsPass { COLS_ELIM = 0, COLS_REDUCED = 1, COLS_BACKSUB = 2 };
// vc_from_v: the Qx sweeps form V_c from V (exact once a step of this Run has stored both;
// see launch_qy_tail), so they need no V_c in memory.
hipError_t launch_cols3_qxx(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                            const Work &w, double f0, const ColsWork &cw, int tile, int pass, bool vc_from_v = false,
                            int rev = 0, bool stream = false);
hipError_t launch_cols3_qyx(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                            const Work &w, const ColsWork &cw, int tile, int pass, int rev = 0, bool stream = false);
bool rows_lds_supported(int n);
int rows_block_for(int n, int rb);
// Qx projector rows reading the Qx predictor's solution from its elimination and reduced
// passes (E0..E2, X of cw) in place of a back-substitution pass.
hipError_t launch_lds_qxy_bs(hipStream_t st, const Consts &k, const State &s, const RunFields &rf, const Work &w,
                             const ColsWork &cw, int rb, bool vc_from_v = false, int rev = 0, bool stream = false);
hipError_t launch_lds_qyy(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                          const Work &w, int rb, int rev = 0, bool stream = false);

// grouped (multi-scenario) launches; rev*: take the blocks last to first (the last
// scenario's last rows first); results do not depend on it
int grouped_cell_blocks(int nx, int ny);
// op 0: drag and viscosity, op 1: the per-Run fields
hipError_t launch_cells_grouped(hipStream_t st, int op, const ScenDesc *D, const int *start, int nscen, int blocks);
// op 2: the Qx tail, op 3: the Qy tail. write_derived = false skips the U_c / V_c stores (a
// batched Run's steps before its last form them from U and V).
int grouped_tail_blocks(int nx, int ny, int op);
hipError_t launch_tails_grouped(hipStream_t st, int op, const ScenDesc *D, const int *start, int nscen, int blocks,
                                int rev = 0, bool write_derived = true);
void grouped_stage_blocks(int nx, int ny, int rb, int *rows, int *cols, int *red, int *redp, int *S_rows, int *S_red,
                          int *L_red);
// vc_from_v: the Qx sweeps form V_c from V (a batched Run's steps after its first)
hipError_t launch_rows_grouped(hipStream_t st, int which, int rb, const ScenDesc *D, const int *start, int nscen,
                               int blocks, int rev = 0, bool vc_from_v = false);
hipError_t launch_cols3_grouped(hipStream_t st, int which, const ScenDesc *D, const int *start_cols,
                                const int *start_red, const int *start_redp, int nscen, int blocks_cols, int blocks_red,
                                int blocks_redp, bool backsub = true, int rev_elim = 0, int rev_backsub = 0,
                                bool vc_from_v = false, bool const_drag = false);

}  // namespace m70
