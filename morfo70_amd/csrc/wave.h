// wave.h -- device arguments and launchers of the wave and rollers driver (class Wave,
// class WaveRow: src/Wave*.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/m70h.h"

namespace m70 {

// The arrays of a WaveRow (include/WaveRow.h:22-75) that the next marching step reads as
// the previous row, ny each: array q of row buffer r at rows[r] + q*ny. The friction term
// is kept only with bottom friction on, the rollers' arrays only with rollers on (a buffer
// is zeroed before each field, and a term that is off is zero).
enum WRowArr {
    WR_H, WR_lE, WR_TanTheta, WR_lK, WR_Cg, WR_U, WR_V, WR_dw, WR_WC, WR_Cgx, WR_Cgy,   // always
    WR_df,                                                                           // friction
    WR_C, WR_WC_r, WR_Cx, WR_Cy, WR_E_r, WR_d_r, WR_Dw,                               // rollers
    WROW_ARRAYS
};

// The public arrays of class Wave (include/Wave.h:33-53), nx*ny each.
struct WField {
    double *K, *Theta, *Cg, *H, *E, *Urms, *Sxx_w, *Sxy_w, *Syy_w, *Dw, *E_r, *D_r, *Sxx_r, *Sxy_r, *Syy_r, *Sxx,
        *Sxy, *Syy, *CosTheta, *SinTheta;
};
constexpr int WFIELD_ARRAYS = 20;

// Dispersion of every mesh node (nx*ny each), solved for the whole field before the march:
// a mesh row's depth and current are known up front, only intermediate rows need their
// own (interpolated) solve.
struct WDisp {
    double *lK, *Sigma, *C, *Cg, *Urms_H;
};

struct WaveArgs {
    int nx, ny;
    double dy, inv_2dy;
    const double *x, *dx;              // node positions, mesh widths
    m70w_params p;
    const double *D, *U, *V, *U_w;     // depth and currents, nx*ny
    WField f;
    double *rows[2];                   // row buffer sets: WKMAX rows of WROW_ARRAYS x ny each
    WDisp disp;
};

// Workgroups of WB threads (one wavefront). The first row: WNI nodes per workgroup, one per
// thread. Marching launches: up to WKMAX rows each; lane t of workgroup k holds node
// (k*WNB + t - WHALO) mod ny for every row of the launch, and owns it when
// WHALO <= t < WHALO + WNB. A row's node needs the previous row at its neighbours and their
// direction at the next ones, so the lanes whose values are exact shrink by one each side
// per row: WHALO = WKMAX + 1 halo lanes keep the owned ones exact through WKMAX rows.
constexpr int WB = 64;
constexpr int WNI = WB - 4;
constexpr int WKMAX = 12;
constexpr int WHALO = WKMAX + 1;
constexpr int WNB = WB - 2 * WHALO;
int wave_blocks(int ny);        // first-row workgroups
int wave_blocks_march(int ny);  // marching workgroups

// Marching control. Launch L (L >= 1; the first row is launch 0) writes ctl[L & 1]: the
// state after its first row (x0, mesh row, step count, done), which follows the
// reference's step rule (src/Wave.cpp:267-290) from the rows before it, and the number of
// rows K it computes. Its rows 2..K are taken to be mesh rows; launch L+1 checks that
// against the rule, from the largest |tan(theta)| of each row, and continues after the
// last row that holds.
struct WaveCtl {
    double x0;      // cross-shore position of the row
    long steps;     // marching steps so far
    int xIndex;     // mesh row of the last mesh row reached
    int done;       // xIndex == nx-1 (or the step limit was hit)
    int err;        // the march exceeded 1000*nx + 1000 steps
    int K;          // rows launch L computes (0: the march was already done)
};
struct WaveCtlBufs {
    WaveCtl *ctl;               // [2]
    // [3][WKMAX + 1]: set L % 3 slot r holds 1 + the bit pattern of the largest |tan(theta)|
    // of launch L's row r (0: no value)
    unsigned long long *acc;
};
constexpr int WACC_SLOTS = WKMAX + 1;

// Wave::calculateWaveField (src/Wave.cpp:80-172): the first row (launch 0), then marching
// launches L0 .. L0+count-1. The caller zeroes acc before the first row, and reads
// ctl[last L & 1] to see whether more launches are needed. The march stores
// Theta, H, E, Dw and E_r of every mesh row; once it is done, launch_wave_field_terms
// forms the field's other arrays.
hipError_t launch_wave_first(hipStream_t st, const WaveArgs &a, double hOff, double thetaOff, double T,
                             const WaveCtlBufs &b);
// WaveRow::dispersion (src/WaveParametrizations.cpp:14-60) of every mesh node into a.disp.
hipError_t launch_wave_dispersion(hipStream_t st, const WaveArgs &a, double T);
hipError_t launch_wave_steps(hipStream_t st, const WaveArgs &a, double T, int s0, int count, const WaveCtlBufs &b);
// Marching launches a field of nx mesh rows takes when every step is a mesh row (the
// rows per launch double from 2 to WKMAX), plus the one that confirms the last rows.
int wave_march_launches(int nx);
// The rest of Wave::storageRow (src/Wave.cpp:175-264) over the whole field: CosTheta,
// SinTheta, Cg, K, the radiation stresses, D_r and Urms from the stored arrays and a.disp.
hipError_t launch_wave_field_terms(hipStream_t st, const WaveArgs &a);
// Wave::setUV (src/Wave.cpp:293-302): U, V copied in; U_w from the stored direction.
hipError_t launch_wave_set_uv(hipStream_t st, const WaveArgs &a, double *U, double *V, double *U_w,
                              const double *Usrc, const double *Vsrc, long n);

}  // namespace m70
