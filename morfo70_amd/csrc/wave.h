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
    double *rows[2];                   // row buffers, WROW_ARRAYS x ny each
    WDisp disp;
};

// Workgroups of WB threads own WNI consecutive nodes of a row (two halo nodes each side).
constexpr int WB = 64;
constexpr int WNI = WB - 4;
int wave_blocks(int ny);

// Marching state after a step; double-buffered (launch s reads ctl[s&1], writes
// ctl[(s+1)&1]).
struct WaveCtl {
    double x0;      // cross-shore position of the last row
    long steps;     // marching steps so far
    int xIndex;     // mesh row of the last mesh row reached
    int done;       // xIndex == nx-1 (or the step limit was hit)
    int err;        // the march exceeded 1000*nx + 1000 steps
};
struct WaveCtlBufs {
    WaveCtl *ctl;               // [2]
    unsigned long long *acc;    // [3] 1 + bit pattern of max |tan(theta)| of a row; 0: no value
};

// Wave::calculateWaveField (src/Wave.cpp:80-172): the first row (with ctl[0], acc[0]),
// then marching steps s0 .. s0+count-1, one launch each. The caller zeroes both row
// buffers and acc before the first row, launches at least nx-1 steps and reads
// ctl[next & 1] to see whether more are needed (intermediate rows). The march stores
// Theta, H, E, Dw and E_r of every mesh row; once it is done, launch_wave_field_terms
// forms the field's other arrays.
hipError_t launch_wave_first(hipStream_t st, const WaveArgs &a, double hOff, double thetaOff, double T,
                             const WaveCtlBufs &b);
// WaveRow::dispersion (src/WaveParametrizations.cpp:14-60) of every mesh node into a.disp.
hipError_t launch_wave_dispersion(hipStream_t st, const WaveArgs &a, double T);
hipError_t launch_wave_steps(hipStream_t st, const WaveArgs &a, double T, int s0, int count, const WaveCtlBufs &b);
// The rest of Wave::storageRow (src/Wave.cpp:175-264) over the whole field: CosTheta,
// SinTheta, Cg, K, the radiation stresses, D_r and Urms from the stored arrays and a.disp.
hipError_t launch_wave_field_terms(hipStream_t st, const WaveArgs &a);
// Wave::setUV (src/Wave.cpp:293-302): U, V copied in; U_w from the stored direction.
hipError_t launch_wave_set_uv(hipStream_t st, const WaveArgs &a, double *U, double *V, double *U_w,
                              const double *Usrc, const double *Vsrc, long n);

}  // namespace m70
