// wave.h -- device arguments and launchers of the wave and rollers driver (class Wave,
// class WaveRow: src/Wave*.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/m70h.h"

namespace m70 {

// One WaveRow (include/WaveRow.h:22-75): per-node arrays of one cross-shore position.
struct WRow {
    double *D, *lK, *U, *V, *U_w, *TanTheta, *CosTheta, *SinTheta, *Theta, *Sigma, *C, *Cg, *dw, *df, *sxx, *sxy,
        *syy, *sxx_r, *sxy_r, *syy_r, *lE, *E, *Dw, *H, *E_r, *d_r, *Cx, *Cy, *Cgx, *Cgy, *Urms_H, *inv_S_i, *Y_i, *G,
        *Gp, *WC, *WC_r;
};
constexpr int WROW_ARRAYS = 37;

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
    WRow rows[2];
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
    unsigned long long *acc;    // [3] bit patterns of max |tan(theta)| of a row
    int *anyv;                  // [3] the row had a non-NaN |tan(theta)|
};

// Wave::calculateWaveField (src/Wave.cpp:80-172): the first row (with ctl[0], acc[0]),
// then marching steps s0 .. s0+count-1, one launch each. The caller zeroes both row
// buffers, acc and anyv before the first row, launches at least nx-1 steps and reads
// ctl[next & 1] to see whether more are needed (intermediate rows).
hipError_t launch_wave_first(hipStream_t st, const WaveArgs &a, double hOff, double thetaOff, double T,
                             const WaveCtlBufs &b);
// WaveRow::dispersion (src/WaveParametrizations.cpp:14-60) of every mesh node into a.disp.
hipError_t launch_wave_dispersion(hipStream_t st, const WaveArgs &a, double T);
hipError_t launch_wave_steps(hipStream_t st, const WaveArgs &a, double T, int s0, int count, const WaveCtlBufs &b);
// Wave::setUV (src/Wave.cpp:293-302): U, V copied in; U_w from the stored direction.
hipError_t launch_wave_set_uv(hipStream_t st, const WaveArgs &a, double *U, double *V, double *U_w,
                              const double *Usrc, const double *Vsrc, long n);

}  // namespace m70
