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

struct WaveArgs {
    int nx, ny;
    double dy, inv_2dy;
    const double *x, *dx;              // node positions, mesh widths
    m70w_params p;
    const double *D, *U, *V, *U_w;     // depth and currents, nx*ny
    WField f;
    WRow rows[2];
    long *status;                      // [0] marching steps, [1] 0 or an error code
};

// Wave::calculateWaveField (src/Wave.cpp:80-172) on one workgroup: the rows march
// offshore to onshore with the reference's control flow (mesh rows and intermediate rows).
hipError_t launch_wave_field(hipStream_t st, const WaveArgs &a, double hOff, double thetaOff, double T);
// Wave::setUV (src/Wave.cpp:293-302): U, V copied in; U_w from the stored direction.
hipError_t launch_wave_set_uv(hipStream_t st, const WaveArgs &a, double *U, double *V, double *U_w,
                              const double *Usrc, const double *Vsrc, long n);

}  // namespace m70
