// sediment.h -- device arguments and launchers of the sediment transport kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/m70h.h"

namespace m70 {

struct SedArgs {
    int nx, ny;
    double Dmin;                          // Parameters::Hydro::Dmin
    double inv_dy;                        // 1/dy
    const double *inv_dx, *inv_dxx;       // 1/dx[i], 1/dx_x[i]
    m70s_params p;
    // hydro state and wave forcing (read; Zs is moved by the bed update)
    const double *D, *U, *V, *Dw;
    double *Zs;
    // sediment fields (include/Sediment.h:21-26) and the bed
    double *Alpha, *gamma, *Gamma, *qx, *qy, *dh_dt, *h, *Zb, *maxDw, *Zs_sh;
    // per-chunk partial results of the column pass: SED_MAX_CHUNKS x ny each
    double *part_max;
    int *part_last;
};

constexpr int SED_MAX_CHUNKS = 128;
// Chunks of rows of the column pass for nx rows (at most SED_MAX_CHUNKS, at least 8
// rows each); *rows_per_chunk receives the chunk height.
int sed_row_chunks(int nx, int *rows_per_chunk);

// Sediment::calculateSedimentFluxes (src/Sediment.cpp:47-65) is the two launches below,
// in order: the per-column running max of Dw and shoreline scan, then the per-cell
// coefficients with the fluxes.
hipError_t launch_sed_columns(hipStream_t st, const SedArgs &a);
hipError_t launch_sed_coeffs_fluxes(hipStream_t st, const SedArgs &a);
// Sediment::updateBottom (src/Sediment.cpp:67-75,105-151); dt = Parameters::Run::dt.
hipError_t launch_sed_update(hipStream_t st, const SedArgs &a, double dt);
// dh_dt from the fluxes in a.qx, a.qy (src/Sediment.cpp:77-103).
hipError_t launch_sed_dhdt(hipStream_t st, const SedArgs &a);

}  // namespace m70
