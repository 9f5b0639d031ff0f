/*********************************************************************************
 * Hydro.h -- C++ drop-in for the reference's Hydro class over the m70h C ABI.
 *
 * Same call surface as the reference (include/Hydro.h:24-124): construct with
 * Hydro(Grid*, Bathymetry*, Wave*), call Run(nIter), read or write the public arrays
 * U, U_c, V, V_c, Zs, Qx, Qy, dQx_dx, dQy_dy and bathy->D between runs. The arrays are
 * host mirrors of the device state: Run() uploads them (and bathy->D and the wave
 * forcing), runs nIter steps on the GPU and downloads them again, as the reference's
 * callers expect (src/Morfo70.cpp:141-153, src/Sediment.cpp:128-148).
 *
 * The reference reads its parameters from the globals Parameters::Hydro::* and
 * Options::Hydro::* at call time; here they live in m70::HydroConfig (hydroConfig()),
 * read by the constructor and at every Run().
 * *******************************************************************************/
#ifndef M70_HOST_HYDRO_H_
#define M70_HOST_HYDRO_H_

#include <string>
#include <vector>

#include "../../include/m70h.h"

namespace m70 {

/* Parameters and model switches (include/OptionsParametersConstants.h:49-113), with the
 * reference's defaults (src/OptionsParametersConstants.cpp:164-212). */
struct HydroConfig {
    m70h_params params{M70H_DRAG_LOGDEPTH, M70H_VISC_DEPTH, M70H_BC_SPONGE, 0.1, 0.01, 1.16, 0.01, 1e-4, 1.0,
                       0.0219, 0.001, 1.0};
    int solver = M70H_SOLVER_FUSED;
};
HydroConfig &hydroConfig();

}  // namespace m70

#ifdef M70_USE_REFERENCE_CLASSES
/* Linked into the reference: its own Grid, Bathymetry and Wave (include/Grid.h:21-54,
 * include/Bathymetry.h:22-53, include/Wave.h:36-53) provide the members read below. */
#include "Grid.h"
#include "Bathymetry.h"
#include "Wave.h"
#else
/* Mesh (include/Grid.h:21-54): node coordinates x[nx], spacing dy, sizes. */
class Grid {
public:
    double *x;
    double *dx;
    double dy;
    int nx;
    int ny;
    int size;
    Grid(const std::vector<double> &x, int ny, double dy);
    ~Grid();
};

/* Bathymetry (include/Bathymetry.h:22-53): Z, Zb = Z - h, h, D = max(Zb + zs0, Dmin). */
class Bathymetry {
public:
    double *Z;
    double *Zb;
    double *h;
    double *D;
    Bathymetry(const double *Z, const double *h, double zs0, Grid *grid);
    ~Bathymetry();

private:
    int size;
};

/* The wave members Hydro reads (include/Wave.h:36-53). */
class Wave {
public:
    double *H;
    double *Urms;
    double *Dw;
    double *D_r;
    double *Sxx;
    double *Sxy;
    double *Syy;
    explicit Wave(Grid *grid);
    ~Wave();
};
#endif /* M70_USE_REFERENCE_CLASSES */

class Hydro {
public:
    double *U;       // cross-shore current at x-faces, nx+1 rows
    double *U_c;     // at centres
    double *V;       // alongshore current at y-faces, nx+1 rows
    double *V_c;     // at centres
    double *Zs;      // free surface
    double *Qx;      // x flux at x-faces, nx+1 rows
    double *Qy;      // y flux at y-faces, nx+1 rows
    double *dQx_dx;
    double *dQy_dy;

    Hydro(Grid *grid, Bathymetry *bathy, Wave *wave);
    ~Hydro();

    /* Hydro::Run (include/Hydro.h:124). Throws std::runtime_error on a library error. */
    void Run(int nIter);

    /* Hydro::setHydroFromResumeFile (include/Hydro.h:117; src/Hydro.cpp:204-227): the
     * state fields of an m70_*.nc snapshot into the public arrays and bathy->D (U_c from
     * "U", U from "U0" and "U_x", Qx from "Qx0" and "Qx_x", V_c, V, Qy alike, Zs, D,
     * dQx_dx, dQy_dy). The next Run uploads them. Throws std::runtime_error when the
     * file or a variable cannot be read. */
    void setHydroFromResumeFile(std::string resumeFile);

    /* Time-varying offshore flux offset f0 (tide), one value per step from now on. */
    void setOffshoreF0Series(const std::vector<double> &f0);

    m70h *handle() { return h_; }

private:
    Grid *grid;
    Bathymetry *bathy;
    Wave *wave;
    m70h *h_ = nullptr;
    void check(int rc) const;
};

#endif
