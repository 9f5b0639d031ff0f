/*********************************************************************************
 * Sediment.h -- C++ drop-in for the reference's Sediment class over the m70s C ABI.
 *
 * Same call surface as the reference (include/Sediment.h:17-115): construct with
 * Sediment(Hydro*, Wave*, Bathymetry*, Grid*), call calculateSedimentFluxes() and
 * updateBottom(), read the public arrays Alpha, gamma, Gamma, qx, qy, dh_dt. The work runs
 * on the GPU on the hydro handle's device state, as left by the last Hydro::Run. The
 * public arrays are host mirrors refreshed after every call; updateBottom() also refreshes
 * bathy->h, bathy->Zb and hydro->Zs, which the reference's updateBottom writes
 * (src/Sediment.cpp:126-148), so the next Hydro::Run uploads the moved Zs.
 *
 * The reference reads Parameters::Bottom::p, Parameters::Sediment::* and
 * Parameters::Run::dt at call time; here they live in m70::SedimentConfig
 * (sedimentConfig()), read at construction (coefficients) and at every updateBottom (dt).
 * *******************************************************************************/
#ifndef M70_HOST_SEDIMENT_H_
#define M70_HOST_SEDIMENT_H_

#include <string>

#include "Hydro.h"

namespace m70 {

/* Defaults of src/OptionsParametersConstants.cpp (Bottom::p, Sediment::CWS::*,
 * Sediment::Swash::*, Run::dt). */
struct SedimentConfig {
    m70s_params params{0.4, 0.002, 0.0, 4.0, 10.0, 0.5, 1.0, 0.3};
    double dt = 0.1;
};
SedimentConfig &sedimentConfig();

}  // namespace m70

class Sediment {
public:
    double *Alpha;   // stirring
    double *gamma;   // downslope
    double *Gamma;   // swash
    double *qx;      // cross-shore transport
    double *qy;      // alongshore transport
    double *dh_dt;   // bed evolution

    Sediment(Hydro *hydro, Wave *wave, Bathymetry *bathy, Grid *grid);
    ~Sediment();

    /* Sediment::calculateSedimentFluxes (include/Sediment.h:76). */
    void calculateSedimentFluxes();
    /* Sediment::updateBottom (include/Sediment.h:81). */
    void updateBottom();
    /* Sediment::setSedimentFluxesFromResumeFile (include/Sediment.h:60; src/Sediment.cpp:77-103):
     * qx and qy of an m70_*.nc snapshot and the dh_dt they imply. Throws
     * std::runtime_error when the file or a variable cannot be read. */
    void setSedimentFluxesFromResumeFile(std::string resumeFile);
    /* Its arithmetic alone, on fluxes the caller already holds. */
    void setSedimentFluxes(const double *qx_in, const double *qy_in);

private:
    Hydro *hydro;
    Wave *wave;
    Bathymetry *bathy;
    Grid *grid;
    m70s *s_ = nullptr;
    void check(int rc) const;
    void pull();
};

#endif
