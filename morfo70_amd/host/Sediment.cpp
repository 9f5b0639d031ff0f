// Sediment.cpp -- C++ drop-in Sediment over the m70s C ABI (see Sediment.h).
#include "Sediment.h"

#include <stdexcept>
#include <vector>

namespace m70 {
SedimentConfig &sedimentConfig()
{
    static SedimentConfig cfg;
    return cfg;
}
}  // namespace m70

void Sediment::check(int rc) const
{
    if (rc != M70H_OK) throw std::runtime_error(std::string("m70s: ") + m70s_last_error(s_));
}

Sediment::Sediment(Hydro *hydro_, Wave *wave_, Bathymetry *bathy_, Grid *grid_)
    : hydro(hydro_), wave(wave_), bathy(bathy_), grid(grid_)
{
    const m70::SedimentConfig &cfg = m70::sedimentConfig();
    int rc = m70s_create(hydro->handle(), bathy->h, bathy->Zb, &cfg.params, &s_);
    if (rc != M70H_OK) throw std::runtime_error(std::string("m70s_create: ") + m70s_last_error(nullptr));
    const int size = grid->size;   // src/Sediment.cpp:13-31
    qx = new double[size]();
    qy = new double[size]();
    dh_dt = new double[size]();
    Alpha = new double[size]();
    Gamma = new double[size]();
    gamma = new double[size]();
}

Sediment::~Sediment()
{
    m70s_destroy(s_);
    delete[] Alpha; delete[] Gamma; delete[] gamma; delete[] qx; delete[] qy; delete[] dh_dt;
}

void Sediment::pull()
{
    check(m70s_get(s_, Alpha, gamma, Gamma, qx, qy, dh_dt, bathy->h, bathy->Zb));
}

void Sediment::calculateSedimentFluxes()
{
    check(m70s_set_bed(s_, bathy->h, bathy->Zb));   // callers may have edited the bed
    check(m70s_calculate_fluxes(s_));
    pull();
}

void Sediment::updateBottom()
{
    check(m70s_set_bed(s_, bathy->h, bathy->Zb));
    check(m70s_update_bottom(s_, m70::sedimentConfig().dt));
    pull();
    check(m70h_get_state(hydro->handle(), hydro->Zs, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr));
}

void Sediment::setSedimentFluxesFromResumeFile(std::string resumeFile)
{
    const long F = (long)grid->size;
    std::vector<double> qx_in(F), qy_in(F);
    if (m70nc_read(resumeFile.c_str(), "qx", qx_in.data(), F) != M70H_OK ||
        m70nc_read(resumeFile.c_str(), "qy", qy_in.data(), F) != M70H_OK)
        throw std::runtime_error(std::string("setSedimentFluxesFromResumeFile: ") + m70nc_last_error());
    setSedimentFluxes(qx_in.data(), qy_in.data());
}

void Sediment::setSedimentFluxes(const double *qx_in, const double *qy_in)
{
    check(m70s_set_fluxes(s_, qx_in, qy_in));
    pull();
}
