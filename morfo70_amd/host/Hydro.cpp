// Hydro.cpp -- C++ drop-in Hydro over the m70h C ABI (see Hydro.h).
#include "Hydro.h"

#include <algorithm>
#include <stdexcept>

namespace m70 {
HydroConfig &hydroConfig()
{
    static HydroConfig cfg;
    return cfg;
}
}  // namespace m70

#ifndef M70_USE_REFERENCE_CLASSES
Grid::Grid(const std::vector<double> &xs, int ny_, double dy_)
{
    nx = (int)xs.size();
    ny = ny_;
    dy = dy_;
    size = nx * ny;
    x = new double[nx];
    std::copy(xs.begin(), xs.end(), x);
    dx = new double[nx > 1 ? nx - 1 : 1];
    for (int i = 0; i < nx - 1; i++) dx[i] = x[i + 1] - x[i];   // src/Grid.cpp:32-34
}

Grid::~Grid()
{
    delete[] x;
    delete[] dx;
}

Bathymetry::Bathymetry(const double *Z_, const double *h_, double zs0, Grid *grid)
{
    size = grid->size;
    Z = new double[size];
    Zb = new double[size];
    h = new double[size];
    D = new double[size];
    const double Dmin = m70::hydroConfig().params.Dmin;
    for (int i = 0; i < size; i++) {   // src/Bathymetry.cpp:33-39
        Z[i] = Z_[i];
        h[i] = h_[i];
        Zb[i] = Z[i] - h[i];
        double d = Zb[i] + zs0;
        if (d < Dmin) d = Dmin;
        D[i] = d;
    }
}

Bathymetry::~Bathymetry()
{
    delete[] Z;
    delete[] Zb;
    delete[] h;
    delete[] D;
}

Wave::Wave(Grid *grid)
{
    const int n = grid->size;
    H = new double[n]();
    Urms = new double[n]();
    Dw = new double[n]();
    D_r = new double[n]();
    Sxx = new double[n]();
    Sxy = new double[n]();
    Syy = new double[n]();
}

Wave::~Wave()
{
    delete[] H; delete[] Urms; delete[] Dw; delete[] D_r; delete[] Sxx; delete[] Sxy; delete[] Syy;
}

#endif /* M70_USE_REFERENCE_CLASSES */

void Hydro::check(int rc) const
{
    if (rc != M70H_OK) throw std::runtime_error(std::string("m70h: ") + m70h_last_error(h_));
}

Hydro::Hydro(Grid *grid_, Bathymetry *bathy_, Wave *wave_) : grid(grid_), bathy(bathy_), wave(wave_)
{
    const m70::HydroConfig &cfg = m70::hydroConfig();
    int rc = m70h_create(grid->nx, grid->ny, grid->x, grid->dy, &cfg.params, &h_);
    if (rc != M70H_OK) throw std::runtime_error(std::string("m70h_create: ") + m70h_last_error(nullptr));
    check(m70h_set_solver(h_, cfg.solver));
    const int size = grid->size, ny = grid->ny;
    // src/Hydro.cpp:25-37
    U = new double[size + ny]();
    Qx = new double[size + ny]();
    U_c = new double[size]();
    V = new double[size + ny]();
    Qy = new double[size + ny]();
    V_c = new double[size]();
    dQx_dx = new double[size]();
    dQy_dy = new double[size]();
    Zs = new double[size]();
    check(m70h_init_from_bathymetry(h_, bathy->D, bathy->Z, bathy->h));
    check(m70h_get_state(h_, Zs, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
}

Hydro::~Hydro()
{
    m70h_destroy(h_);
    delete[] U; delete[] Qx; delete[] U_c; delete[] V; delete[] Qy; delete[] V_c;
    delete[] dQx_dx; delete[] dQy_dy; delete[] Zs;
}

void Hydro::setHydroFromResumeFile(std::string resumeFile)
{
    const long F = (long)grid->size, ny = grid->ny;
    const char *path = resumeFile.c_str();
    struct { const char *var; double *dst; long n; } reads[] = {
        {"U", U_c, F},          {"U0", U, ny},        {"Qx0", Qx, ny},       {"Qx_x", Qx + ny, F},
        {"U_x", U + ny, F},     {"V", V_c, F},        {"V0", V, ny},         {"Qy0", Qy, ny},
        {"Qy_y", Qy + ny, F},   {"V_y", V + ny, F},   {"Zs", Zs, F},         {"D", bathy->D, F},
        {"dQx_dx", dQx_dx, F},  {"dQy_dy", dQy_dy, F}};
    for (auto &r : reads)
        if (m70nc_read(path, r.var, r.dst, r.n) != M70H_OK)
            throw std::runtime_error(std::string("setHydroFromResumeFile: ") + m70nc_last_error());
}

void Hydro::setOffshoreF0Series(const std::vector<double> &f0)
{
    check(m70h_set_offshore_f0_series(h_, f0.data(), (long)f0.size()));
}

void Hydro::Run(int nIter)
{
    // callers may have changed Zs (src/Sediment.cpp:128,137,148), D or the forcing
    check(m70h_set_state(h_, Zs, bathy->D, Qx, Qy, U, V, U_c, V_c, dQx_dx, dQy_dy));
    check(m70h_set_params(h_, &m70::hydroConfig().params));   // read at call time, as the reference
    check(m70h_set_forcing(h_, wave->Sxx, wave->Sxy, wave->Syy, wave->Urms, wave->Dw, wave->D_r, wave->H));
    const int rc = m70h_run(h_, nIter);
    // a Run that ends with NaN/Inf in the state still leaves that state in the public
    // arrays, as the reference does; the facade then reports it
    if (rc == M70H_OK || rc == M70H_ENONFINITE)
        check(m70h_get_state(h_, Zs, bathy->D, Qx, Qy, U, V, U_c, V_c, dQx_dx, dQy_dy));
    check(rc);
}
