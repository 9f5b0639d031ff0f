// facade_run.cpp -- drives the C++ Hydro facade the way the reference's main loop drives
// its Hydro (src/main.cpp:17-21, src/Morfo70.cpp:141-153): Grid, Bathymetry, Wave,
// Hydro(grid, bathy, wave), fill the forcing, Run(n) repeatedly, read the public arrays.
//
// With sediment = 1 it also drives the C++ Sediment after every Run, as the outer loop
// does inside the bed evolution window (src/Morfo70.cpp:155-163):
// calculateSedimentFluxes(), then updateBottom().
//
// With resume_runs > 0 it then writes the state as a resume file (the m70_*.nc variables
// Hydro::setHydroFromResumeFile and Sediment::setSedimentFluxesFromResumeFile read,
// src/Hydro.cpp:204-227, src/Sediment.cpp:77-103), builds a second Hydro (and Sediment)
// from that file, and runs both the original and the resumed Hydro resume_runs more
// times without sediment.
//
// Usage: facade_run <input.bin> <output.bin>
//   input:  int32 nx, ny, runs, steps_per_run, drag, visc, bc, solver, sediment, resume_runs;
//           float64 dy, zs0, dt, Dmin, alpha_f, cD0, nu0, M, ka, z0, alpha_r, bed_dt;
//           x[nx]; Z[nx*ny]; h[nx*ny]; Sxx, Sxy, Syy, Urms, Dw, D_r, H [nx*ny each]
//   output: Zs, D [nx*ny]; Qx, Qy, U, V [(nx+1)*ny]; U_c, V_c, dQx_dx, dQy_dy [nx*ny];
//           with sediment: Alpha, gamma, Gamma, qx, qy, dh_dt, h, Zb [nx*ny each];
//           with resume_runs > 0: the hydro fields above of the original, then of the
//           resumed Hydro, after the extra runs; with sediment also qx, qy, dh_dt of the
//           original Sediment when the file was written and of the resumed one just after
//           reading it
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "Hydro.h"
#include "Sediment.h"

static void rd(FILE *f, void *p, size_t sz, size_t n)
{
    if (fread(p, sz, n, f) != n) throw std::runtime_error("short read");
}

static void write_hydro(FILE *out, const Hydro &hy, const Bathymetry &b, size_t F, size_t ny)
{
    const size_t Fp = F + ny;
    fwrite(hy.Zs, sizeof(double), F, out);
    fwrite(b.D, sizeof(double), F, out);
    fwrite(hy.Qx, sizeof(double), Fp, out);
    fwrite(hy.Qy, sizeof(double), Fp, out);
    fwrite(hy.U, sizeof(double), Fp, out);
    fwrite(hy.V, sizeof(double), Fp, out);
    fwrite(hy.U_c, sizeof(double), F, out);
    fwrite(hy.V_c, sizeof(double), F, out);
    fwrite(hy.dQx_dx, sizeof(double), F, out);
    fwrite(hy.dQy_dy, sizeof(double), F, out);
}

// The resume variables of an m70_*.nc snapshot (src/Morfo70.cpp:225-276 writes them;
// src/Hydro.cpp:204-227 and src/Sediment.cpp:77-103 read them).
static void write_resume(const char *path, const Grid &grid, const Hydro &hy, const Bathymetry &b,
                         const Sediment *sed)
{
    const int nx = grid.nx, ny = grid.ny;
    std::vector<double> y(ny);
    for (int j = 0; j < ny; ++j) y[j] = j * grid.dy;
    std::vector<const char *> names = {"U", "U0", "U_x", "Qx0", "Qx_x", "V", "V0", "V_y", "Qy0", "Qy_y",
                                       "Zs", "D", "dQx_dx", "dQy_dy"};
    std::vector<const double *> data = {hy.U_c, hy.U, hy.U + ny, hy.Qx, hy.Qx + ny, hy.V_c, hy.V, hy.V + ny,
                                        hy.Qy, hy.Qy + ny, hy.Zs, b.D, hy.dQx_dx, hy.dQy_dy};
    std::vector<int> rows = {nx, 1, nx, 1, nx, nx, 1, nx, 1, nx, nx, nx, nx, nx};
    if (sed) {
        names.push_back("qx"); data.push_back(sed->qx); rows.push_back(nx);
        names.push_back("qy"); data.push_back(sed->qy); rows.push_back(nx);
    }
    const char *att = "time";
    const double t = 0.0;
    if (m70nc_write(path, nx, ny, grid.x, y.data(), (int)names.size(), names.data(), data.data(), rows.data(), 1,
                    &att, &t) != M70H_OK)
        throw std::runtime_error(std::string("resume file: ") + m70nc_last_error());
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    try {
        FILE *in = fopen(argv[1], "rb");
        if (!in) throw std::runtime_error("cannot open input");
        int hdr[10];
        rd(in, hdr, sizeof(int), 10);
        const int nx = hdr[0], ny = hdr[1], runs = hdr[2], steps = hdr[3], with_sed = hdr[8], resume_runs = hdr[9];
        double sc[12];
        rd(in, sc, sizeof(double), 12);
        m70::sedimentConfig().dt = sc[11];
        m70::HydroConfig &cfg = m70::hydroConfig();
        cfg.params = m70h_params{hdr[4], hdr[5], hdr[6], sc[2], sc[3], sc[4], sc[5], sc[6], sc[7], sc[8], sc[9], sc[10]};
        cfg.solver = hdr[7];
        const size_t F = (size_t)nx * ny;
        std::vector<double> x(nx), Z(F), h(F);
        rd(in, x.data(), sizeof(double), nx);
        rd(in, Z.data(), sizeof(double), F);
        rd(in, h.data(), sizeof(double), F);
        Grid grid(x, ny, sc[0]);
        Bathymetry bathy(Z.data(), h.data(), sc[1], &grid);
        Wave wave(&grid);
        double *wf[7] = {wave.Sxx, wave.Sxy, wave.Syy, wave.Urms, wave.Dw, wave.D_r, wave.H};
        for (double *p : wf) rd(in, p, sizeof(double), F);
        fclose(in);
        Hydro hydro(&grid, &bathy, &wave);
        Sediment *sed = with_sed ? new Sediment(&hydro, &wave, &bathy, &grid) : nullptr;
        for (int r = 0; r < runs; ++r) {
            hydro.Run(steps);
            if (sed) {
                sed->calculateSedimentFluxes();
                sed->updateBottom();
            }
        }
        FILE *out = fopen(argv[2], "wb");
        if (!out) throw std::runtime_error("cannot open output");
        if (resume_runs > 0) {
            const std::string path = std::string(argv[2]) + ".resume.nc";
            write_resume(path.c_str(), grid, hydro, bathy, sed);
            Bathymetry bathy2(Z.data(), h.data(), sc[1], &grid);   // the bed as the Sediment left it
            std::copy(bathy.h, bathy.h + F, bathy2.h);
            std::copy(bathy.Zb, bathy.Zb + F, bathy2.Zb);
            Hydro hydro2(&grid, &bathy2, &wave);
            hydro2.setHydroFromResumeFile(path);
            std::vector<double> sed_before, sed_after;
            if (sed) {
                Sediment sed2(&hydro2, &wave, &bathy2, &grid);
                sed2.setSedimentFluxesFromResumeFile(path);
                for (const double *p : {sed->qx, sed->qy, sed->dh_dt}) sed_before.insert(sed_before.end(), p, p + F);
                for (const double *p : {sed2.qx, sed2.qy, sed2.dh_dt}) sed_after.insert(sed_after.end(), p, p + F);
            }
            for (int r = 0; r < resume_runs; ++r) {
                hydro.Run(steps);
                hydro2.Run(steps);
            }
            write_hydro(out, hydro, bathy, F, ny);
            write_hydro(out, hydro2, bathy2, F, ny);
            fwrite(sed_before.data(), sizeof(double), sed_before.size(), out);
            fwrite(sed_after.data(), sizeof(double), sed_after.size(), out);
            delete sed;
            fclose(out);
            return 0;
        }
        write_hydro(out, hydro, bathy, F, ny);
        if (sed) {
            const double *sf[8] = {sed->Alpha, sed->gamma, sed->Gamma, sed->qx, sed->qy, sed->dh_dt, bathy.h, bathy.Zb};
            for (const double *p : sf) fwrite(p, sizeof(double), F, out);
            delete sed;
        }
        fclose(out);
    } catch (const std::exception &e) {
        fprintf(stderr, "facade_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
