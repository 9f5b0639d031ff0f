// facade_run.cpp -- drives the C++ Hydro facade the way the reference's main loop drives
// its Hydro (src/main.cpp:17-21, src/Morfo70.cpp:141-153): Grid, Bathymetry, Wave,
// Hydro(grid, bathy, wave), fill the forcing, Run(n) repeatedly, read the public arrays.
//
// With sediment = 1 it also drives the C++ Sediment after every Run, as the outer loop
// does inside the bed evolution window (src/Morfo70.cpp:155-163):
// calculateSedimentFluxes(), then updateBottom().
//
// Usage: facade_run <input.bin> <output.bin>
//   input:  int32 nx, ny, runs, steps_per_run, drag, visc, bc, solver, sediment;
//           float64 dy, zs0, dt, Dmin, alpha_f, cD0, nu0, M, ka, z0, alpha_r, bed_dt;
//           x[nx]; Z[nx*ny]; h[nx*ny]; Sxx, Sxy, Syy, Urms, Dw, D_r, H [nx*ny each]
//   output: Zs, D [nx*ny]; Qx, Qy, U, V [(nx+1)*ny]; U_c, V_c, dQx_dx, dQy_dy [nx*ny];
//           with sediment: Alpha, gamma, Gamma, qx, qy, dh_dt, h, Zb [nx*ny each]
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "Hydro.h"
#include "Sediment.h"

static void rd(FILE *f, void *p, size_t sz, size_t n)
{
    if (fread(p, sz, n, f) != n) throw std::runtime_error("short read");
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
        int hdr[9];
        rd(in, hdr, sizeof(int), 9);
        const int nx = hdr[0], ny = hdr[1], runs = hdr[2], steps = hdr[3], with_sed = hdr[8];
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
        const size_t Fp = F + ny;
        fwrite(hydro.Zs, sizeof(double), F, out);
        fwrite(bathy.D, sizeof(double), F, out);
        fwrite(hydro.Qx, sizeof(double), Fp, out);
        fwrite(hydro.Qy, sizeof(double), Fp, out);
        fwrite(hydro.U, sizeof(double), Fp, out);
        fwrite(hydro.V, sizeof(double), Fp, out);
        fwrite(hydro.U_c, sizeof(double), F, out);
        fwrite(hydro.V_c, sizeof(double), F, out);
        fwrite(hydro.dQx_dx, sizeof(double), F, out);
        fwrite(hydro.dQy_dy, sizeof(double), F, out);
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
