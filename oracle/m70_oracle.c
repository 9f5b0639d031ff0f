/*
 * m70_oracle.c -- CPU restatement of the Morfo70 Hydro ADI step (TEST INFRASTRUCTURE).
 *
 * Checker only: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg. The product library never links or calls it.
 *
 * Each function follows the reference function named in its comment, in the same
 * floating-point operation order (build with -ffp-contract=off). Loops that the reference
 * runs serially with cross-row reads (the OpenMP races in src/HydroQx.cpp:36-46 and
 * src/HydroQy.cpp:27-41) are split into phases here so that the OpenMP version is
 * race-free and bitwise identical to the serial one for any thread count.
 */
#include "m70_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Mutation hooks for the oracle's own tests (tests/test_oracle_consistency.py). A build
 * with -DORC_MUTANT=k scales term k of the assembly by 1.1; the default build multiplies
 * by 1.0, which is exact, so its results are those of the unmarked expressions. Terms:
 *   1 Qx predictor advection     2 Qx predictor x-viscosity    3 Qx predictor cross terms
 *   4 Qx predictor dt*d(dQy_dy)/dx coupling                    5 Qx predictor 1/dt in d
 *   6 Qx predictor implicit gravity                            7 Qx projector advection
 *   8 Qx projector y-viscosity   9 Qy predictor advection     10 Qy predictor y-viscosity
 *  11 Qy predictor cross terms  12 Qy predictor dt*d(dQx_dx)/dy coupling
 *  13 Qy projector advection    14 Qy projector x-viscosity   15 drag term
 *  16 Qx forcing                17 Qy forcing                 18 Qx explicit gravity
 *  19 Qy explicit gravity       20 Qy predictor implicit gravity
 *  21 Qy predictor 1/dt in d    22 Qx projector 1/dt form (1 + dt*drag) in d */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(k) ((ORC_MUTANT) == (k) ? 1.1 : 1.0)

/* include/OptionsParametersConstants.h:14-19 */
static const double g = 9.8100000000000000;
static const double rho = 1025.00000000000;
static const double sqrt_2 = 1.414213562373095;

struct orc_hydro {
    int nx, ny;
    long size; /* nx*ny */
    double dy;
    orc_params p;
    /* Grid metrics (src/Grid.cpp:32-92) */
    double *dx, *dx_x;
    double *difx_x_up, *difx_x_d, *difx_x_low, *dif2x_x_up, *dif2x_x_d, *dif2x_x_low;
    double *difx_c_up, *difx_c_d, *difx_c_low, *dif2x_c_up, *dif2x_c_d, *dif2x_c_low;
    double *x2c_up, *x2c_d;
    double *metrics; /* backing store */
    /* Bathymetry (include/Bathymetry.h:29-32) */
    double *Z, *h, *D;
    /* Wave forcing (include/Wave.h:36-53) */
    double *Sxx, *Sxy, *Syy, *Urms, *Dw, *D_r, *Hw;
    /* State (include/Hydro.h:26-34) */
    double *U, *U_c, *V, *V_c, *Zs, *Qx, *Qy, *dQx_dx, *dQy_dy;
    /* Per-Run (include/Hydro.h:44-59) */
    double *cD, *nuD, *Forcing_x, *Forcing_y, *Urms_x, *Urms_y, *cD_x, *nuD_x, *dnuD_dx,
        *dnuDx_dy, *cD_y, *nuD_y, *dnuD_dy, *dnuDy_dx;
    /* Per-step scratch (include/Hydro.h:62-79) */
    double *guess_Zs, *guess_D, *Di, *inv_Di, *Ui_j, *DragTerm;
    unsigned char *IsWater;
    double *v_tmp, *low, *d, *up, *f;
    /* Boundary constants (include/Hydro.h:89-91) */
    double bcOffshoreQx_f1, bcOffshoreQx_f0, bcOffshoreQy;
    double *f0_series;
    long f0_len, f0_pos;
};

static double *zalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---------------------------------------------------------------------------------- */
/* Grid metrics: src/Grid.cpp:127-145 (closed forms) and :32-92 (assembly).           */
/* ---------------------------------------------------------------------------------- */
static void dif_coefficients(double dx0, double dx1, double *up, double *d, double *low)
{   /* Grid::calculateDifCoefficients, src/Grid.cpp:127-136 */
    double dx02 = dx0 * dx0;
    double dx12 = dx1 * dx1;
    double dx0dx1 = dx0 * dx1;
    double dx0_p_dx1 = dx0 + dx1;
    *d = (dx12 - dx02) / (dx0_p_dx1 * dx0dx1);
    *up = dx02 / (dx0_p_dx1 * dx0dx1);
    *low = -dx12 / (dx0_p_dx1 * dx0dx1);
}

static void dif2_coefficients(double dx0, double dx1, double *up, double *d, double *low)
{   /* Grid::calculateDif2Coefficients, src/Grid.cpp:138-145 */
    double dx0dx1 = dx0 * dx1;
    double dx0_p_dx1 = dx0 + dx1;
    *d = -2 / (dx0dx1);
    *up = 2 / (dx0_p_dx1 * dx1);
    *low = 2 / (dx0_p_dx1 * dx0);
}

void orc_grid_metrics(int nx, const double *x, double *out)
{
    double *dx = out, *dx_x = dx + (nx - 1);
    double *difx_x_up = dx_x + nx, *difx_x_d = difx_x_up + (nx - 1), *difx_x_low = difx_x_d + (nx - 1);
    double *dif2x_x_up = difx_x_low + (nx - 1), *dif2x_x_d = dif2x_x_up + (nx - 1),
           *dif2x_x_low = dif2x_x_d + (nx - 1);
    double *difx_c_up = dif2x_x_low + (nx - 1), *difx_c_d = difx_c_up + nx, *difx_c_low = difx_c_d + nx;
    double *dif2x_c_up = difx_c_low + nx, *dif2x_c_d = dif2x_c_up + nx, *dif2x_c_low = dif2x_c_d + nx;
    double *x2c_up = dif2x_c_low + nx, *x2c_d = x2c_up + nx;
    int i;
    /* src/Grid.cpp:32-34 */
    for (i = 0; i < nx - 1; i++) dx[i] = x[i + 1] - x[i];
    /* src/Grid.cpp:36-40 */
    dx_x[0] = dx[0];
    for (i = 1; i < nx - 1; i++) dx_x[i] = (dx[i] + dx[i - 1]) / 2;
    dx_x[nx - 1] = dx[nx - 2];
    /* src/Grid.cpp:49-53 */
    for (i = 0; i < nx - 1; i++) {
        dif_coefficients(dx_x[i], dx_x[i + 1], &difx_x_up[i], &difx_x_d[i], &difx_x_low[i]);
        dif2_coefficients(dx_x[i], dx_x[i + 1], &dif2x_x_up[i], &dif2x_x_d[i], &dif2x_x_low[i]);
    }
    /* src/Grid.cpp:63-73 */
    dif_coefficients(dx[0], dx[1], &difx_c_up[0], &difx_c_d[0], &difx_c_low[0]);
    dif2_coefficients(dx[0], dx[1], &dif2x_c_up[0], &dif2x_c_d[0], &dif2x_c_low[0]);
    for (i = 1; i < nx - 1; i++) {
        dif_coefficients(dx[i - 1], dx[i], &difx_c_up[i], &difx_c_d[i], &difx_c_low[i]);
        dif2_coefficients(dx[i - 1], dx[i], &dif2x_c_up[i], &dif2x_c_d[i], &dif2x_c_low[i]);
    }
    dif_coefficients(dx[nx - 2], dx[nx - 2] / 2, &difx_c_up[nx - 1], &difx_c_d[nx - 1], &difx_c_low[nx - 1]);
    dif2_coefficients(dx[nx - 2], dx[nx - 2] / 2, &dif2x_c_up[nx - 1], &dif2x_c_d[nx - 1], &dif2x_c_low[nx - 1]);
    /* src/Grid.cpp:76-92 */
    {
        double dx0 = dx[0];
        double dx1 = dx0;
        double inv_dx1pdx0 = 1 / (dx1 + dx0);
        x2c_up[0] = dx0 * inv_dx1pdx0;
        x2c_d[0] = dx1 * inv_dx1pdx0;
        for (i = 1; i < nx - 1; i++) {
            dx0 = dx1;
            dx1 = dx[i];
            inv_dx1pdx0 = 1 / (dx1 + dx0);
            x2c_up[i] = dx0 * inv_dx1pdx0;
            x2c_d[i] = dx1 * inv_dx1pdx0;
        }
        x2c_up[nx - 1] = 0.5;
        x2c_d[nx - 1] = 0.5;
    }
}

/* ---------------------------------------------------------------------------------- */
/* Solvers: src/Math.cpp                                                              */
/* ---------------------------------------------------------------------------------- */
void orc_tridiag(double *Y, double *V, const double *A, const double *B, const double *C,
                 const double *F, int n, int m)
{
    /* Math::TridiagonalSolver, src/Math.cpp:15-65. m independent lines, element stride m.
     * Columns are independent, so OpenMP splits the j range; each column runs the same
     * serial recurrence as the reference. */
#pragma omp parallel
    {
        int j0 = 0, j1 = m;
#ifdef _OPENMP
        int nt = omp_get_num_threads(), t = omp_get_thread_num();
        long chunk = ((long)m + nt - 1) / nt;
        j0 = (int)(chunk * t); if (j0 > m) j0 = m;
        j1 = (int)(chunk * (t + 1)); if (j1 > m) j1 = m;
#endif
        int w = j1 - j0;
        if (w > 0) {
            double *inv_w = (double *)malloc(sizeof(double) * (size_t)w);
            const double *a = A + j0, *b = B + j0, *c = C + j0, *f = F + j0;
            double *y = Y + j0, *v = V + j0, *ylow;
            int i, j;
            for (j = 0; j < w; j++) {
                inv_w[j] = 1 / a[j];
                y[j] = f[j] * inv_w[j];
            }
            for (i = 1; i < n; i++) {
                ylow = y;
                a += m; b += m; y += m; f += m;
                for (j = 0; j < w; j++) {
                    v[j] = c[j] * inv_w[j];
                    inv_w[j] = 1 / (a[j] - b[j] * v[j]);
                    y[j] = (f[j] - b[j] * ylow[j]) * inv_w[j];
                }
                c += m; v += m;
            }
            ylow = Y + j0 + (long)(n - 2) * m;
            y = ylow + m;
            v = V + j0 + (long)(n - 2) * m;
            for (i = n - 2; i > -1; i--) {
                for (j = 0; j < w; j++) ylow[j] -= v[j] * y[j];
                v -= m;
                y = ylow;
                ylow -= m;
            }
            free(inv_w);
        }
    }
}

/* Math::TridiagonalSolverPeriodic (El-Mikkawy 2005), src/Math.cpp:68-118, using caller
 * scratch ws[5n]. a = diagonal, b = lower (b[0] is the corner A(0,n-1)), c = upper
 * (c[n-1] is the corner A(n-1,0)). Requires n >= 3. */
static void tridiag_periodic_ws(const double *a, const double *b, const double *c, const double *f,
                                int n, double *_y, double *ws)
{
    double *k = ws, *inv_k = ws + n, *v = ws + 2 * n, *hh = ws + 3 * n, *r = ws + 4 * n;
    double *y = _y; /* the reference solves into a temporary then copies; same values */
    double t = b[0];
    double s = c[n - 1];
    int i;
    k[0] = a[0];
    v[0] = t;
    inv_k[0] = 1 / k[0];
    hh[0] = s * inv_k[0];
    for (i = 1; i < n - 2; i++) {
        k[i] = a[i] - b[i] * c[i - 1] * inv_k[i - 1];
        inv_k[i] = 1 / k[i];
        v[i] = -b[i] * v[i - 1] * inv_k[i - 1];
        hh[i] = -c[i - 1] * hh[i - 1] * inv_k[i];
    }
    k[n - 2] = a[n - 2] - b[n - 2] * c[n - 3] * inv_k[n - 3];
    inv_k[n - 2] = 1 / k[n - 2];
    v[n - 2] = c[n - 2] - b[n - 2] * v[n - 3] * inv_k[n - 3];
    hh[n - 2] = (b[n - 1] - hh[n - 3] * c[n - 3]) * inv_k[n - 2];

    r[0] = f[0];
    {
        double rh = r[0] * hh[0];
        double vh = v[0] * hh[0];
        for (i = 1; i < n - 1; i++) {
            r[i] = f[i] - b[i] * r[i - 1] * inv_k[i - 1];
            rh += hh[i] * r[i];
            vh += hh[i] * v[i];
        }
        k[n - 1] = a[n - 1] - vh;
        r[n - 1] = f[n - 1] - rh;
    }
    y[n - 1] = r[n - 1] / k[n - 1];
    y[n - 2] = (r[n - 2] - v[n - 2] * y[n - 1]) * inv_k[n - 2];
    for (i = n - 3; i > -1; i--) y[i] = (r[i] - c[i] * y[i + 1] - v[i] * y[n - 1]) * inv_k[i];
}

void orc_tridiag_periodic(const double *a, const double *b, const double *c, const double *f,
                          int n, double *y)
{
    double *ws = zalloc(5L * n);
    double *tmp = zalloc(n);
    tridiag_periodic_ws(a, b, c, f, n, tmp, ws);
    memcpy(y, tmp, sizeof(double) * (size_t)n);
    free(tmp);
    free(ws);
}

/* ---------------------------------------------------------------------------------- */
/* Mesh operations: include/MeshOperations.h, src/HydroMeshOperations.cpp             */
/* ---------------------------------------------------------------------------------- */
static void C2X(double *var_x, const double *var_c, int nx, int ny)
{   /* MeshOperations::C2X, include/MeshOperations.h:18-27 */
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++)
        for (int j = 0; j < ny; j++)
            var_x[j + (long)i * ny] = 0.5 * (var_c[j + (long)i * ny] + var_c[j + (long)(i + 1) * ny]);
}

static void X2C(double *var_c, const double *var_x, const double *x2c_d, const double *x2c_up,
                int nx, int ny)
{   /* MeshOperations::X2C, include/MeshOperations.h:36-47 */
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        const double xd = x2c_d[i];
        const double xup = x2c_up[i];
        for (int j = 0; j < ny; j++)
            var_c[j + (long)i * ny] = xd * var_x[j + (long)i * ny] + xup * var_x[j + (long)(i + 1) * ny];
    }
}

static void C2Y_row(double *var_y, const double *var_c, int ny)
{   /* Hydro::C2Y(row), src/HydroMeshOperations.cpp:63-70 */
    for (int j = 0; j < ny - 1; j++) var_y[j] = 0.5 * (var_c[j + 1] + var_c[j]);
    var_y[ny - 1] = 0.5 * (var_c[0] + var_c[ny - 1]);
}

static void C2Y(double *var_y, const double *var_c, int nx, int ny)
{   /* Hydro::C2Y(nx), src/HydroMeshOperations.cpp:47-61 */
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) C2Y_row(var_y + (long)i * ny, var_c + (long)i * ny, ny);
}

static void Y2C_row(double *var_c, const double *var_y, int ny)
{   /* Hydro::Y2C(row), src/HydroMeshOperations.cpp:88-97 */
    for (int j = 1; j < ny; j++) var_c[j] = 0.5 * (var_y[j] + var_y[j - 1]);
    var_c[0] = 0.5 * (var_y[0] + var_y[ny - 1]);
}

static void difC2X(double *df_dx, const double *var_c, const double *dx, int nx, int ny)
{   /* Hydro::difC2X, src/HydroMeshOperations.cpp:99-114 (rows 0..nx-2) */
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) {
        const double *v = var_c + (long)i * ny, *v_up = v + ny;
        double *d_dx = df_dx + (long)i * ny;
        double inv_x = 1 / dx[i];
        for (int j = 0; j < ny; j++) d_dx[j] = inv_x * (v_up[j] - v[j]);
    }
}

static void difC2Y(double *_df_dy, const double *var_c, double dy, int nx, int ny)
{   /* Hydro::difC2Y, src/HydroMeshOperations.cpp:116-130 */
    double inv_dy = 1 / dy;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        const double *v = var_c + (long)i * ny;
        double *df_dy = _df_dy + (long)i * ny;
        for (int j = 0; j < ny - 1; j++) df_dy[j] = (v[j + 1] - v[j]) * inv_dy;
        df_dy[ny - 1] = (v[0] - v[ny - 1]) * inv_dy;
    }
}

static void difY2C_row(double *df_dy, const double *var_y, double dy, int ny)
{   /* Hydro::difY2C(row), src/HydroMeshOperations.cpp:148-159 */
    double inv_dy = 1 / dy;
    for (int j = 1; j < ny; j++) df_dy[j] = (var_y[j] - var_y[j - 1]) * inv_dy;
    df_dy[0] = (var_y[0] - var_y[ny - 1]) * inv_dy;
}

static void difC2C_x_addBoundary(double *_df_dx, const double *var_c, const double *difx_d,
                                 const double *difx_up, const double *difx_low, int nx, int ny)
{   /* Hydro::difC2C_x_addBoundary, src/HydroMeshOperations.cpp:161-200 */
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        double *df_dx = _df_dx + (long)i * ny;
        const double *v = var_c + (long)i * ny;
        if (i == 0) {
            const double *v_up = v + ny;
            double dup = difx_up[0], dlow = difx_low[0], dd = difx_d[0];
            for (int j = 0; j < ny; j++) df_dx[j] = dlow * (2 * v[j] - v_up[j]) + dd * v[j] + v_up[j] * dup;
        } else if (i < nx - 1) {
            const double *v_low = v - ny, *v_up = v + ny;
            double dup = difx_up[i], dd = difx_d[i], dlow = difx_low[i];
            for (int j = 0; j < ny; j++) df_dx[j] = dlow * v_low[j] + dd * v[j] + v_up[j] * dup;
        } else {
            const double *v_low = v - ny;
            double dd = difx_d[nx - 1], dlow = difx_low[nx - 1];
            for (int j = 0; j < ny; j++) df_dx[j] = dlow * v_low[j] + dd * v[j];
        }
    }
}

static void difC2C_y(double *_df_dy, const double *var_c, double dy, int nx, int ny)
{   /* Hydro::difC2C_y, src/HydroMeshOperations.cpp:202-221 */
    double inv_2dy = 0.5 / dy;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        const double *v = var_c + (long)i * ny;
        double *df_dy = _df_dy + (long)i * ny;
        for (int j = 1; j < ny - 1; j++) df_dy[j] = inv_2dy * (v[j + 1] - v[j - 1]);
        df_dy[0] = inv_2dy * (v[1] - v[ny - 1]);
        df_dy[ny - 1] = inv_2dy * (v[0] - v[ny - 2]);
    }
}

/* ---------------------------------------------------------------------------------- */
/* Parametrizations: src/HydroParametrizations.cpp                                    */
/* ---------------------------------------------------------------------------------- */
static void calculateDragTerm(double *drag, const double *cd, const double *urms, const double *u,
                              const double *v, const double *inv_d, int ny, double alpha_f)
{   /* Hydro::calculateDragTerm, src/HydroParametrizations.cpp:19-28 */
    const double alpha_f2 = (alpha_f * alpha_f);
    const double inv_sqrt2 = 1 / sqrt_2;
    for (int j = 0; j < ny; j++)
        drag[j] = MUT(15) * (cd[j] * urms[j] * inv_sqrt2 *
                  sqrt(alpha_f2 + 2 * (v[j] * v[j] + u[j] * u[j]) / (urms[j] * urms[j])) * inv_d[j]);
}

static void calculateDragAndViscosity(orc_hydro *H)
{   /* Hydro::calculateDragAndViscosity, src/HydroParametrizations.cpp:30-47, with the
     * model functions of :49-113 */
    const int nx = H->nx, ny = H->ny;
    const orc_params *p = &H->p;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        const double *D = H->D + (long)i * ny;
        double *cD = H->cD + (long)i * ny;
        const double *Dw = H->Dw + (long)i * ny, *D_r = H->D_r + (long)i * ny, *Hw = H->Hw + (long)i * ny;
        double *nuD = H->nuD + (long)i * ny;
        int j;
        switch (p->drag_model) {
        case ORC_DRAG_CONSTANT: /* :49-51 */
            for (j = 0; j < ny; j++) cD[j] = p->cD0;
            break;
        case ORC_DRAG_LOGDEPTH: /* :53-63 */
            for (j = 0; j < ny; j++) {
                double d = D[j];
                d = (d < p->z0 + 0.001) ? (p->z0 + 0.001) : d; /* std::max(d, z0+0.001) */
                const double cd = 0.4 / (log(d / p->z0) - 1.0);
                cD[j] = cd * cd;
            }
            break;
        default: /* ORC_DRAG_MANNING, :65-71 */
            for (j = 0; j < ny; j++) cD[j] = 0.015 * cbrt(p->ka / D[j]);
            break;
        }
        switch (p->visc_model) {
        case ORC_VISC_CONSTANT: /* :73-77 */
            for (j = 0; j < ny; j++) nuD[j] = D[j] * p->nu0;
            break;
        case ORC_VISC_ROLLERS: { /* :79-90 */
            double inv_rho = 1 / rho;
            for (j = 0; j < ny; j++)
                nuD[j] = D[j] * (p->nu0 + p->M * D[j] * cbrt(((1 - p->alpha_r) * Dw[j] + D_r[j]) * inv_rho));
            break;
        }
        case ORC_VISC_DEPTH: { /* :92-101 */
            double inv_rho = 1 / rho;
            for (j = 0; j < ny; j++) nuD[j] = D[j] * (p->nu0 + p->M * D[j] * cbrt(Dw[j] * inv_rho));
            break;
        }
        default: { /* ORC_VISC_HRMS, :103-113; note the float inv_rho at :107 */
            float inv_rho = (float)(1 / rho);
            for (j = 0; j < ny; j++) nuD[j] = D[j] * (p->nu0 + p->M * Hw[j] * cbrt(Dw[j] * inv_rho));
            break;
        }
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Forcing: src/HydroQx.cpp:357-388, src/HydroQy.cpp:367-424                         */
/* ---------------------------------------------------------------------------------- */
static void getForcing_x(orc_hydro *H)
{
    const int nx = H->nx, ny = H->ny;
    double inv_rho = 1 / rho;
    double inv_4dy = 0.25 / H->dy;
    inv_4dy *= inv_rho;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) {
        const double *Sxy = H->Sxy + (long)i * ny, *Sxy_up = Sxy + ny;
        const double *Sxx = H->Sxx + (long)i * ny, *Sxx_up = Sxx + ny;
        double *forcing_x = H->Forcing_x + (long)i * ny;
        double inv_rhodx = inv_rho / H->dx[i];
        for (int j = 1; j < ny - 1; j++)
            forcing_x[j] = -(inv_4dy * (Sxy[j + 1] - Sxy[j - 1] + Sxy_up[j + 1] - Sxy_up[j - 1]) +
                             (Sxx_up[j] - Sxx[j]) * inv_rhodx);
        forcing_x[0] = -(inv_4dy * (Sxy[1] - Sxy[ny - 1] + Sxy_up[1] - Sxy_up[ny - 1]) + (Sxx_up[0] - Sxx[0]) * inv_rhodx);
        forcing_x[ny - 1] = -(inv_4dy * (Sxy[0] - Sxy[ny - 2] + Sxy_up[0] - Sxy_up[ny - 2]) +
                              (Sxx_up[ny - 1] - Sxx[ny - 1]) * inv_rhodx);
    }
}

static void forcing_y_row(double *forcing_y, const double *Syy, const double *Sxy_low, const double *Sxy,
                          const double *Sxy_up, double inv_dy, double d_dx_low, double d_dx_d,
                          double d_dx_up, int ny)
{   /* inner loops of Hydro::getForcing_y, src/HydroQy.cpp:389-393, 402-405, 418-421 */
    if (Sxy_up) {
        for (int j = 0; j < ny - 1; j++)
            forcing_y[j] = -(inv_dy * (Syy[j + 1] - Syy[j]) +
                             0.5 * ((Sxy_up[j] + Sxy_up[j + 1]) * d_dx_up + (Sxy[j] + Sxy[j + 1]) * d_dx_d +
                                    (Sxy_low[j] + Sxy_low[j + 1]) * d_dx_low));
        forcing_y[ny - 1] = -(inv_dy * (Syy[0] - Syy[ny - 1]) +
                              0.5 * ((Sxy_up[0] + Sxy_up[ny - 1]) * d_dx_up + (Sxy[0] + Sxy[ny - 1]) * d_dx_d +
                                     (Sxy_low[0] + Sxy_low[ny - 1]) * d_dx_low));
    } else {
        for (int j = 0; j < ny - 1; j++)
            forcing_y[j] = -(inv_dy * (Syy[j + 1] - Syy[j]) +
                             0.5 * ((Sxy[j] + Sxy[j + 1]) * d_dx_d + (Sxy_low[j] + Sxy_low[j + 1]) * d_dx_low));
        forcing_y[ny - 1] = -(inv_dy * (Syy[0] - Syy[ny - 1]) +
                              0.5 * ((Sxy[0] + Sxy[ny - 1]) * d_dx_d + (Sxy_low[0] + Sxy_low[ny - 1]) * d_dx_low));
    }
}

static void getForcing_y(orc_hydro *H)
{   /* Hydro::getForcing_y, src/HydroQy.cpp:367-424. Row 0 uses Sxy rows 0,1,2 with the
     * row-0 coefficients (:385-393); the last row has no "up" term (:413-421). */
    const int nx = H->nx, ny = H->ny;
    double inv_rho = 1 / rho;
    double inv_dy = 1 / H->dy;
    inv_dy *= inv_rho;
    const double *S = H->Sxy;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        double d_dx_d = H->difx_c_d[i] * inv_rho;
        double d_dx_low = H->difx_c_low[i] * inv_rho;
        double *fy = H->Forcing_y + (long)i * ny;
        const double *Syy = H->Syy + (long)i * ny;
        if (i == 0) {
            double d_dx_up = H->difx_c_up[0] * inv_rho;
            forcing_y_row(fy, Syy, S, S + ny, S + 2L * ny, inv_dy, d_dx_low, d_dx_d, d_dx_up, ny);
        } else if (i < nx - 1) {
            double d_dx_up = H->difx_c_up[i] * inv_rho;
            forcing_y_row(fy, Syy, S + (long)(i - 1) * ny, S + (long)i * ny, S + (long)(i + 1) * ny, inv_dy,
                          d_dx_low, d_dx_d, d_dx_up, ny);
        } else {
            forcing_y_row(fy, Syy, S + (long)(i - 1) * ny, S + (long)i * ny, NULL, inv_dy, d_dx_low, d_dx_d,
                          0.0, ny);
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Depth interpolation and wet/dry masks                                              */
/* ---------------------------------------------------------------------------------- */
int orc_wet_rule(double d, double dup, double zs, double zsup, double s, double sup, double Dmin)
{   /* body of Hydro::setWaterNodes_x, src/HydroQx.cpp:245-253 (identical in _y,
     * src/HydroQy.cpp:103-109); integer arithmetic on bools, then bool conversion */
    int dup_gt_min = dup > Dmin;
    int d_gt_min = d > Dmin;
    int zsup_gt_zs = zsup > zs;
    int zs_gt_zsup = zs > zsup;
    int sink = s > 0;
    int sinkup = sup > 0;
    return (d_gt_min * dup_gt_min +
            (dup_gt_min * (!d_gt_min) * zsup_gt_zs + d_gt_min * (!dup_gt_min) * zs_gt_zsup) * !(sink + sinkup)) != 0;
}

static void getD_x0(orc_hydro *H)
{   /* Hydro::getD_x0, src/HydroQx.cpp:390-401 */
    const int ny = H->ny;
    const double *d = H->D;
    for (int j = 0; j < ny; j++) {
        H->Di[j] = 1.5 * d[j] - 0.5 * d[j + ny];
        H->inv_Di[j] = 1.0 / H->Di[j];
    }
}

static void getD_x(orc_hydro *H, int i)
{   /* Hydro::getD_x, src/HydroQx.cpp:403-418 */
    const int ny = H->ny;
    const double *d = H->D + (long)i * ny;
    double *d_x = H->Di + (long)(i + 1) * ny, *inv_d_x = H->inv_Di + (long)(i + 1) * ny;
    for (int j = 0; j < ny; j++) {
        d_x[j] = 0.5 * (d[j] + d[j + ny]);
        inv_d_x[j] = 1 / d_x[j];
    }
}

static void setWaterNodes_x(orc_hydro *H, int i)
{   /* Hydro::setWaterNodes_x, src/HydroQx.cpp:231-256 */
    const int ny = H->ny;
    const double Dmin = H->p.Dmin + 1E-4;
    const double *d = H->D + (long)i * ny, *zs = H->Zs + (long)i * ny, *dqx_dx = H->dQx_dx + (long)i * ny;
    unsigned char *isWater = H->IsWater + (long)i * ny;
    for (int j = 0; j < ny; j++)
        isWater[j] = (unsigned char)orc_wet_rule(d[j], d[j + ny], zs[j], zs[j + ny], dqx_dx[j], dqx_dx[j + ny], Dmin);
}

static void setWaterNodes_y(orc_hydro *H, int i)
{   /* Hydro::setWaterNodes_y, src/HydroQy.cpp:92-122 */
    const int ny = H->ny;
    const double Dmin = H->p.Dmin + 1E-4;
    const double *d = H->guess_D + (long)i * ny, *zs = H->guess_Zs + (long)i * ny,
                 *dqy_dy = H->dQy_dy + (long)i * ny;
    unsigned char *isWater = H->IsWater + (long)i * ny;
    for (int j = 0; j < ny - 1; j++)
        isWater[j] = (unsigned char)orc_wet_rule(d[j], d[j + 1], zs[j], zs[j + 1], dqy_dy[j], dqy_dy[j + 1], Dmin);
    isWater[ny - 1] =
        (unsigned char)orc_wet_rule(d[ny - 1], d[0], zs[ny - 1], zs[0], dqy_dy[ny - 1], dqy_dy[0], Dmin);
}

static void getD_y0(orc_hydro *H)
{   /* Hydro::getD_y0, src/HydroQy.cpp:426-440. The wrap element uses d[0] (row 0) where
     * the interior uses row 1: reproduced as written (:438). */
    const int ny = H->ny;
    const double *d = H->guess_D;
    double *d_y = H->Di, *inv_d_y = H->inv_Di;
    for (int j = 0; j < ny - 1; j++) {
        d_y[j] = 1 * (d[j + 1] + d[j]) - 0.5 * (d[ny + j + 1] + d[ny + j]);
        inv_d_y[j] = 1 / d_y[j];
    }
    d_y[ny - 1] = 1 * (d[0] + d[ny - 1]) - 0.5 * (d[0] + d[2 * ny - 1]);
    inv_d_y[ny - 1] = 1 / d_y[ny - 1];
}

static void getD_y(orc_hydro *H, int i)
{   /* Hydro::getD_y, src/HydroQy.cpp:442-456 */
    const int ny = H->ny;
    const double *d = H->guess_D + (long)i * ny;
    double *d_y = H->Di + (long)(i + 1) * ny, *inv_d_y = H->inv_Di + (long)(i + 1) * ny;
    for (int j = 0; j < ny - 1; j++) {
        d_y[j] = 0.5 * (d[j + 1] + d[j]);
        inv_d_y[j] = 1 / d_y[j];
    }
    d_y[ny - 1] = 0.5 * (d[0] + d[ny - 1]);
    inv_d_y[ny - 1] = 1 / d_y[ny - 1];
}

/* ---------------------------------------------------------------------------------- */
/* Qx predictor coefficients: src/HydroQx.cpp:99-229, BC src/HydroBoundaryConditions.cpp */
/* ---------------------------------------------------------------------------------- */
static void setSolverQxx_f(orc_hydro *H, int i)
{   /* src/HydroQx.cpp:99-133 */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double inv_dt = 1 / dt;
    const double inv_dy = 1 / H->dy;
    const double inv_dx = 1 / H->dx[i];
    const double *zs = H->Zs + (long)i * ny, *dqy_dy = H->dQy_dy + (long)i * ny, *v = H->V + (long)(i + 1) * ny,
                 *qx = H->Qx + (long)(i + 1) * ny, *forcing_x = H->Forcing_x + (long)i * ny,
                 *d_x = H->Di + (long)(i + 1) * ny, *nud_x = H->nuD_x + (long)i * ny,
                 *dnud_dy = H->dnuDx_dy + (long)i * ny;
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    double *_f = H->f + (long)i * ny;
    for (int j = 1; j < ny; j++) {
        const double dzs_dx = inv_dx * (zs[ny + j] - zs[j]);
        const double d2qy_dxy = inv_dx * (dqy_dy[ny + j] - dqy_dy[j]);
        const double dV_dx = 0.5 * inv_dx * ((v[ny + j] + v[ny + j - 1]) - (v[j] + v[j - 1]));
        const double d2V_dxy = inv_dx * inv_dy * ((v[ny + j] - v[ny + j - 1]) - (v[j] - v[j - 1]));
        _f[j] = isWater[j] * (MUT(16) * forcing_x[j] + qx[j] * inv_dt +
                              g * d_x[j] * (MUT(4) * (dt * d2qy_dxy) - MUT(18) * dzs_dx) +
                              MUT(3) * (dnud_dy[j] * dV_dx) + MUT(3) * (nud_x[j] * d2V_dxy));
    }
    {
        const double dzs_dx = inv_dx * (zs[ny] - zs[0]);
        const double d2qy_dxy = inv_dx * (dqy_dy[ny] - dqy_dy[0]);
        const double dV_dx = 0.5 * inv_dx * ((v[ny] + v[2 * ny - 1]) - (v[0] + v[ny - 1]));
        const double d2V_dxy = inv_dx * inv_dy * ((v[ny] - v[2 * ny - 1]) - (v[0] - v[ny - 1]));
        _f[0] = isWater[0] * (MUT(16) * forcing_x[0] + qx[0] * inv_dt +
                              g * d_x[0] * (MUT(4) * (dt * d2qy_dxy) - MUT(18) * dzs_dx) +
                              MUT(3) * (dnud_dy[0] * dV_dx) + MUT(3) * (nud_x[0] * d2V_dxy));
    }
}

static void setSolverQxx_d(orc_hydro *H, int i)
{   /* src/HydroQx.cpp:135-162 */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double inv_dt = 1 / dt;
    const double difx = H->difx_x_d[i];
    const double dif2x = H->dif2x_x_d[i];
    const double *d_x = H->Di + (long)(i + 1) * ny, *inv_d_x = H->inv_Di + (long)(i + 1) * ny,
                 *nud_x = H->nuD_x + (long)i * ny, *dnud_dx = H->dnuD_dx + (long)i * ny,
                 *dragTerm = H->DragTerm + (long)i * ny, *u = H->U + (long)(i + 1) * ny;
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    double *_d = H->d + (long)i * ny;
    for (int j = 0; j < ny; j++) {
        int isl = !isWater[j];
        _d[j] = (MUT(5) * inv_dt + dragTerm[j] + difx * (MUT(1) * u[j] - MUT(2) * (2 * dnud_dx[j] * inv_d_x[j])) -
                 dif2x * (MUT(6) * (g * d_x[j] * dt) + MUT(2) * (2 * nud_x[j] * inv_d_x[j]))) * isWater[j] + isl;
    }
}

static void setSolverQxx_up(orc_hydro *H, int i)
{   /* src/HydroQx.cpp:164-188 */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double difx = H->difx_x_up[i];
    const double dif2x = H->dif2x_x_up[i];
    const double *uup = H->U + (long)(i + 2) * ny, *d_x = H->Di + (long)(i + 1) * ny,
                 *inv_d_xup = H->inv_Di + (long)(i + 2) * ny, *nud_x = H->nuD_x + (long)i * ny,
                 *dnud_dx = H->dnuD_dx + (long)i * ny;
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    double *_up = H->up + (long)i * ny;
    for (int j = 0; j < ny; j++) {
        double _isWater = isWater[j] && isWater[j + ny];
        _up[j] = (difx * (MUT(1) * uup[j] - MUT(2) * (2 * dnud_dx[j] * inv_d_xup[j])) -
                  dif2x * (MUT(6) * (g * d_x[j] * dt) + MUT(2) * (2 * nud_x[j] * inv_d_xup[j]))) *
                 _isWater;
    }
}

static void setSolverQxx_low(orc_hydro *H, int i)
{   /* src/HydroQx.cpp:190-212 */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double difx = H->difx_x_low[i];
    const double dif2x = H->dif2x_x_low[i];
    const double *ulow = H->U + (long)i * ny, *d_x = H->Di + (long)(i + 1) * ny,
                 *inv_d_xlow = H->inv_Di + (long)i * ny, *nud_x = H->nuD_x + (long)i * ny,
                 *dnud_dx = H->dnuD_dx + (long)i * ny;
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    double *_low = H->low + (long)i * ny;
    for (int j = 0; j < ny; j++) {
        double _isWater = isWater[j] && isWater[j - ny];
        _low[j] = (difx * (MUT(1) * ulow[j] - MUT(2) * (2 * dnud_dx[j] * inv_d_xlow[j])) -
                   dif2x * (MUT(6) * (g * d_x[j] * dt) + MUT(2) * (2 * nud_x[j] * inv_d_xlow[j]))) *
                  _isWater;
    }
}

static void setSolverQxx0_low(orc_hydro *H)
{   /* src/HydroQx.cpp:214-229 (row 0, unmasked) */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double difx = H->difx_x_low[0];
    const double dif2x = H->dif2x_x_low[0];
    const double *ulow = H->U, *d_x = H->Di + ny, *inv_d_xlow = H->inv_Di, *nud_x = H->nuD_x, *dnud_dx = H->dnuD_dx;
    double *_low = H->low;
    for (int j = 0; j < ny; j++)
        _low[j] = (difx * (MUT(1) * ulow[j] - MUT(2) * (2 * dnud_dx[j] * inv_d_xlow[j])) -
                   dif2x * (MUT(6) * (g * d_x[j] * dt) + MUT(2) * (2 * nud_x[j] * inv_d_xlow[j])));
}

static void setSolverQxx_BC(orc_hydro *H)
{   /* Hydro::setSolverQxx_BC, src/HydroBoundaryConditions.cpp:25-39 */
    const int ny = H->ny;
    double f1 = H->bcOffshoreQx_f1, f0 = H->bcOffshoreQx_f0;
    for (int j = 0; j < ny; j++) {
        H->d[j] += f1 * H->low[j];
        H->f[j] -= f0 * H->low[j];
    }
}

/* ---------------------------------------------------------------------------------- */
/* Qx projector coefficients: src/HydroQx.cpp:262-349                                 */
/* ---------------------------------------------------------------------------------- */
static void setSolverQxy(orc_hydro *H, int i)
{
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double *dragTerm = H->DragTerm + (long)i * ny, *qx = H->Qx + (long)(i + 1) * ny,
                 *nud_x = H->nuD_x + (long)i * ny, *inv_d_x = H->inv_Di + (long)(i + 1) * ny,
                 *dnudx_dy = H->dnuDx_dy + (long)i * ny, *v = H->Ui_j + (long)i * ny;
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    double *_f = H->f + (long)i * ny, *_d = H->d + (long)i * ny, *_up = H->up + (long)i * ny,
           *_low = H->low + (long)i * ny;
    int j;
    /* setSolverQxy_f, :262-276 */
    for (j = 0; j < ny; j++) _f[j] = isWater[j] * qx[j] * (1 + dt * dragTerm[j]);
    /* setSolverQxy_d, :278-295 */
    {
        const double _2inv_dy2 = 2 / (H->dy * H->dy);
        for (j = 0; j < ny; j++) {
            int isl = !isWater[j];
            _d[j] = isWater[j] * (MUT(22) * (1 + dt * dragTerm[j]) + MUT(8) * (dt * inv_d_x[j] * nud_x[j] * _2inv_dy2)) + isl;
        }
    }
    {
        const double inv_2dy = 0.5 / H->dy;
        const double inv_dy2 = 4 * inv_2dy * inv_2dy;
        /* setSolverQxy_up, :297-322 */
        for (j = 0; j < ny - 1; j++) {
            double _isWater = isWater[j] && isWater[j + 1];
            _up[j] = dt * (MUT(7) * (v[j + 1] * inv_2dy) - MUT(8) * (inv_d_x[j + 1] * (nud_x[j] * inv_dy2 + dnudx_dy[j] * inv_2dy))) * _isWater;
        }
        {
            double _isWater = isWater[ny - 1] && isWater[0];
            _up[ny - 1] = dt * (MUT(7) * (v[0] * inv_2dy) - MUT(8) * (inv_d_x[0] * (nud_x[ny - 1] * inv_dy2 + dnudx_dy[ny - 1] * inv_2dy))) *
                          _isWater;
        }
        /* setSolverQxy_low, :324-349 */
        for (j = 1; j < ny; j++) {
            double _isWater = isWater[j] && isWater[j - 1];
            _low[j] = dt * (MUT(7) * (-v[j - 1] * inv_2dy) - MUT(8) * (inv_d_x[j - 1] * (nud_x[j] * inv_dy2 - dnudx_dy[j] * inv_2dy))) * _isWater;
        }
        {
            double _isWater = isWater[0] && isWater[ny - 1];
            _low[0] = dt * (MUT(7) * (-v[ny - 1] * inv_2dy) - MUT(8) * (inv_d_x[ny - 1] * (nud_x[0] * inv_dy2 - dnudx_dy[0] * inv_2dy))) * _isWater;
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Post-solver: src/HydroQx.cpp:420-451, src/Hydro.cpp:230-271, src/HydroQy.cpp:458-487 */
/* ---------------------------------------------------------------------------------- */
static void solveZs_guess(orc_hydro *H, int i)
{   /* src/Hydro.cpp:230-250 */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double *dqy_dy = H->dQy_dy + (long)i * ny, *dqx_dx = H->dQx_dx + (long)i * ny,
                 *zs = H->Zs + (long)i * ny, *d = H->D + (long)i * ny;
    double *guess_zs = H->guess_Zs + (long)i * ny, *guess_d = H->guess_D + (long)i * ny;
    for (int j = 0; j < ny; j++) {
        const double incrZs = -dt * (dqx_dx[j] + dqy_dy[j]);
        guess_zs[j] = zs[j] + incrZs;
        guess_d[j] = d[j] + incrZs;
    }
}

static void solveZs(orc_hydro *H, int i)
{   /* src/Hydro.cpp:253-271 */
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double *dqy_dy = H->dQy_dy + (long)i * ny, *dqx_dx = H->dQx_dx + (long)i * ny;
    double *zs = H->Zs + (long)i * ny, *d = H->D + (long)i * ny;
    for (int j = 0; j < ny; j++) {
        double incrZs = -dt * (dqx_dx[j] + dqy_dy[j]);
        zs[j] += incrZs;
        d[j] += incrZs;
    }
}

static void getU0(orc_hydro *H)
{   /* src/HydroQx.cpp:420-431 */
    const int ny = H->ny;
    const double *d = H->guess_D, *qx = H->Qx;
    for (int j = 0; j < ny; j++) {
        const double d_x0 = 1.5 * d[j] - 0.5 * d[j + ny];
        H->U[j] = qx[j] / d_x0;
    }
}

static void getU(orc_hydro *H, int i)
{   /* src/HydroQx.cpp:433-451 */
    const int ny = H->ny;
    const double *d = H->guess_D + (long)i * ny, *qx = H->Qx + (long)(i + 1) * ny;
    double *u = H->U + (long)(i + 1) * ny;
    for (int j = 0; j < ny; j++) {
        const double d_x = 0.5 * (d[j] + d[j + ny]);
        u[j] = qx[j] / d_x;
    }
}

static void getV0(orc_hydro *H)
{   /* src/HydroQy.cpp:458-469 */
    const int ny = H->ny;
    const double *d = H->D, *qy = H->Qy;
    for (int j = 0; j < ny; j++) {
        const double d_y0 = 2 * d[j] - d[j + ny];
        H->V[j] = qy[j] / d_y0;
    }
}

static void getV(orc_hydro *H, int i)
{   /* src/HydroQy.cpp:471-487 */
    const int ny = H->ny;
    const double *d = H->D + (long)i * ny, *qy = H->Qy + (long)(i + 1) * ny;
    double *v = H->V + (long)(i + 1) * ny;
    for (int j = 0; j < ny - 1; j++) {
        const double d_y = 0.5 * (d[j] + d[j + 1]);
        v[j] = qy[j] / d_y;
    }
    {
        const double d_y = 0.5 * (d[ny - 1] + d[0]);
        v[ny - 1] = qy[ny - 1] / d_y;
    }
}

/* ---------------------------------------------------------------------------------- */
/* solveQx: src/HydroQx.cpp:18-93                                                     */
/* ---------------------------------------------------------------------------------- */
static void solveQx(orc_hydro *H)
{
    const int nx = H->nx, ny = H->ny;
    int i;
    C2X(H->Ui_j, H->V_c, nx, ny); /* :23 */

    /* :29-47. Serially the reference interleaves getD_x(i), setWaterNodes_x(i),
     * setSolverQxx_up(i-1), drag, f, d, low(i); each call only reads rows written earlier
     * in that order, so computing each quantity over all rows first gives the same values. */
    getD_x0(H);
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) getD_x(H, i);
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) setWaterNodes_x(H, i);
    /* IsWater row nx-1 is never written by the reference (src/Hydro.cpp:91); it stays
     * false here (calloc), matching the "must be zero" intent at src/HydroQx.cpp:47. */
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) {
        calculateDragTerm(H->DragTerm + (long)i * ny, H->cD_x + (long)i * ny, H->Urms_x + (long)i * ny,
                          H->U + (long)(i + 1) * ny, H->Ui_j + (long)i * ny, H->inv_Di + (long)(i + 1) * ny, ny,
                          H->p.alpha_f);
        setSolverQxx_f(H, i);
        setSolverQxx_d(H, i);
        if (i > 0) setSolverQxx_low(H, i);
        setSolverQxx_up(H, i);
    }
    setSolverQxx0_low(H);
    setSolverQxx_BC(H); /* :48 */

    /* :51 */
    orc_tridiag(H->Qx + ny, H->v_tmp, H->d, H->low, H->up, H->f, nx - 1, ny);

    /* :59-70 projector, one periodic solve per row */
#pragma omp parallel
    {
        double *ws = zalloc(5L * ny);
        double *tmp = zalloc(ny);
        int ii;
#pragma omp for schedule(static)
        for (ii = 0; ii < nx - 1; ii++) {
            setSolverQxy(H, ii);
            tridiag_periodic_ws(H->d + (long)ii * ny, H->low + (long)ii * ny, H->up + (long)ii * ny,
                                H->f + (long)ii * ny, ny, tmp, ws);
            memcpy(H->Qx + (long)(ii + 1) * ny, tmp, sizeof(double) * (size_t)ny);
        }
        free(tmp);
        free(ws);
    }

    /* :76 setQx_BC, src/HydroBoundaryConditions.cpp:13-23 */
    for (int j = 0; j < ny; j++) H->Qx[j] = H->bcOffshoreQx_f1 * H->Qx[ny + j] + H->bcOffshoreQx_f0;

    /* :80 difC2X over nx+1 rows with dx_x */
    difC2X(H->dQx_dx, H->Qx, H->dx_x, nx + 1, ny);
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) solveZs_guess(H, i); /* :81-83 */

    getU0(H); /* :86 */
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) getU(H, i); /* :87-89 */

    X2C(H->U_c, H->U, H->x2c_d, H->x2c_up, nx, ny); /* :92 */
}

/* ---------------------------------------------------------------------------------- */
/* Qy predictor coefficients: src/HydroQy.cpp:124-230                                 */
/* ---------------------------------------------------------------------------------- */
static void setSolverQyy(orc_hydro *H, int i)
{
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double inv_dt = 1 / dt;
    double *_f = H->f + (long)i * ny, *_d = H->d + (long)i * ny, *_up = H->up + (long)i * ny,
           *_low = H->low + (long)i * ny;
    const unsigned char *isWater0 = H->IsWater; /* row 0 for f and d: src/HydroQy.cpp:140,170 */
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    int j;
    { /* setSolverQyy_f, :124-158 */
        const double inv_dy = 1 / H->dy;
        const double inv_dx = 1 / H->dx_x[i];
        const double x2c_d = H->x2c_d[i];
        const double x2c_up = H->x2c_up[i];
        const double *zs = H->Zs + (long)i * ny, *dqx_dx = H->dQx_dx + (long)i * ny,
                     *forcing_y = H->Forcing_y + (long)i * ny, *qy = H->Qy + (long)(i + 1) * ny,
                     *d_y = H->Di + (long)(i + 1) * ny, *u = H->U + (long)i * ny, *nud_y = H->nuD_y + (long)i * ny,
                     *dnudy_dx = H->dnuDy_dx + (long)i * ny;
        for (j = 0; j < ny - 1; j++) {
            const double dzs_dy = inv_dy * (zs[j + 1] - zs[j]);
            const double d2qx_dxy = inv_dy * (dqx_dx[j + 1] - dqx_dx[j]);
            const double du_dy = inv_dy * (x2c_up * (u[ny + j + 1] - u[ny + j]) + x2c_d * (u[j + 1] - u[j]));
            const double d2u_dxy = inv_dx * inv_dy * ((u[ny + j + 1] - u[ny + j]) - (u[j + 1] - u[j]));
            _f[j] = isWater0[j] * (MUT(17) * forcing_y[j] + qy[j] * inv_dt +
                                   g * d_y[j] * (MUT(12) * (dt * d2qx_dxy) - MUT(19) * dzs_dy) +
                                   MUT(11) * (dnudy_dx[j] * du_dy) + MUT(11) * (nud_y[j] * d2u_dxy));
        }
        {
            const double dzs_dy = inv_dy * (zs[0] - zs[ny - 1]);
            const double d2qx_dxy = inv_dy * (dqx_dx[0] - dqx_dx[ny - 1]);
            const double du_dy = inv_dy * (x2c_up * (u[ny] - u[2 * ny - 1]) + x2c_d * (u[0] - u[ny - 1]));
            const double d2u_dxy = inv_dx * inv_dy * ((u[ny] - u[2 * ny - 1]) - (u[0] - u[ny - 1]));
            _f[ny - 1] = isWater0[ny - 1] * (MUT(17) * forcing_y[ny - 1] + qy[ny - 1] * inv_dt +
                                             g * d_y[ny - 1] * (MUT(12) * (dt * d2qx_dxy) - MUT(19) * dzs_dy) +
                                             MUT(11) * (dnudy_dx[ny - 1] * du_dy) + MUT(11) * (nud_y[ny - 1] * d2u_dxy));
        }
    }
    /* :34 C2Y(U_c) -> Ui_j (U at y-faces), :35 drag */
    C2Y_row(H->Ui_j + (long)i * ny, H->U_c + (long)i * ny, ny);
    calculateDragTerm(H->DragTerm + (long)i * ny, H->cD_y + (long)i * ny, H->Urms_y + (long)i * ny,
                      H->Ui_j + (long)i * ny, H->V + (long)(i + 1) * ny, H->inv_Di + (long)(i + 1) * ny, ny,
                      H->p.alpha_f);
    { /* setSolverQyy_d, :160-181 */
        const double _2inv_dy2 = 2 / (H->dy * H->dy);
        const double *d_y = H->Di + (long)(i + 1) * ny, *inv_d_y = H->inv_Di + (long)(i + 1) * ny,
                     *dragTerm = H->DragTerm + (long)i * ny, *nud_y = H->nuD_y + (long)i * ny;
        for (j = 0; j < ny; j++) {
            int isl = !isWater0[j];
            _d[j] = (MUT(21) * inv_dt + dragTerm[j] + MUT(20) * (g * d_y[j] * dt * _2inv_dy2) +
                     MUT(10) * (2 * nud_y[j] * inv_d_y[j] * _2inv_dy2)) *
                        isWater0[j] + isl;
        }
    }
    {
        const double inv_dy = 1 / H->dy;
        const double inv_dy2 = inv_dy * inv_dy;
        const double *d_y = H->Di + (long)(i + 1) * ny, *inv_d_y = H->inv_Di + (long)(i + 1) * ny,
                     *v = H->V + (long)(i + 1) * ny, *nud_y = H->nuD_y + (long)i * ny,
                     *dnud_dy = H->dnuD_dy + (long)i * ny;
        /* setSolverQyy_low, :206-230 */
        for (j = 1; j < ny; j++) {
            double _isWater = isWater[j] && isWater[j - 1];
            _low[j] = (0.5 * inv_dy * (MUT(9) * -v[j - 1] + MUT(10) * (2 * dnud_dy[j] * inv_d_y[j - 1])) -
                       inv_dy2 * (MUT(20) * (g * d_y[j] * dt) + MUT(10) * (2 * nud_y[j] * inv_d_y[j - 1]))) * _isWater;
        }
        {
            double _isWater = isWater[0] && isWater[ny - 1];
            _low[0] = (0.5 * inv_dy * (MUT(9) * -v[ny - 1] + MUT(10) * (2 * dnud_dy[0] * inv_d_y[ny - 1])) -
                       inv_dy2 * (MUT(20) * (g * d_y[0] * dt) + MUT(10) * (2 * nud_y[0] * inv_d_y[ny - 1]))) * _isWater;
        }
        /* setSolverQyy_up, :183-204 */
        for (j = 0; j < ny - 1; j++) {
            double _isWater = isWater[j] && isWater[j + 1];
            _up[j] = (0.5 * inv_dy * (MUT(9) * v[j + 1] - MUT(10) * (2 * dnud_dy[j] * inv_d_y[j + 1])) -
                      inv_dy2 * (MUT(20) * (g * d_y[j] * dt) + MUT(10) * (2 * nud_y[j] * inv_d_y[j + 1]))) * _isWater;
        }
        {
            double _isWater = isWater[ny - 1] && isWater[0];
            _up[ny - 1] = (0.5 * inv_dy * (MUT(9) * v[0] - MUT(10) * (2 * dnud_dy[ny - 1] * inv_d_y[0])) -
                           inv_dy2 * (MUT(20) * (g * d_y[ny - 1] * dt) + MUT(10) * (2 * nud_y[ny - 1] * inv_d_y[0]))) * _isWater;
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* Qy projector coefficients: src/HydroQy.cpp:235-362                                 */
/* ---------------------------------------------------------------------------------- */
static void setSolverQyx(orc_hydro *H, int i)
{
    const int ny = H->ny;
    const double dt = H->p.dt;
    const double *dragTerm = H->DragTerm + (long)i * ny, *nud = H->nuD_y + (long)i * ny,
                 *dnud_dx = H->dnuDy_dx + (long)i * ny;
    const unsigned char *isWater = H->IsWater + (long)i * ny;
    double *_f = H->f + (long)i * ny, *_d = H->d + (long)i * ny, *_up = H->up + (long)i * ny,
           *_low = H->low + (long)i * ny;
    int j;
    { /* setSolverQyx_f, :235-248 */
        const double *Qy = H->Qy + (long)(i + 1) * ny;
        for (j = 0; j < ny; j++) _f[j] = isWater[j] * Qy[j] * (1 + dt * dragTerm[j]);
    }
    { /* setSolverQyx_d, :250-285 */
        const double x2c_d = H->x2c_d[i];
        const double x2c_up = H->x2c_up[i];
        const double *inv_d = H->inv_Di + (long)(i + 1) * ny, *u = H->U + (long)i * ny;
        const double dx_d = H->difx_c_d[i];
        const double d2x_d = H->dif2x_c_d[i];
        const double inv_dx = 1 / H->dx_x[i];
        for (j = 0; j < ny - 1; j++) {
            const int isl = !isWater[j];
            const double du_dx = 0.5 * inv_dx * ((u[j + ny + 1] + u[j + ny]) - (u[j + 1] + u[j]));
            const double u_y = 0.5 * (x2c_up * (u[ny + j + 1] + u[ny + j]) + x2c_d * (u[j + 1] + u[j]));
            _d[j] = isWater[j] * ((1 + dt * dragTerm[j]) + MUT(13) * (dt * (u_y * dx_d + du_dx)) -
                                  MUT(14) * (dt * (inv_d[j] * nud[j] * d2x_d + inv_d[j] * dnud_dx[j] * dx_d))) + isl;
        }
        {
            const int isl = !isWater[ny - 1];
            const double du_dx = 0.5 * inv_dx * ((u[ny] + u[2 * ny - 1]) - (u[0] + u[ny - 1]));
            const double u_y = 0.5 * (x2c_up * (u[ny] + u[2 * ny - 1]) + x2c_d * (u[0] + u[ny - 1]));
            _d[ny - 1] = isWater[ny - 1] * ((1 + dt * dragTerm[ny - 1]) + MUT(13) * (dt * (u_y * dx_d + du_dx)) -
                                            MUT(14) * (dt * (inv_d[ny - 1] * nud[ny - 1] * d2x_d + inv_d[ny - 1] * dnud_dx[ny - 1] * dx_d))) +
                         isl;
        }
    }
    { /* setSolverQyx_up, :287-311 */
        const double *inv_d_y_up = H->inv_Di + (long)(2 + i) * ny, *u = H->Ui_j + (long)i * ny;
        const double dx_up = H->difx_c_up[i];
        const double d2x_up = H->dif2x_c_up[i];
        for (j = 0; j < ny; j++) {
            double _isWater = isWater[j] && isWater[j + ny];
            _up[j] = _isWater * dt * (MUT(13) * (u[j] * dx_up) - MUT(14) * (nud[j] * d2x_up * inv_d_y_up[j]) -
                                      MUT(14) * (dnud_dx[j] * dx_up * inv_d_y_up[j]));
        }
    }
    { /* setSolverQyx_low, :313-338, and setSolverQyx0_low, :340-362 */
        const double *inv_d_y_low = H->inv_Di + (long)i * ny, *u = H->Ui_j + (long)i * ny;
        const double dx_low = H->difx_c_low[i];
        const double d2x_low = H->dif2x_c_low[i];
        if (i > 0) {
            for (j = 0; j < ny; j++) {
                double _isWater = isWater[j] && isWater[j - ny];
                _low[j] = _isWater * dt * (MUT(13) * (u[j] * dx_low) - MUT(14) * (nud[j] * d2x_low * inv_d_y_low[j]) -
                                           MUT(14) * (dnud_dx[j] * dx_low * inv_d_y_low[j]));
            }
        } else {
            for (j = 0; j < ny; j++)
                _low[j] = isWater[j] * dt * (MUT(13) * (u[j] * dx_low) - MUT(14) * (nud[j] * d2x_low * inv_d_y_low[j]) -
                                             MUT(14) * (dnud_dx[j] * dx_low * inv_d_y_low[j]));
        }
    }
}

/* ---------------------------------------------------------------------------------- */
/* solveQy: src/HydroQy.cpp:17-85                                                     */
/* ---------------------------------------------------------------------------------- */
static void solveQy(orc_hydro *H)
{
    const int nx = H->nx, ny = H->ny;
    const long size = H->size;
    int i;

    /* :26-42. Masks first: row 0 of IsWater is read by every row's f and d (Q1). */
    getD_y0(H);
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) {
        setWaterNodes_y(H, i);
        getD_y(H, i);
    }
    getD_y(H, nx - 1); /* :42 */
#pragma omp parallel
    {
        double *ws = zalloc(5L * ny);
        double *tmp = zalloc(ny);
        int ii;
#pragma omp for schedule(static)
        for (ii = 0; ii < nx - 1; ii++) {
            setSolverQyy(H, ii);
            tridiag_periodic_ws(H->d + (long)ii * ny, H->low + (long)ii * ny, H->up + (long)ii * ny,
                                H->f + (long)ii * ny, ny, tmp, ws);
            memcpy(H->Qy + (long)(ii + 1) * ny, tmp, sizeof(double) * (size_t)ny);
        }
        free(tmp);
        free(ws);
    }

    /* :52-64 projector coefficients */
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx - 1; i++) setSolverQyx(H, i);
    { /* setSolverQyx_BC, src/HydroBoundaryConditions.cpp:41-60 */
        const double bc = H->bcOffshoreQy;
        for (int j = 0; j < ny; j++) H->d[j] += bc * H->low[j];
        const double inv_3 = 1.0 / 3.0;
        double *d = H->d + size - 2L * ny, *up = H->up + size - 2L * ny;
        for (int j = 0; j < ny; j++) d[j] += up[j] * inv_3;
    }
    /* :67 */
    orc_tridiag(H->Qy + ny, H->v_tmp, H->d, H->low, H->up, H->f, nx - 1, ny);

    { /* :73 setQy_BC, src/HydroBoundaryConditions.cpp:62-79 */
        const double bc = H->bcOffshoreQy;
        for (int j = 0; j < ny; j++) H->Qy[j] = bc * H->Qy[ny + j];
        const double inv_3 = 1.0 / 3.0;
        double *qy0 = H->Qy + size;
        const double *qy1 = qy0 - ny;
        for (int j = 0; j < ny; j++) qy0[j] = qy1[j] * inv_3;
    }

    /* :76-83 */
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        difY2C_row(H->dQy_dy + (long)i * ny, H->Qy + (long)(i + 1) * ny, H->dy, ny);
        solveZs(H, i);
        getV(H, i);
        Y2C_row(H->V_c + (long)i * ny, H->V + (long)(i + 1) * ny, ny);
    }
    getV0(H); /* :84 */
}

/* ---------------------------------------------------------------------------------- */
/* Hydro::Run, src/Hydro.cpp:170-202                                                  */
/* ---------------------------------------------------------------------------------- */
int orc_run(orc_hydro *H, int nIter)
{
    const int nx = H->nx, ny = H->ny;
    getForcing_x(H);
    getForcing_y(H);
    C2X(H->Urms_x, H->Urms, nx, ny);
    C2Y(H->Urms_y, H->Urms, nx, ny);
    calculateDragAndViscosity(H);
    C2X(H->cD_x, H->cD, nx, ny);
    C2Y(H->cD_y, H->cD, nx, ny);
    difC2X(H->dnuD_dx, H->nuD, H->dx, nx, ny);
    C2X(H->nuD_x, H->nuD, nx, ny);
    C2Y(H->nuD_y, H->nuD, nx, ny);
    difC2C_y(H->dnuDx_dy, H->nuD_x, H->dy, nx - 1, ny);
    difC2Y(H->dnuD_dy, H->nuD, H->dy, nx, ny);
    difC2C_x_addBoundary(H->dnuDy_dx, H->nuD_y, H->difx_c_d, H->difx_c_up, H->difx_c_low, nx, ny);
    for (int k = 0; k < nIter; k++) {
        if (H->f0_series && H->f0_pos < H->f0_len) H->bcOffshoreQx_f0 = H->f0_series[H->f0_pos];
        H->f0_pos++;
        solveQx(H);
        solveQy(H);
    }
    return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Construction: src/Bathymetry.cpp:33-39, src/Hydro.cpp:17-133                      */
/* ---------------------------------------------------------------------------------- */
orc_hydro *orc_create(int nx, int ny, const double *x, double dy, const orc_params *p,
                      const double *Z, const double *h, double zs0)
{
    if (nx < 3 || ny < 3 || !x || !p || !Z || !h) return NULL;
    orc_hydro *H = (orc_hydro *)calloc(1, sizeof(orc_hydro));
    long size = (long)nx * ny;
    H->nx = nx; H->ny = ny; H->size = size; H->dy = dy; H->p = *p;
    H->metrics = zalloc(20L * nx);
    orc_grid_metrics(nx, x, H->metrics);
    {
        double *q = H->metrics;
        H->dx = q; q += nx - 1; H->dx_x = q; q += nx;
        H->difx_x_up = q; q += nx - 1; H->difx_x_d = q; q += nx - 1; H->difx_x_low = q; q += nx - 1;
        H->dif2x_x_up = q; q += nx - 1; H->dif2x_x_d = q; q += nx - 1; H->dif2x_x_low = q; q += nx - 1;
        H->difx_c_up = q; q += nx; H->difx_c_d = q; q += nx; H->difx_c_low = q; q += nx;
        H->dif2x_c_up = q; q += nx; H->dif2x_c_d = q; q += nx; H->dif2x_c_low = q; q += nx;
        H->x2c_up = q; q += nx; H->x2c_d = q;
    }
    H->Z = zalloc(size); H->h = zalloc(size); H->D = zalloc(size);
    memcpy(H->Z, Z, sizeof(double) * (size_t)size);
    memcpy(H->h, h, sizeof(double) * (size_t)size);
    for (long q = 0; q < size; q++) { /* src/Bathymetry.cpp:33-39 */
        double Zb = Z[q] - h[q];
        double d = Zb + zs0;
        if (d < p->Dmin) d = p->Dmin;
        H->D[q] = d;
    }
    H->Sxx = zalloc(size); H->Sxy = zalloc(size); H->Syy = zalloc(size); H->Urms = zalloc(size);
    H->Dw = zalloc(size); H->D_r = zalloc(size); H->Hw = zalloc(size);
    H->U = zalloc(size + ny); H->Qx = zalloc(size + ny); H->U_c = zalloc(size);
    H->V = zalloc(size + ny); H->Qy = zalloc(size + ny); H->V_c = zalloc(size);
    H->dQx_dx = zalloc(size); H->dQy_dy = zalloc(size); H->Zs = zalloc(size);
    for (long q = 0; q < size; q++) H->Zs[q] = H->D[q] - Z[q] + h[q]; /* src/Hydro.cpp:35-37 */
    { /* src/Hydro.cpp:44-60 */
        const double lSponge = H->dx_x[0] * 5;
        const double exp_dx_2ls = exp(-0.5 * H->dx_x[0] / lSponge);
        const double exp_dx_ls = exp(-H->dx_x[0] / lSponge);
        if (p->offshore_bc == ORC_BC_SPONGE) {
            H->bcOffshoreQx_f1 = 0.5 * exp_dx_2ls / (1 - 0.5 * exp_dx_2ls);
            H->bcOffshoreQx_f0 = 0;
        } else {
            H->bcOffshoreQx_f1 = 1;
            H->bcOffshoreQx_f0 = 0;
        }
        H->bcOffshoreQy = exp_dx_ls;
    }
    H->cD = zalloc(size); H->nuD = zalloc(size); H->cD_x = zalloc(size - ny); H->cD_y = zalloc(size);
    H->Urms_x = zalloc(size - ny); H->Urms_y = zalloc(size); H->nuD_x = zalloc(size - ny); H->nuD_y = zalloc(size);
    H->dnuD_dx = zalloc(size - ny); H->dnuD_dy = zalloc(size); H->dnuDx_dy = zalloc(size - ny);
    H->dnuDy_dx = zalloc(size); H->Forcing_x = zalloc(size); H->Forcing_y = zalloc(size);
    H->guess_D = zalloc(size + ny); H->guess_Zs = zalloc(size + ny); H->Di = zalloc(size + ny);
    H->inv_Di = zalloc(size + ny); H->Ui_j = zalloc(size); H->DragTerm = zalloc(size);
    H->IsWater = (unsigned char *)calloc((size_t)size, 1);
    H->d = zalloc(size); H->f = zalloc(size); H->up = zalloc(size); H->low = zalloc(size); H->v_tmp = zalloc(size);
    return H;
}

void orc_destroy(orc_hydro *H)
{
    if (!H) return;
    double **fields[] = {&H->metrics, &H->Z, &H->h, &H->D, &H->Sxx, &H->Sxy, &H->Syy, &H->Urms, &H->Dw,
                         &H->D_r, &H->Hw, &H->U, &H->Qx, &H->U_c, &H->V, &H->Qy, &H->V_c, &H->dQx_dx,
                         &H->dQy_dy, &H->Zs, &H->cD, &H->nuD, &H->cD_x, &H->cD_y, &H->Urms_x, &H->Urms_y,
                         &H->nuD_x, &H->nuD_y, &H->dnuD_dx, &H->dnuD_dy, &H->dnuDx_dy, &H->dnuDy_dx,
                         &H->Forcing_x, &H->Forcing_y, &H->guess_D, &H->guess_Zs, &H->Di, &H->inv_Di,
                         &H->Ui_j, &H->DragTerm, &H->d, &H->f, &H->up, &H->low, &H->v_tmp, &H->f0_series};
    for (size_t q = 0; q < sizeof(fields) / sizeof(fields[0]); q++) free(*fields[q]);
    free(H->IsWater);
    free(H);
}

void orc_set_forcing(orc_hydro *H, const double *Sxx, const double *Sxy, const double *Syy,
                     const double *Urms, const double *Dw, const double *D_r, const double *Hw)
{
    size_t nb = sizeof(double) * (size_t)H->size;
    if (Sxx) memcpy(H->Sxx, Sxx, nb);
    if (Sxy) memcpy(H->Sxy, Sxy, nb);
    if (Syy) memcpy(H->Syy, Syy, nb);
    if (Urms) memcpy(H->Urms, Urms, nb);
    if (Dw) memcpy(H->Dw, Dw, nb);
    if (D_r) memcpy(H->D_r, D_r, nb);
    if (Hw) memcpy(H->Hw, Hw, nb);
}

void orc_set_f0(orc_hydro *H, double f0)
{
    H->bcOffshoreQx_f0 = f0;
    free(H->f0_series);
    H->f0_series = NULL;
    H->f0_len = H->f0_pos = 0;
}

int orc_set_f0_series(orc_hydro *H, const double *f0, long n)
{
    free(H->f0_series);
    H->f0_series = NULL;
    H->f0_len = H->f0_pos = 0;
    if (!f0 || n <= 0) return 0;
    H->f0_series = zalloc(n);
    memcpy(H->f0_series, f0, sizeof(double) * (size_t)n);
    H->f0_len = n;
    return 0;
}

void orc_bc_constants(const orc_hydro *H, double *f1, double *f0, double *bcQy)
{
    if (f1) *f1 = H->bcOffshoreQx_f1;
    if (f0) *f0 = H->bcOffshoreQx_f0;
    if (bcQy) *bcQy = H->bcOffshoreQy;
}

void orc_get_state(const orc_hydro *H, double *Zs, double *D, double *Qx, double *Qy, double *U,
                   double *V, double *U_c, double *V_c, double *dQx_dx, double *dQy_dy)
{
    size_t F = sizeof(double) * (size_t)H->size, Fp = sizeof(double) * (size_t)(H->size + H->ny);
    if (Zs) memcpy(Zs, H->Zs, F);
    if (D) memcpy(D, H->D, F);
    if (Qx) memcpy(Qx, H->Qx, Fp);
    if (Qy) memcpy(Qy, H->Qy, Fp);
    if (U) memcpy(U, H->U, Fp);
    if (V) memcpy(V, H->V, Fp);
    if (U_c) memcpy(U_c, H->U_c, F);
    if (V_c) memcpy(V_c, H->V_c, F);
    if (dQx_dx) memcpy(dQx_dx, H->dQx_dx, F);
    if (dQy_dy) memcpy(dQy_dy, H->dQy_dy, F);
}

void orc_set_state(orc_hydro *H, const double *Zs, const double *D, const double *Qx, const double *Qy,
                   const double *U, const double *V, const double *U_c, const double *V_c,
                   const double *dQx_dx, const double *dQy_dy)
{
    size_t F = sizeof(double) * (size_t)H->size, Fp = sizeof(double) * (size_t)(H->size + H->ny);
    if (Zs) memcpy(H->Zs, Zs, F);
    if (D) memcpy(H->D, D, F);
    if (Qx) memcpy(H->Qx, Qx, Fp);
    if (Qy) memcpy(H->Qy, Qy, Fp);
    if (U) memcpy(H->U, U, Fp);
    if (V) memcpy(H->V, V, Fp);
    if (U_c) memcpy(H->U_c, U_c, F);
    if (V_c) memcpy(H->V_c, V_c, F);
    if (dQx_dx) memcpy(H->dQx_dx, dQx_dx, F);
    if (dQy_dy) memcpy(H->dQy_dy, dQy_dy, F);
}

long orc_get_field(const orc_hydro *H, const char *name, double *out)
{
    struct { const char *n; const double *p; long len; } tab[] = {
        {"cD", H->cD, H->size}, {"nuD", H->nuD, H->size}, {"Forcing_x", H->Forcing_x, H->size - H->ny},
        {"Forcing_y", H->Forcing_y, H->size}, {"Urms_x", H->Urms_x, H->size - H->ny}, {"Urms_y", H->Urms_y, H->size},
        {"cD_x", H->cD_x, H->size - H->ny}, {"cD_y", H->cD_y, H->size}, {"nuD_x", H->nuD_x, H->size - H->ny},
        {"nuD_y", H->nuD_y, H->size}, {"dnuD_dx", H->dnuD_dx, H->size - H->ny},
        {"dnuDx_dy", H->dnuDx_dy, H->size - H->ny}, {"dnuD_dy", H->dnuD_dy, H->size},
        {"dnuDy_dx", H->dnuDy_dx, H->size}};
    for (size_t q = 0; q < sizeof(tab) / sizeof(tab[0]); q++)
        if (strcmp(tab[q].n, name) == 0) {
            if (out) memcpy(out, tab[q].p, sizeof(double) * (size_t)tab[q].len);
            return tab[q].len;
        }
    return -1;
}

/* ================================================================================== */
/* Sediment transport and bed update (src/Sediment.cpp)                               */
/* ================================================================================== */
/* include/OptionsParametersConstants.h:15,17,22-23 */
static const double pi_c = 3.141592653589793;
static const double s_rel = 2.65;
static const double tanPhi_i = 0.6249;

struct orc_sediment {
    orc_hydro *H;
    orc_sed_params sp;
    double *Alpha, *gamma, *Gamma, *qx, *qy, *dh_dt, *h, *Zb, *maxDw, *Zs_sh;
};

orc_sediment *orc_sed_create(orc_hydro *H, const orc_sed_params *sp)
{   /* Sediment::Sediment, src/Sediment.cpp:13-31: zeroed arrays */
    orc_sediment *S = (orc_sediment *)calloc(1, sizeof(orc_sediment));
    if (!S) return NULL;
    const long size = H->size;
    long q;
    S->H = H;
    S->sp = *sp;
    S->Alpha = zalloc(size); S->gamma = zalloc(size); S->Gamma = zalloc(size);
    S->qx = zalloc(size); S->qy = zalloc(size); S->dh_dt = zalloc(size);
    S->h = zalloc(size); S->Zb = zalloc(size);
    S->maxDw = zalloc(H->ny); S->Zs_sh = zalloc(H->ny);
    for (q = 0; q < size; q++) {   /* bathy->h, bathy->Zb = Z - h (src/Bathymetry.cpp:34) */
        S->h[q] = H->h[q];
        S->Zb[q] = H->Z[q] - H->h[q];
    }
    return S;
}

void orc_sed_destroy(orc_sediment *S)
{
    if (!S) return;
    free(S->Alpha); free(S->gamma); free(S->Gamma); free(S->qx); free(S->qy); free(S->dh_dt);
    free(S->h); free(S->Zb); free(S->maxDw); free(S->Zs_sh);
    free(S);
}

/* Sediment::getMaxWaveBreakingDissipation, :252-262: running maximum over every call */
static void sed_max_dw(orc_sediment *S)
{
    const int nx = S->H->nx, ny = S->H->ny;
    const double *dw = S->H->Dw;
    int i, j;
#pragma omp parallel for private(i) schedule(static)
    for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++) {
            const double a = S->maxDw[j], b = dw[(long)i * ny + j];
            S->maxDw[j] = (a < b) ? b : a;   /* std::max(a, b) */
        }
}

/* Sediment::getShorelineZs, :264-283: Zs of the last row i >= 1 where rows i-1 and i
 * are not both dry; the previous value stays otherwise */
static void sed_shoreline(orc_sediment *S)
{
    const int nx = S->H->nx, ny = S->H->ny;
    const double *zs = S->H->Zs, *d = S->H->D;
    const double Dmin = S->H->p.Dmin + 1E-4;
    int i, j;
#pragma omp parallel for private(i) schedule(static)
    for (j = 0; j < ny; j++)
        for (i = 1; i < nx; i++) {
            const int isDry = (d[(long)i * ny + j] <= Dmin);
            const int isDryLow = (d[(long)(i - 1) * ny + j] <= Dmin);
            const int isSh = !(isDry * isDryLow);
            S->Zs_sh[j] = (1 - isSh) * S->Zs_sh[j] + zs[(long)i * ny + j] * isSh;
        }
}

/* Sediment::getWaveStirringCWS, :222-239 */
static void sed_stirring(orc_sediment *S, int i)
{
    const int ny = S->H->ny;
    const double alpha0_min = S->sp.alpha0_min;
    const double incrAlpha = S->sp.alpha0_max - alpha0_min;
    const double Dmax = S->sp.D_alpha0_max;
    const double Dmax_min = S->sp.D_alpha0_min - Dmax;
    const double *d = S->H->D + (long)i * ny;
    double *alpha = S->Alpha + (long)i * ny;
    int j;
    for (j = 0; j < ny; j++) {
        double x = 1 + 2 * (Dmax - d[j]) / Dmax_min;
        alpha[j] = alpha0_min + incrAlpha * 0.5 * (1 + tanh(pi_c * x));
    }
}

/* Sediment::getGammaDownslopeCWS, :194-220 (the EGU form; sf_drySwash is unused) */
static void sed_downslope(orc_sediment *S, int i)
{
    const int ny = S->H->ny;
    const double inv_tanPhi_i = 1 / tanPhi_i;
    const double gamma0_tanPhi = S->sp.gamma0 * inv_tanPhi_i;
    const double *d = S->H->D + (long)i * ny;
    const double Dmin = S->H->p.Dmin;
    const double A_swash = S->sp.A_swash;
    double *g_ = S->gamma + (long)i * ny;
    int j;
    for (j = 0; j < ny; j++) {
        const double u_bsl = (tanh(pi_c * (2 * (d[j] - Dmin) / (A_swash - Dmin) - 1)) + 1) / 2;
        g_[j] = u_bsl * gamma0_tanPhi;
    }
}

/* Sediment::getGammaSwash, :241-266 (abs resolves to the double overload) */
static void sed_swash(orc_sediment *S, int i)
{
    const int ny = S->H->ny;
    const double G0 = S->sp.Gamma0;
    const double Dmin = S->H->p.Dmin + 1E-4;
    const double A_swash = S->sp.A_swash;
    const double inv_rho_g_sm1 = 1 / (g * rho * (s_rel - 1));
    const double *d = S->H->D + (long)i * ny, *zb = S->Zb + (long)i * ny;
    double *G_ = S->Gamma + (long)i * ny;
    const double _2_Doff_on = 2 / (A_swash - Dmin);
    int j;
    for (j = 0; j < ny; j++) {
        const int isDry = (d[j] <= Dmin);
        const double d_dry = (1 - isDry) * d[j] + isDry * (fabs(zb[j] + S->Zs_sh[j]));
        const double sf_swash = 1 - (tanh(pi_c * (_2_Doff_on * (d_dry - Dmin) - 1)) + 1) * 0.5;
        G_[j] = G0 * inv_rho_g_sm1 * sf_swash * S->maxDw[j];
    }
}

/* Sediment::getqx, :156-174: rows 0..nx-2, U row i+1 */
static void sed_qx(orc_sediment *S, int i)
{
    const int ny = S->H->ny;
    const double inv_dx = 1 / S->H->dx[i];
    const double *g_ = S->gamma + (long)i * ny, *A = S->Alpha + (long)i * ny, *G_ = S->Gamma + (long)i * ny;
    const double *h = S->h + (long)i * ny, *U = S->H->U + (long)(i + 1) * ny;
    double *qx = S->qx + (long)i * ny;
    int j;
    for (j = 0; j < ny; j++) {
        double A_x = 0.5 * (A[j] + A[j + ny]);
        double g_x = 0.5 * (g_[j] + g_[j + ny]);
        double G_x = 0.5 * (G_[j] + G_[j + ny]);
        double dh_dx = (h[j + ny] - h[j]) * inv_dx;
        qx[j] = A_x * (U[j] - g_x * dh_dx) - G_x * dh_dx;
    }
}

/* Sediment::getqy, :176-192: rows 0..nx-1, V row i+1, periodic in j */
static void sed_qy(orc_sediment *S, int i)
{
    const int ny = S->H->ny;
    const double inv_dy = 1 / S->H->dy;
    const double *g_ = S->gamma + (long)i * ny, *A = S->Alpha + (long)i * ny, *G_ = S->Gamma + (long)i * ny;
    const double *h = S->h + (long)i * ny, *V = S->H->V + (long)(i + 1) * ny;
    double *qy = S->qy + (long)i * ny;
    int j;
    for (j = 0; j < ny; j++) {
        const int jp = (j == ny - 1) ? 0 : j + 1;
        double A_y = 0.5 * (A[j] + A[jp]);
        double g_y = 0.5 * (g_[j] + g_[jp]);
        double G_y = 0.5 * (G_[j] + G_[jp]);
        double dh_dy = (h[jp] - h[j]) * inv_dy;
        qy[j] = A_y * (V[j] - g_y * dh_dy) - G_y * dh_dy;
    }
}

void orc_sed_fluxes(orc_sediment *S)
{   /* :47-65. The reference's OpenMP loop computes row i's coefficients and then qx of row
     * i-1, which reads row i-1's coefficients that another thread may not have written
     * yet; its serial order is the intended one. Every row's coefficients read only
     * inputs, so computing all of them first and then the fluxes gives the serial
     * order's values, race-free in parallel. */
    const int nx = S->H->nx;
    int i;
    sed_max_dw(S);
    sed_shoreline(S);
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        sed_stirring(S, i);
        sed_downslope(S, i);
        sed_swash(S, i);
    }
#pragma omp parallel for schedule(static)
    for (i = 0; i < nx; i++) {
        if (i < nx - 1) sed_qx(S, i);
        sed_qy(S, i);
    }
}

void orc_sed_update_bottom(orc_sediment *S, double dt)
{   /* Sediment::updateBottomColumn, :105-151, rows 1..nx-1 (qx row nx-1 is never
     * computed and stays zero) */
    const int nx = S->H->nx, ny = S->H->ny;
    const double dt_2 = 0.5 * dt;
    const double inv_dy = 1 / S->H->dy;
    const double p = S->sp.p;
    const double _1_m1p = 1 / (1 - p);
    int i;
#pragma omp parallel for schedule(static)
    for (i = 1; i < nx; i++) {
        const double inv_dx = 1 / S->H->dx_x[i];
        const double *qx = S->qx + (long)(i - 1) * ny, *qy = S->qy + (long)i * ny;
        double *h = S->h + (long)i * ny, *Zb = S->Zb + (long)i * ny, *Zs = S->H->Zs + (long)i * ny;
        double *m1 = S->dh_dt + (long)i * ny;
        int j;
        for (j = 0; j < ny; j++) {
            const int jm = (j == 0) ? ny - 1 : j - 1;
            double dqy_dy = inv_dy * (qy[j] - qy[jm]);
            double dqx_dx = inv_dx * (qx[ny + j] - qx[j]);
            double dh_dt = -_1_m1p * (dqx_dx + dqy_dy);
            double incr_h = dt_2 * (3 * dh_dt - m1[j]);   /* Adams-Bashforth 2 */
            h[j] += incr_h;
            Zb[j] -= incr_h;
            Zs[j] += incr_h;
            m1[j] = dh_dt;
        }
    }
}

void orc_sed_set_fluxes(orc_sediment *S, const double *qx, const double *qy)
{
    const int nx = S->H->nx, ny = S->H->ny;
    const double p = S->sp.p;
    const double _1_m1p = 1 / (1 - p);
    const double inv_dy = 1 / S->H->dy;
    int i, j;
    memcpy(S->qx, qx, sizeof(double) * (size_t)S->H->size);
    memcpy(S->qy, qy, sizeof(double) * (size_t)S->H->size);
    for (i = 1; i < nx; i++) {
        const double inv_dx = 1 / S->H->dx_x[i];
        for (j = 0; j < ny; j++) {
            const int jm = (j == 0) ? ny - 1 : j - 1;
            double dqy_dy = inv_dy * (S->qy[(long)i * ny + j] - S->qy[(long)i * ny + jm]);
            double dqx_dx = inv_dx * (S->qx[(long)i * ny + j] - S->qx[(long)(i - 1) * ny + j]);
            S->dh_dt[(long)i * ny + j] = -_1_m1p * (dqx_dx + dqy_dy);
        }
    }
}

long orc_sed_get(orc_sediment *S, const char *name, double *out)
{
    const long F = S->H->size, ny = S->H->ny;
    struct { const char *n; double *p; long len; } tab[] = {
        {"Alpha", S->Alpha, F}, {"gamma", S->gamma, F}, {"Gamma", S->Gamma, F}, {"qx", S->qx, F},
        {"qy", S->qy, F}, {"dh_dt", S->dh_dt, F}, {"h", S->h, F}, {"Zb", S->Zb, F},
        {"maxDw", S->maxDw, ny}, {"Zs_sh", S->Zs_sh, ny}};
    size_t q;
    for (q = 0; q < sizeof(tab) / sizeof(tab[0]); q++)
        if (strcmp(tab[q].n, name) == 0) {
            if (out) memcpy(out, tab[q].p, sizeof(double) * (size_t)tab[q].len);
            return tab[q].len;
        }
    return -1;
}
