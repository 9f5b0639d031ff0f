/*
 * wave_oracle.c -- CPU restatement of the wave and rollers driver (TEST INFRASTRUCTURE;
 * see wave_oracle.h for scope, the three places given their evident meaning, and how the
 * restatement is pinned). Serial in x as the reference; OpenMP over the nodes of a row in
 * each phase, which are independent within the phase.
 */
#include "wave_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* include/OptionsParametersConstants.h:14-21 */
static const double g = 9.8100000000000000;
static const double pi = 3.141592653589793;
static const double rho = 1025.00000000000;
static const double sqrt_pi = 1.772453850905516;
#define CTE_ENERGY (0.125 * rho * g)
#define INV_2PI (0.5 / pi)

static double smin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static double smax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* One WaveRow (include/WaveRow.h:22-75): per-node arrays of one cross-shore position. */
typedef struct {
    double dx;
    double *D, *lK, *U, *V, *U_w, *TanTheta, *CosTheta, *SinTheta, *Theta, *Sigma, *C, *Cg, *dw, *df, *sxx, *sxy,
        *syy, *sxx_r, *sxy_r, *syy_r, *lE, *E, *Dw, *H, *E_r, *d_r, *Cx, *Cy, *Cgx, *Cgy, *Urms_H, *inv_S_i, *Y_i, *G,
        *Gp, *WC, *WC_r;
} wrow;
#define WROW_ARRAYS 37

struct orc_wave {
    int nx, ny;
    double dy, inv_2dy;
    double *x, *dx;
    orc_wave_params p;
    double *D, *U, *V, *U_w;
    double *K, *Theta, *Cg, *H, *E, *Urms, *Sxx_w, *Sxy_w, *Syy_w, *Dw, *E_r, *D_r, *Sxx_r, *Sxy_r, *Syy_r, *Sxx, *Sxy,
        *Syy, *CosTheta, *SinTheta;
    wrow rows[2];
};

static double *zalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static double **row_arrays(wrow *r)
{
    return &r->D;   /* the WROW_ARRAYS pointers are laid out consecutively after dx */
}

/* ---------------------------------------------------------------------------------
 * Per-node parametrizations (src/WaveParametrizations.cpp)
 * --------------------------------------------------------------------------------- */

/* WaveRow::dispersion (:14-60): Doppler-shifted linear dispersion by Newton-Raphson on kD. */
void orc_wave_dispersion(double d, double u_w, double w, double z0, double *lK, double *sigma_o, double *C_o,
                         double *Cg_o, double *Urms_H)
{
    const int maxIterations = 10;
    const double tolerance = 1E-14;
    const double inv_sqrtgh = 1 / sqrt(g * d);
    double alpha = u_w * inv_sqrtgh;
    const double mu = w * w * d / g;
    const double sqrtmu = sqrt(mu);
    double k = w * inv_sqrtgh;
    double kD = k * d;
    double tanhkD = tanh(kD);
    int iter = 1;
    double err = 9999;
    while (iter < maxIterations && err > tolerance) {
        const double mu_ = (sqrtmu - alpha * kD);
        const double f = kD * tanhkD - mu_ * mu_;
        const double df_dkD = tanhkD + kD * (1 - tanhkD * tanhkD) + 2 * alpha * mu_;
        const double incrkD = f / df_dkD;
        kD += -incrkD;
        tanhkD = tanh(kD);
        err = fabs(f);
        iter += 1;
    }
    k = kD / d;
    const double sigma = sqrt(g * k * tanhkD);
    *sigma_o = sigma;
    *lK = log(k);
    const double cc = sigma / k;
    *C_o = cc;
    *Cg_o = 0.5 * (1 + k * d * (1.0 / tanhkD - tanhkD)) * cc;
    const double coshkd_2 = 1 / (1 - tanhkD * tanhkD);
    *Urms_H = 0.5 * g * cosh(k * z0) / (cc * sqrt(coshkd_2));
}

/* waveBreakingDissipationCT (:155-203), TG (:206-251), Regular (:254-302): Dw/E and its
 * derivative in H. */
void orc_wave_breaking(const orc_wave_params *p, double h, double sigma, double D, double *dw, double *ddw_dh)
{
    if (p->breaking_model == ORC_BREAK_CT) {
        double inv_gamma_b = 1 / p->gamma_b;
        double inv_4sqrtpi = 0.25 / sqrt_pi;
        double B33 = 3 * p->B3;
        double inv_D = 1 / D;
        double f1 = B33 * sigma * inv_4sqrtpi * inv_D;
        double dcc_dh = inv_D * inv_gamma_b;
        double cc = h * dcc_dh;
        double cc_2 = cc * cc;
        double inv_1pcc_2 = 1 / (1 + cc_2);
        double pow_2d5 = inv_1pcc_2 * inv_1pcc_2 * sqrt(inv_1pcc_2);
        double pow_3d5 = inv_1pcc_2 * pow_2d5;
        double f2 = 1 - pow_2d5;
        double df2_h = -2.5 * pow_3d5 * 2 * cc_2;
        double tanh8ccm1 = tanh(8 * (cc - 1));
        double f3 = 1 + tanh8ccm1;
        double tanh8ccm1_2 = tanh8ccm1 * tanh8ccm1;
        double df3_h = (1 - tanh8ccm1_2) * 8. * dcc_dh;
        *dw = h * f1 * f2 * f3;
        *ddw_dh = f1 * (f2 * f3 + h * (df2_h * f3 + df3_h * f2));
    } else if (p->breaking_model == ORC_BREAK_TG) {
        const double inv_gamma_b = 1 / p->gamma_b;
        const double kappa_sigma = 3 * p->B3 / (4 * sqrt_pi * p->gamma_b * p->gamma_b);
        const double kappa = kappa_sigma * sigma;
        const double inv_D = 1 / D;
        const double gamma = h * inv_D;
        const double g2 = gamma * gamma;
        const double g3 = g2 * gamma;
        const double dcc_dh = inv_D * inv_gamma_b;
        const double cc = h * dcc_dh;
        const double ff = (1 + cc * cc);
        const double ff1 = pow(ff, -2.5);
        const double ff2 = (1 - ff1);
        const double ff3 = ff1 / ff;
        *dw = kappa * g3 * ff2;
        *ddw_dh = kappa * g2 * (3 * ff2 * inv_D + 2.5 * gamma * ff3 * 2 * cc * dcc_dh);
    } else {   /* ORC_BREAK_REGULAR */
        const double inv_gamma_b = 1 / p->gamma_b;
        const double B3_pi = p->B3 / pi;
        const double m_mon = p->m_mon, n_mon = p->n_mon;
        const double inv_D = 1 / D;
        const double dcc_dh = inv_D * inv_gamma_b;
        const double cc = h * dcc_dh;
        const double cc_to_n = pow(cc, n_mon);
        const double cc_to_m = pow(cc, m_mon);
        const double cc_to_mm1 = (smax(0.0, cc_to_m / cc));
        const double cc_to_nm1 = (smax(0.0, cc_to_n / cc));
        const double f1 = B3_pi * inv_D * sigma;
        const double f2 = cc_to_m;
        const double f3 = -expm1(-cc_to_n);
        const double df2_dh = m_mon * (cc_to_mm1) * dcc_dh;
        const double df3_dh = (f3 - 1) * (-n_mon * (cc_to_nm1)) * dcc_dh;
        *dw = f1 * h * f2 * f3;
        *ddw_dh = f1 * (f2 * f3 + h * df2_dh * f3 + h * df3_dh * f2);
    }
}

/* bottomFrictionConstant (:305-339) and bottomFrictionSoulsby97 (:342-388), per node
 * (see the header for the indexing); sigma0 is the row's first-node frequency. */
static void wave_friction(const orc_wave_params *p, double h, double sigma0, double urms_h, double *df, double *ddf_dh)
{
    if (p->friction_model == ORC_FRIC_CONSTANT) {
        double _64fw_6pig = 21.333333333 * p->fw0 * INV_2PI / g;
        double urms_h_cube = urms_h;
        urms_h_cube = urms_h_cube * urms_h_cube * urms_h_cube;
        *ddf_dh = _64fw_6pig * urms_h_cube;
        *df = *ddf_dh * h;
    } else {
        double _64_6pig = 21.3333333 * INV_2PI / g;
        double urms = urms_h * h;
        double fw = 1.39 * sqrt(p->z0 * sigma0 / urms);
        double urms_h_cube = urms_h * urms_h * urms_h;
        *ddf_dh = _64_6pig * fw * urms_h_cube;
        *df = *ddf_dh * h;
    }
}

/* ---------------------------------------------------------------------------------
 * Row operations
 * --------------------------------------------------------------------------------- */

static void row_dispersion(orc_wave *W, wrow *r, double w)
{   /* WaveRow::dispersion_parallel (:108-128) */
    const int ny = W->ny;
    int j;
#pragma omp parallel for schedule(static)
    for (j = 0; j < ny; j++)
        orc_wave_dispersion(r->D[j], r->U_w[j], w, W->p.z0, &r->lK[j], &r->Sigma[j], &r->C[j], &r->Cg[j],
                            &r->Urms_H[j]);
}

/* Radiation-stress factors of a direction (src/WaveRow.cpp:171-190, src/WaveDir.cpp:74-99). */
static void row_direction_terms(wrow *r, int j, double theta, double cosTheta, double sinTheta, double cossinTheta)
{
    const double c = r->C[j], cg = r->Cg[j];
    r->Cx[j] = c * cosTheta;
    r->Cgx[j] = cg * cosTheta;
    r->Cy[j] = c * sinTheta;
    r->Cgy[j] = cg * sinTheta;
    const double cg_c = cg / c;
    const double cosTheta2 = cosTheta * cosTheta;
    r->sxx[j] = ((1 + cosTheta2) * cg_c - 0.5);
    r->syy[j] = ((2 - cosTheta2) * cg_c - 0.5);
    r->sxy[j] = (cossinTheta * cg_c);
    r->sxx_r[j] = 2 * cosTheta2;
    r->syy_r[j] = 2 * (1 - cosTheta2);
    r->sxy_r[j] = 2 * cossinTheta;
    r->Theta[j] = theta;
    r->CosTheta[j] = cosTheta;
    r->SinTheta[j] = sinTheta;
    r->TanTheta[j] = sinTheta / cosTheta;
}

/* WaveRow::initFirstRow (src/WaveRow.cpp:129-232). */
static void first_row(orc_wave *W, wrow *r, double hOff, double thetaOff, double T)
{
    const int ny = W->ny;
    int j;
    memcpy(r->D, W->D, sizeof(double) * (size_t)ny);
    memcpy(r->U, W->U, sizeof(double) * (size_t)ny);
    memcpy(r->V, W->V, sizeof(double) * (size_t)ny);
    memcpy(r->U_w, W->U_w, sizeof(double) * (size_t)ny);
    const double w = 2 * pi / T;
    row_dispersion(W, r, w);
    /* node 0's depth, the absolute frequency and the offshore height (:143-146) */
    double dw0 = 0, ddw_dh0 = 0, df0 = 0, ddf_dh0 = 0;
    orc_wave_breaking(&W->p, hOff, w, r->D[0], &dw0, &ddw_dh0);
    if (W->p.bottom_friction) wave_friction(&W->p, hOff, w, r->Urms_H[0], &df0, &ddf_dh0);
    const double EOff = hOff * hOff * CTE_ENERGY;
    const double lEOff = log(EOff);
#pragma omp parallel for schedule(static)
    for (j = 0; j < ny; j++) {
        const double cosTheta = cos(thetaOff);
        const double sinTheta = sin(thetaOff);
        row_direction_terms(r, j, thetaOff, cosTheta, sinTheta, sinTheta * cosTheta);
        r->H[j] = hOff;
        r->E[j] = EOff;
        r->Dw[j] = dw0 * EOff;
        r->lE[j] = lEOff;
        r->dw[j] = dw0;
        r->df[j] = df0;
    }
}

/* WaveRow::initRow, both forms (src/WaveRow.cpp:235-296). interp: the row lies dx into
 * mesh cell xIndex (width dx_grid); D, U, V, U_w move from row xIndex by the fraction
 * dx/dx_grid of the step itself (not of the distance from row xIndex). */
static void init_row(orc_wave *W, wrow *r, const wrow *rm1, double dx, int interp, double dx_grid, long index,
                     long index_m1)
{
    const int ny = W->ny;
    int j;
    r->dx = dx;
    memcpy(r->H, rm1->H, sizeof(double) * (size_t)ny);
    memcpy(r->lE, rm1->lE, sizeof(double) * (size_t)ny);
    if (!interp) {
        memcpy(r->D, W->D + index, sizeof(double) * (size_t)ny);
        memcpy(r->U, W->U + index, sizeof(double) * (size_t)ny);
        memcpy(r->V, W->V + index, sizeof(double) * (size_t)ny);
        memcpy(r->U_w, W->U_w + index, sizeof(double) * (size_t)ny);
        return;
    }
    const double dx_dx = dx / dx_grid;
#pragma omp parallel for schedule(static)
    for (j = 0; j < ny; j++) {
        r->D[j] = W->D[index_m1 + j] + (W->D[index + j] - W->D[index_m1 + j]) * dx_dx;
        r->U[j] = W->U[index_m1 + j] + (W->U[index + j] - W->U[index_m1 + j]) * dx_dx;
        r->V[j] = W->V[index_m1 + j] + (W->V[index + j] - W->V[index_m1 + j]) * dx_dx;
        r->U_w[j] = W->U_w[index_m1 + j] + (W->U_w[index + j] - W->U_w[index_m1 + j]) * dx_dx;
    }
}

/* WaveRow::solveDir (src/WaveDir.cpp:12-126) with the periodic extension of
 * solveDir_parallel (:129-163). */
static void solve_dir(orc_wave *W, wrow *r, const wrow *rm1)
{
    const int ny = W->ny;
    const double dx = r->dx, inv_2dy = W->inv_2dy;
    const double dx2 = dx * dx;
    const double inv_dx = 1 / dx;
    const double thetaMax = W->p.theta_max;
    int j;
#pragma omp parallel for schedule(static)
    for (j = 0; j < ny; j++) {
        const int jp = (j == ny - 1) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
        double tt_p1 = rm1->TanTheta[jp];
        double tt_m1 = rm1->TanTheta[jm];
        double dttheta_dy = (tt_p1 - tt_m1) * inv_2dy;
        double ttheta_med = (tt_p1 + tt_m1) * 0.5;
        double y_i = -ttheta_med / (dttheta_dy + inv_dx);
        double s_i = sqrt(dx2 + y_i * y_i);
        double inv_s_i = 1 / s_i;
        double sinTheta_i = -y_i * inv_s_i;
        double cosTheta_i = dx * inv_s_i;
        double theta_i = asin(sinTheta_i);
        r->Y_i[j] = y_i;
        r->inv_S_i[j] = inv_s_i;
        double dlk_dx = (r->lK[j] - rm1->lK[j]) * inv_dx;
        double dlk_dy_c = 0.5 * inv_2dy * (r->lK[jp] + rm1->lK[jp] - r->lK[jm] - rm1->lK[jm]);
        double dlk_ds = dlk_dx * cosTheta_i + dlk_dy_c * sinTheta_i;
        double dlk_dst = -dlk_dx * sinTheta_i + dlk_dy_c * cosTheta_i;
        double theta = theta_i + dlk_dst / (inv_s_i + dlk_ds * 0.5);
        theta = smin(theta, thetaMax);
        theta = smax(theta, -thetaMax);
        const double sinTheta = sin(theta);
        const double cosTheta = cos(theta);
        row_direction_terms(r, j, theta, cosTheta, sinTheta, cosTheta * sinTheta);
    }
}

/* Value at the foot of the incident ray: linear interpolation between the neighbours
 * j+1, j-1 of the previous row (src/WaveEnergy.cpp:188-193). */
static double at_ray(const double *f, int up, int low, double y_i_dy)
{
    return (f[up] + f[low]) * 0.5 + (f[up] - f[low]) * y_i_dy;
}

/* WaveRow::energyPreprocessor (src/WaveEnergy.cpp:130-287). */
static void energy_preprocessor(orc_wave *W, wrow *r, const wrow *rm1)
{
    const int ny = W->ny;
    const double inv_dx = 1 / r->dx, inv_2dy = W->inv_2dy;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ny; i++) {
        const int up = (i == ny - 1) ? 0 : i + 1, low = (i == 0) ? ny - 1 : i - 1;
        double y_i_dy = r->Y_i[i] * inv_2dy;
        double inv_s_i = r->inv_S_i[i];
        double cg_c = 0.5 * (r->Cg[i] + at_ray(rm1->Cg, up, low, y_i_dy));
        double u_c = 0.5 * (r->U[i] + at_ray(rm1->U, up, low, y_i_dy));
        double v_c = 0.5 * (r->V[i] + at_ray(rm1->V, up, low, y_i_dy));
        double dw_i = at_ray(rm1->dw, up, low, y_i_dy);
        double df_i = at_ray(rm1->df, up, low, y_i_dy);
        double le_i = at_ray(rm1->lE, up, low, y_i_dy);
        double du_dx = (r->U[i] - rm1->U[i]) * inv_dx;
        double du_dy = (r->U[up] - r->U[low]) * inv_2dy;
        double dv_dx = (r->V[i] - rm1->V[i]) * inv_dx;
        double dv_dy = (r->V[up] - r->V[low]) * inv_2dy;
        double wc = 2 * r->sxx[i] * du_dx + r->sxy[i] * (du_dy + dv_dx) + 2 * r->syy[i] * dv_dy;
        double wc_c = 0.5 * (wc + at_ray(rm1->WC, up, low, y_i_dy));
        r->WC[i] = wc;
        r->WC_r[i] = r->sxx_r[i] * du_dx + r->sxy_r[i] * (du_dy + dv_dx) + r->syy_r[i] * dv_dy;
        double dcgxu_dx = (r->Cgx[i] - rm1->Cgx[i]) * inv_dx + du_dx;
        double dcgyv_dy_c = 0.5 * (r->Cgy[up] + rm1->Cgy[up] - r->Cgy[low] - rm1->Cgy[low]) * inv_2dy + dv_dy;
        double div_cgv_c = dcgxu_dx + dcgyv_dy_c;
        double dle_dy_c = (rm1->lE[up] - rm1->lE[low]) * inv_2dy;
        r->G[i] = v_c * dle_dy_c + wc_c + div_cgv_c + 0.5 * (dw_i + df_i) - cg_c * inv_s_i * le_i -
                  u_c * rm1->lE[i] * inv_dx;
        r->Gp[i] = cg_c * inv_s_i + u_c * inv_dx;
        r->lE[i] = rm1->lE[i];
    }
}

/* WaveRow::solveEnergy (src/WaveEnergy.cpp:32-127): Newton-Raphson on log(E) per node,
 * a node leaving the iteration once its update is within the tolerance. */
static void solve_energy(orc_wave *W, wrow *r)
{
    const int ny = W->ny;
    const orc_wave_params *p = &W->p;
    const double logCteEnergy = log(CTE_ENERGY);
    const double sigma0 = r->Sigma[0];
    int j;
#pragma omp parallel for schedule(static)
    for (j = 0; j < ny; j++) {
        double le_old = r->lE[j], H = r->H[j], E = 0.0, Dw = 0.0;
        double dw = 0.0, ddw_dh = 0.0, df = 0.0, ddf_dh = 0.0;
        for (int k = 0; k < p->max_iterations; k++) {
            orc_wave_breaking(p, H, r->Sigma[j], r->D[j], &dw, &ddw_dh);
            if (p->bottom_friction) wave_friction(p, H, sigma0, r->Urms_H[j], &df, &ddf_dh);
            double gg = r->G[j];
            double gp = r->Gp[j];
            double le = le_old;
            gg += gp * le + 0.5 * (dw + df);
            gp += gp + 0.25 * (ddf_dh + ddw_dh) * H;
            le -= gg / gp;
            double h = exp((le - logCteEnergy) * 0.5);
            H = h;
            E = CTE_ENERGY * h * h;
            const int more = fabs(le - le_old) > p->tolerance;
            le_old = le;
            Dw = dw * E;
            if (!more) break;
        }
        r->lE[j] = le_old;
        r->E[j] = E;
        r->H[j] = H;
        r->dw[j] = dw;
        r->df[j] = df;
        r->Dw[j] = Dw;
    }
}

/* WaveRow::solveRollers (src/WaveRollers.cpp:10-106). */
static void solve_rollers(orc_wave *W, wrow *r, const wrow *rm1)
{
    const int ny = W->ny;
    const double inv_dx = 1 / r->dx, inv_2dy = W->inv_2dy;
    const double d_r_div_c = 2 * g * W->p.sinBeta_r;
    const double alpha_r = W->p.alpha_r;
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ny; i++) {
        const int up = (i == ny - 1) ? 0 : i + 1, low = (i == 0) ? ny - 1 : i - 1;
        double y_i_dy = r->Y_i[i] * inv_2dy;
        double inv_s_i = r->inv_S_i[i];
        double c_c = 0.5 * (r->C[i] + at_ray(rm1->C, up, low, y_i_dy));
        double u_c = 0.5 * (r->U[i] + at_ray(rm1->U, up, low, y_i_dy));
        double v_c = 0.5 * (r->V[i] + at_ray(rm1->V, up, low, y_i_dy));
        double du_dx = (r->U[i] - rm1->U[i]) * inv_dx;
        double dv_dy = (r->V[up] - r->V[low]) * inv_2dy;
        double wc_r_c = 0.5 * (r->WC_r[i] + at_ray(rm1->WC_r, up, low, y_i_dy));
        double dcxu_dx = (r->Cx[i] - rm1->Cx[i]) * inv_dx + du_dx;
        double dcyv_dy_c = 0.5 * (r->Cy[up] + rm1->Cy[up] - r->Cy[low] - rm1->Cy[low]) * inv_2dy + dv_dy;
        double div_cv_c = dcxu_dx + dcyv_dy_c;
        double der_dy_c = (rm1->E_r[up] - rm1->E_r[low]) * inv_2dy;
        const double d_r = d_r_div_c / r->C[i];
        r->d_r[i] = d_r;
        double d_r_c = 0.5 * (d_r + at_ray(rm1->d_r, up, low, y_i_dy));
        double Dw_c = 0.5 * (r->Dw[i] + at_ray(rm1->Dw, up, low, y_i_dy));
        double er_i = at_ray(rm1->E_r, up, low, y_i_dy);
        double upv = c_c * inv_s_i * er_i + u_c * rm1->E_r[i] * inv_dx - v_c * der_dy_c + 0.5 * alpha_r * Dw_c;
        double lowv = c_c * inv_s_i + u_c * inv_dx + div_cv_c + 0.5 * wc_r_c + 0.5 * d_r_c;
        r->E_r[i] = upv / lowv;
    }
}

/* Wave::storageRow (src/Wave.cpp:175-264). */
static void store_row(orc_wave *W, int xIndex, const wrow *r)
{
    const int ny = W->ny;
    const long o = (long)xIndex * ny;
    int j;
#pragma omp parallel for schedule(static)
    for (j = 0; j < ny; j++) {
        W->Theta[o + j] = r->Theta[j];
        W->CosTheta[o + j] = r->CosTheta[j];
        W->SinTheta[o + j] = r->SinTheta[j];
        W->Cg[o + j] = r->Cg[j];
        W->H[o + j] = r->H[j];
        W->Dw[o + j] = r->Dw[j];
        W->K[o + j] = r->lK[j];
        const double e = r->E[j], e_r = r->E_r[j];
        const double sxx_w = r->sxx[j] * e, sxy_w = r->sxy[j] * e, syy_w = r->syy[j] * e;
        const double sxx_r = r->sxx_r[j] * e_r, sxy_r = r->sxy_r[j] * e_r, syy_r = r->syy_r[j] * e_r;
        W->E[o + j] = e;
        W->Sxx_w[o + j] = sxx_w;
        W->Sxy_w[o + j] = sxy_w;
        W->Syy_w[o + j] = syy_w;
        W->E_r[o + j] = e_r;
        W->D_r[o + j] = r->d_r[j] * e_r;
        W->Sxx_r[o + j] = sxx_r;
        W->Sxy_r[o + j] = sxy_r;
        W->Syy_r[o + j] = syy_r;
        W->Sxx[o + j] = sxx_w + sxx_r;
        W->Sxy[o + j] = sxy_w + sxy_r;
        W->Syy[o + j] = syy_w + syy_r;
        W->Urms[o + j] = r->Urms_H[j] * r->H[j];
    }
}

/* Wave::getDx (src/Wave.cpp:267-290). */
static int get_dx(const orc_wave *W, double x0, int xIndex, const wrow *rm1, double *dx_row)
{
    double tanThetaMax = -999;
    for (int i = 0; i < W->ny; i++) {
        const double t = fabs(rm1->TanTheta[i]);
        if (t > tanThetaMax) tanThetaMax = t;
    }
    const double dx_remain = W->x[xIndex + 1] - x0;
    if (tanThetaMax == 0) {
        *dx_row = dx_remain;
        return 1;
    }
    *dx_row = W->dy / tanThetaMax;
    if (*dx_row < dx_remain) {
        *dx_row = *dx_row * 0.9;
        return 0;
    }
    *dx_row = dx_remain;
    return 1;
}

/* ---------------------------------------------------------------------------------
 * API
 * --------------------------------------------------------------------------------- */
orc_wave *orc_wave_create(int nx, int ny, const double *x, double dy, const orc_wave_params *p)
{
    orc_wave *W = (orc_wave *)calloc(1, sizeof(orc_wave));
    if (!W) return NULL;
    const long F = (long)nx * ny;
    W->nx = nx;
    W->ny = ny;
    W->dy = dy;
    W->inv_2dy = 0.5 / dy;
    W->p = *p;
    W->x = zalloc(nx);
    W->dx = zalloc(nx);
    memcpy(W->x, x, sizeof(double) * (size_t)nx);
    for (int i = 0; i < nx - 1; i++) W->dx[i] = x[i + 1] - x[i];
    double **f[] = {&W->D, &W->U, &W->V, &W->U_w, &W->K, &W->Theta, &W->Cg, &W->H, &W->E, &W->Urms, &W->Sxx_w,
                    &W->Sxy_w, &W->Syy_w, &W->Dw, &W->E_r, &W->D_r, &W->Sxx_r, &W->Sxy_r, &W->Syy_r, &W->Sxx,
                    &W->Sxy, &W->Syy, &W->CosTheta, &W->SinTheta};
    for (size_t q = 0; q < sizeof(f) / sizeof(f[0]); q++) *f[q] = zalloc(F);
    for (int r = 0; r < 2; r++) {
        double **a = row_arrays(&W->rows[r]);
        for (int q = 0; q < WROW_ARRAYS; q++) a[q] = zalloc(ny);
    }
    return W;
}

void orc_wave_destroy(orc_wave *W)
{
    if (!W) return;
    double *f[] = {W->x, W->dx, W->D, W->U, W->V, W->U_w, W->K, W->Theta, W->Cg, W->H, W->E, W->Urms, W->Sxx_w,
                   W->Sxy_w, W->Syy_w, W->Dw, W->E_r, W->D_r, W->Sxx_r, W->Sxy_r, W->Syy_r, W->Sxx, W->Sxy, W->Syy,
                   W->CosTheta, W->SinTheta};
    for (size_t q = 0; q < sizeof(f) / sizeof(f[0]); q++) free(f[q]);
    for (int r = 0; r < 2; r++) {
        double **a = row_arrays(&W->rows[r]);
        for (int q = 0; q < WROW_ARRAYS; q++) free(a[q]);
    }
    free(W);
}

void orc_wave_set_depth(orc_wave *W, const double *D)
{
    memcpy(W->D, D, sizeof(double) * (size_t)W->nx * (size_t)W->ny);
}

void orc_wave_set_uv(orc_wave *W, const double *U, const double *V)
{
    const long F = (long)W->nx * W->ny;
    long q;
    memcpy(W->U, U, sizeof(double) * (size_t)F);
    memcpy(W->V, V, sizeof(double) * (size_t)F);
    for (q = 0; q < F; q++) W->U_w[q] = W->CosTheta[q] * W->U[q] + W->SinTheta[q] * W->V[q];
}

long orc_wave_calculate(orc_wave *W, double hOff, double thetaOff, double T)
{
    const int nx = W->nx, ny = W->ny;
    if (W->p.breaking_model < ORC_BREAK_CT || W->p.breaking_model > ORC_BREAK_REGULAR) return -1;
    double x0 = W->x[0];
    int xIndex = 0;
    int prev = 0, cur = 1;
    long steps = 0;
    const long max_steps = 1000L * nx + 1000;
    /* the reference builds both rows afresh for every field (src/Wave.cpp:115-116), with
     * the arrays the first row does not set (WC, WC_r, E_r, d_r) zero */
    for (int rr = 0; rr < 2; rr++) {
        double **a = row_arrays(&W->rows[rr]);
        for (int q = 0; q < WROW_ARRAYS; q++) memset(a[q], 0, sizeof(double) * (size_t)ny);
    }
    first_row(W, &W->rows[prev], hOff, thetaOff, T);
    store_row(W, 0, &W->rows[prev]);
    const double w = 2 * pi / T;
    while (xIndex < nx - 1) {
        double dx_row;
        wrow *rm1 = &W->rows[prev], *r = &W->rows[cur];
        const int store = get_dx(W, x0, xIndex, rm1, &dx_row);
        const long index = (long)(xIndex + 1) * ny, index_m1 = (long)xIndex * ny;
        init_row(W, r, rm1, dx_row, !store, W->dx[xIndex], index, index_m1);
        row_dispersion(W, r, w);
        solve_dir(W, r, rm1);
        energy_preprocessor(W, r, rm1);
        solve_energy(W, r);
        if (W->p.rollers) solve_rollers(W, r, rm1);
        if (store) {
            xIndex = xIndex + 1;
            store_row(W, xIndex, r);
        }
        prev ^= 1;
        cur ^= 1;
        x0 = x0 + dx_row;
        if (++steps > max_steps) return -2;
    }
    return steps;
}

long orc_wave_get(const orc_wave *W, const char *name, double *out)
{
    const long F = (long)W->nx * W->ny;
    struct { const char *n; const double *p; } tab[] = {
        {"K", W->K}, {"Theta", W->Theta}, {"Cg", W->Cg}, {"H", W->H}, {"E", W->E}, {"Urms", W->Urms},
        {"Sxx_w", W->Sxx_w}, {"Sxy_w", W->Sxy_w}, {"Syy_w", W->Syy_w}, {"Dw", W->Dw}, {"E_r", W->E_r},
        {"D_r", W->D_r}, {"Sxx_r", W->Sxx_r}, {"Sxy_r", W->Sxy_r}, {"Syy_r", W->Syy_r}, {"Sxx", W->Sxx},
        {"Sxy", W->Sxy}, {"Syy", W->Syy}, {"CosTheta", W->CosTheta}, {"SinTheta", W->SinTheta}, {"U_w", W->U_w}};
    for (size_t q = 0; q < sizeof(tab) / sizeof(tab[0]); q++)
        if (strcmp(tab[q].n, name) == 0) {
            if (out) memcpy(out, tab[q].p, sizeof(double) * (size_t)F);
            return F;
        }
    return -1;
}
