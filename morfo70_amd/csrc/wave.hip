// wave.hip -- the wave and rollers driver (class Wave, class WaveRow: src/Wave.cpp,
// src/WaveRow.cpp, src/WaveDir.cpp, src/WaveEnergy.cpp, src/WaveParametrizations.cpp,
// src/WaveRollers.cpp) for gfx950.
//
// The field marches offshore to onshore, one row after the other; a row depends on the
// whole previous row (the step length takes the maximum |tan(theta)| over it), and within
// a row the direction needs the neighbours' wave numbers and the energy balance the
// neighbours' directions. So one workgroup owns a field: its threads take the nodes of a
// row in strides and barriers separate (depth, dispersion) -> (direction) -> (energy,
// rollers, storage); the row arrays live in global memory, resident in L2 (37 arrays of
// ny doubles per row, two rows). Control flow (mesh rows and intermediate rows, the step
// count) follows src/Wave.cpp:80-172, with the two row buffers swapped after every step
// (the reference's brace-less `else` at :158-163 swaps them only after intermediate rows,
// aliasing the buffers otherwise). Each node's Newton-Raphson, friction terms included,
// is its own (src/WaveEnergy.cpp:62-115 with the per-node indexing; friction is off by
// default). Arithmetic keeps the reference's operation order (-ffp-contract=off);
// transcendental functions are the device library's.
#include "wave.h"

namespace m70 {

// include/OptionsParametersConstants.h:14-21
constexpr double WG = 9.8100000000000000;
constexpr double WPI = 3.141592653589793;
constexpr double WRHO = 1025.00000000000;
constexpr double WSQRT_PI = 1.772453850905516;
constexpr double CTE_ENERGY = 0.125 * WRHO * WG;
constexpr double WINV_2PI = 0.5 / WPI;

__device__ __forceinline__ double smin_(double a, double b) { return (b < a) ? b : a; }   // std::min
__device__ __forceinline__ double smax_(double a, double b) { return (a < b) ? b : a; }   // std::max

// WaveRow::dispersion (src/WaveParametrizations.cpp:14-60).
__device__ void w_dispersion(double d, double u_w, double w, double z0, double &lK, double &sigma_o, double &C_o,
                             double &Cg_o, double &Urms_H)
{
    const int maxIterations = 10;
    const double tolerance = 1E-14;
    const double inv_sqrtgh = 1 / sqrt(WG * d);
    double alpha = u_w * inv_sqrtgh;
    const double mu = w * w * d / WG;
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
    const double sigma = sqrt(WG * k * tanhkD);
    sigma_o = sigma;
    lK = log(k);
    const double cc = sigma / k;
    C_o = cc;
    Cg_o = 0.5 * (1 + k * d * (1.0 / tanhkD - tanhkD)) * cc;
    const double coshkd_2 = 1 / (1 - tanhkD * tanhkD);
    Urms_H = 0.5 * WG * cosh(k * z0) / (cc * sqrt(coshkd_2));
}

// Breaking dissipation Dw/E and d(Dw/E)/dH: Church-Thornton (:155-203), Thornton-Guza
// (:206-251), regular waves (:254-302).
__device__ void w_breaking(const m70w_params &p, double h, double sigma, double D, double &dw, double &ddw_dh)
{
    if (p.breaking_model == M70W_BREAK_CT) {
        double inv_gamma_b = 1 / p.gamma_b;
        double inv_4sqrtpi = 0.25 / WSQRT_PI;
        double B33 = 3 * p.B3;
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
        dw = h * f1 * f2 * f3;
        ddw_dh = f1 * (f2 * f3 + h * (df2_h * f3 + df3_h * f2));
    } else if (p.breaking_model == M70W_BREAK_TG) {
        const double inv_gamma_b = 1 / p.gamma_b;
        const double kappa_sigma = 3 * p.B3 / (4 * WSQRT_PI * p.gamma_b * p.gamma_b);
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
        dw = kappa * g3 * ff2;
        ddw_dh = kappa * g2 * (3 * ff2 * inv_D + 2.5 * gamma * ff3 * 2 * cc * dcc_dh);
    } else {
        const double inv_gamma_b = 1 / p.gamma_b;
        const double B3_pi = p.B3 / WPI;
        const double m_mon = p.m_mon, n_mon = p.n_mon;
        const double inv_D = 1 / D;
        const double dcc_dh = inv_D * inv_gamma_b;
        const double cc = h * dcc_dh;
        const double cc_to_n = pow(cc, n_mon);
        const double cc_to_m = pow(cc, m_mon);
        const double cc_to_mm1 = (smax_(0.0, cc_to_m / cc));
        const double cc_to_nm1 = (smax_(0.0, cc_to_n / cc));
        const double f1 = B3_pi * inv_D * sigma;
        const double f2 = cc_to_m;
        const double f3 = -expm1(-cc_to_n);
        const double df2_dh = m_mon * (cc_to_mm1) * dcc_dh;
        const double df3_dh = (f3 - 1) * (-n_mon * (cc_to_nm1)) * dcc_dh;
        dw = f1 * h * f2 * f3;
        ddw_dh = f1 * (f2 * f3 + h * df2_dh * f3 + h * df3_dh * f2);
    }
}

// Bottom friction Df/E and d(Df/E)/dH, constant fw (:305-339) or Soulsby 1997 (:342-388);
// sigma0 is the frequency of the row's first node, as the reference passes it.
__device__ void w_friction(const m70w_params &p, double h, double sigma0, double urms_h, double &df, double &ddf_dh)
{
    if (p.friction_model == M70W_FRIC_CONSTANT) {
        double _64fw_6pig = 21.333333333 * p.fw0 * WINV_2PI / WG;
        double urms_h_cube = urms_h;
        urms_h_cube = urms_h_cube * urms_h_cube * urms_h_cube;
        ddf_dh = _64fw_6pig * urms_h_cube;
        df = ddf_dh * h;
    } else {
        double _64_6pig = 21.3333333 * WINV_2PI / WG;
        double urms = urms_h * h;
        double fw = 1.39 * sqrt(p.z0 * sigma0 / urms);
        double urms_h_cube = urms_h * urms_h * urms_h;
        ddf_dh = _64_6pig * fw * urms_h_cube;
        df = ddf_dh * h;
    }
}

// Value at the foot of the incident ray (src/WaveEnergy.cpp:188-193).
__device__ __forceinline__ double at_ray(const double *f, int up, int low, double y_i_dy)
{
    return (f[up] + f[low]) * 0.5 + (f[up] - f[low]) * y_i_dy;
}

// Direction and radiation-stress factors of a node (src/WaveRow.cpp:171-190,
// src/WaveDir.cpp:74-99).
__device__ __forceinline__ void dir_terms(const WRow &r, int j, double theta, double cosTheta, double sinTheta,
                                          double cossinTheta)
{
    const double c = r.C[j], cg = r.Cg[j];
    r.Cx[j] = c * cosTheta;
    r.Cgx[j] = cg * cosTheta;
    r.Cy[j] = c * sinTheta;
    r.Cgy[j] = cg * sinTheta;
    const double cg_c = cg / c;
    const double cosTheta2 = cosTheta * cosTheta;
    r.sxx[j] = ((1 + cosTheta2) * cg_c - 0.5);
    r.syy[j] = ((2 - cosTheta2) * cg_c - 0.5);
    r.sxy[j] = (cossinTheta * cg_c);
    r.sxx_r[j] = 2 * cosTheta2;
    r.syy_r[j] = 2 * (1 - cosTheta2);
    r.sxy_r[j] = 2 * cossinTheta;
    r.Theta[j] = theta;
    r.CosTheta[j] = cosTheta;
    r.SinTheta[j] = sinTheta;
    r.TanTheta[j] = sinTheta / cosTheta;
}

// Wave::storageRow (src/Wave.cpp:175-264) of node j into row xIndex of the field.
__device__ __forceinline__ void store_node(const WField &f, long o, const WRow &r, int j)
{
    f.Theta[o] = r.Theta[j];
    f.CosTheta[o] = r.CosTheta[j];
    f.SinTheta[o] = r.SinTheta[j];
    f.Cg[o] = r.Cg[j];
    f.H[o] = r.H[j];
    f.Dw[o] = r.Dw[j];
    f.K[o] = r.lK[j];
    const double e = r.E[j], e_r = r.E_r[j];
    const double sxx_w = r.sxx[j] * e, sxy_w = r.sxy[j] * e, syy_w = r.syy[j] * e;
    const double sxx_r = r.sxx_r[j] * e_r, sxy_r = r.sxy_r[j] * e_r, syy_r = r.syy_r[j] * e_r;
    f.E[o] = e;
    f.Sxx_w[o] = sxx_w;
    f.Sxy_w[o] = sxy_w;
    f.Syy_w[o] = syy_w;
    f.E_r[o] = e_r;
    f.D_r[o] = r.d_r[j] * e_r;
    f.Sxx_r[o] = sxx_r;
    f.Sxy_r[o] = sxy_r;
    f.Syy_r[o] = syy_r;
    f.Sxx[o] = sxx_w + sxx_r;
    f.Sxy[o] = sxy_w + sxy_r;
    f.Syy[o] = syy_w + syy_r;
    f.Urms[o] = r.Urms_H[j] * r.H[j];
}

template <int BS>
__global__ void __launch_bounds__(BS) k_wave_field(WaveArgs a, double hOff, double thetaOff, double T)
{
    __shared__ double red[BS];
    const int t = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const m70w_params &p = a.p;
    const double inv_2dy = a.inv_2dy;
    const double w = 2 * WPI / T;
    // Both rows start from zero, as the reference builds them afresh for every field
    // (src/Wave.cpp:115-116; WC, WC_r, E_r, d_r of the first row are read as zero).
    for (int q = 0; q < 2; ++q) {
        double *const *arr = &a.rows[q].D;
        for (int k = 0; k < WROW_ARRAYS; ++k)
            for (int j = t; j < ny; j += BS) arr[k][j] = 0.0;
    }
    // WaveRow::initFirstRow (src/WaveRow.cpp:129-232)
    {
        const WRow &r = a.rows[0];
        for (int j = t; j < ny; j += BS) {
            r.D[j] = a.D[j];
            r.U[j] = a.U[j];
            r.V[j] = a.V[j];
            r.U_w[j] = a.U_w[j];
            w_dispersion(r.D[j], r.U_w[j], w, p.z0, r.lK[j], r.Sigma[j], r.C[j], r.Cg[j], r.Urms_H[j]);
        }
        __syncthreads();
        double dw0 = 0, ddw_dh0 = 0, df0 = 0, ddf_dh0 = 0;   // node 0, absolute frequency, offshore H
        w_breaking(p, hOff, w, r.D[0], dw0, ddw_dh0);
        if (p.bottom_friction) w_friction(p, hOff, w, r.Urms_H[0], df0, ddf_dh0);
        const double EOff = hOff * hOff * CTE_ENERGY;
        const double lEOff = log(EOff);
        for (int j = t; j < ny; j += BS) {
            const double cosTheta = cos(thetaOff);
            const double sinTheta = sin(thetaOff);
            dir_terms(r, j, thetaOff, cosTheta, sinTheta, sinTheta * cosTheta);
            r.H[j] = hOff;
            r.E[j] = EOff;
            r.Dw[j] = dw0 * EOff;
            r.lE[j] = lEOff;
            r.dw[j] = dw0;
            r.df[j] = df0;
            store_node(a.f, j, r, j);
        }
        __syncthreads();
    }
    double x0 = a.x[0];
    int xIndex = 0, prev = 0;
    long steps = 0, err = 0;
    const long max_steps = 1000L * nx + 1000;
    const double logCteEnergy = log(CTE_ENERGY);
    while (xIndex < nx - 1) {
        const WRow &rm1 = a.rows[prev];
        const WRow &r = a.rows[prev ^ 1];
        // Wave::getDx (src/Wave.cpp:267-290): the largest |tan(theta)| of the previous row
        double m = -999;
        for (int j = t; j < ny; j += BS) {
            const double tt = fabs(rm1.TanTheta[j]);
            if (tt > m) m = tt;
        }
        red[t] = m;
        __syncthreads();
        for (int s = BS / 2; s > 0; s >>= 1) {
            if (t < s && red[t + s] > red[t]) red[t] = red[t + s];
            __syncthreads();
        }
        const double tanThetaMax = red[0];
        const double dx_remain = a.x[xIndex + 1] - x0;
        double dx_row;
        int store;
        if (tanThetaMax == 0) {
            dx_row = dx_remain;
            store = 1;
        } else {
            dx_row = a.dy / tanThetaMax;
            if (dx_row < dx_remain) {
                dx_row = dx_row * 0.9;
                store = 0;
            } else {
                dx_row = dx_remain;
                store = 1;
            }
        }
        // WaveRow::initRow (src/WaveRow.cpp:235-296) and dispersion_parallel (:108-128)
        const long index = (long)(xIndex + 1) * ny, index_m1 = (long)xIndex * ny;
        const double dx_dx = dx_row / a.dx[xIndex];
        for (int j = t; j < ny; j += BS) {
            r.H[j] = rm1.H[j];
            r.lE[j] = rm1.lE[j];
            double d, u, v, uw;
            if (store) {
                d = a.D[index + j];
                u = a.U[index + j];
                v = a.V[index + j];
                uw = a.U_w[index + j];
            } else {
                d = a.D[index_m1 + j] + (a.D[index + j] - a.D[index_m1 + j]) * dx_dx;
                u = a.U[index_m1 + j] + (a.U[index + j] - a.U[index_m1 + j]) * dx_dx;
                v = a.V[index_m1 + j] + (a.V[index + j] - a.V[index_m1 + j]) * dx_dx;
                uw = a.U_w[index_m1 + j] + (a.U_w[index + j] - a.U_w[index_m1 + j]) * dx_dx;
            }
            r.D[j] = d;
            r.U[j] = u;
            r.V[j] = v;
            r.U_w[j] = uw;
            w_dispersion(d, uw, w, p.z0, r.lK[j], r.Sigma[j], r.C[j], r.Cg[j], r.Urms_H[j]);
        }
        __syncthreads();
        // WaveRow::solveDir (src/WaveDir.cpp:12-126)
        const double dx = dx_row;
        const double dx2 = dx * dx;
        const double inv_dx = 1 / dx;
        const double thetaMax = p.theta_max;
        for (int j = t; j < ny; j += BS) {
            const int jp = (j == ny - 1) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
            double tt_p1 = rm1.TanTheta[jp];
            double tt_m1 = rm1.TanTheta[jm];
            double dttheta_dy = (tt_p1 - tt_m1) * inv_2dy;
            double ttheta_med = (tt_p1 + tt_m1) * 0.5;
            double y_i = -ttheta_med / (dttheta_dy + inv_dx);
            double s_i = sqrt(dx2 + y_i * y_i);
            double inv_s_i = 1 / s_i;
            double sinTheta_i = -y_i * inv_s_i;
            double cosTheta_i = dx * inv_s_i;
            double theta_i = asin(sinTheta_i);
            r.Y_i[j] = y_i;
            r.inv_S_i[j] = inv_s_i;
            double dlk_dx = (r.lK[j] - rm1.lK[j]) * inv_dx;
            double dlk_dy_c = 0.5 * inv_2dy * (r.lK[jp] + rm1.lK[jp] - r.lK[jm] - rm1.lK[jm]);
            double dlk_ds = dlk_dx * cosTheta_i + dlk_dy_c * sinTheta_i;
            double dlk_dst = -dlk_dx * sinTheta_i + dlk_dy_c * cosTheta_i;
            double theta = theta_i + dlk_dst / (inv_s_i + dlk_ds * 0.5);
            theta = smin_(theta, thetaMax);
            theta = smax_(theta, -thetaMax);
            const double sinTheta = sin(theta);
            const double cosTheta = cos(theta);
            dir_terms(r, j, theta, cosTheta, sinTheta, cosTheta * sinTheta);
        }
        __syncthreads();
        // energyPreprocessor (src/WaveEnergy.cpp:130-287), solveEnergy (:32-127),
        // solveRollers (src/WaveRollers.cpp:10-106) and storageRow, node by node
        const double sigma0 = r.Sigma[0];
        const double d_r_div_c = 2 * WG * p.sinBeta_r;
        const double alpha_r = p.alpha_r;
        for (int i = t; i < ny; i += BS) {
            const int up = (i == ny - 1) ? 0 : i + 1, low = (i == 0) ? ny - 1 : i - 1;
            const double y_i_dy = r.Y_i[i] * inv_2dy;
            const double inv_s_i = r.inv_S_i[i];
            double G, Gp;
            {
                double cg_c = 0.5 * (r.Cg[i] + at_ray(rm1.Cg, up, low, y_i_dy));
                double u_c = 0.5 * (r.U[i] + at_ray(rm1.U, up, low, y_i_dy));
                double v_c = 0.5 * (r.V[i] + at_ray(rm1.V, up, low, y_i_dy));
                double dw_i = at_ray(rm1.dw, up, low, y_i_dy);
                double df_i = at_ray(rm1.df, up, low, y_i_dy);
                double le_i = at_ray(rm1.lE, up, low, y_i_dy);
                double du_dx = (r.U[i] - rm1.U[i]) * inv_dx;
                double du_dy = (r.U[up] - r.U[low]) * inv_2dy;
                double dv_dx = (r.V[i] - rm1.V[i]) * inv_dx;
                double dv_dy = (r.V[up] - r.V[low]) * inv_2dy;
                double wc = 2 * r.sxx[i] * du_dx + r.sxy[i] * (du_dy + dv_dx) + 2 * r.syy[i] * dv_dy;
                double wc_c = 0.5 * (wc + at_ray(rm1.WC, up, low, y_i_dy));
                r.WC[i] = wc;
                r.WC_r[i] = r.sxx_r[i] * du_dx + r.sxy_r[i] * (du_dy + dv_dx) + r.syy_r[i] * dv_dy;
                double dcgxu_dx = (r.Cgx[i] - rm1.Cgx[i]) * inv_dx + du_dx;
                double dcgyv_dy_c = 0.5 * (r.Cgy[up] + rm1.Cgy[up] - r.Cgy[low] - rm1.Cgy[low]) * inv_2dy + dv_dy;
                double div_cgv_c = dcgxu_dx + dcgyv_dy_c;
                double dle_dy_c = (rm1.lE[up] - rm1.lE[low]) * inv_2dy;
                G = v_c * dle_dy_c + wc_c + div_cgv_c + 0.5 * (dw_i + df_i) - cg_c * inv_s_i * le_i -
                    u_c * rm1.lE[i] * inv_dx;
                Gp = cg_c * inv_s_i + u_c * inv_dx;
                r.G[i] = G;
                r.Gp[i] = Gp;
            }
            {   // Newton-Raphson on log(E); the Jacobian's doubled Gp is the reference's (:97)
                double le_old = rm1.lE[i], H = r.H[i], E = 0.0, Dw = 0.0;
                double dw = 0.0, ddw_dh = 0.0, df = 0.0, ddf_dh = 0.0;
                const double sig = r.Sigma[i], Dn = r.D[i], urms_h = r.Urms_H[i];
                for (int k = 0; k < p.max_iterations; k++) {
                    w_breaking(p, H, sig, Dn, dw, ddw_dh);
                    if (p.bottom_friction) w_friction(p, H, sigma0, urms_h, df, ddf_dh);
                    double gg = G;
                    double gp = Gp;
                    double le = le_old;
                    gg += gp * le + 0.5 * (dw + df);
                    gp += gp + 0.25 * (ddf_dh + ddw_dh) * H;
                    le -= gg / gp;
                    double h = exp((le - logCteEnergy) * 0.5);
                    H = h;
                    E = CTE_ENERGY * h * h;
                    const bool more = fabs(le - le_old) > p.tolerance;
                    le_old = le;
                    Dw = dw * E;
                    if (!more) break;
                }
                r.lE[i] = le_old;
                r.E[i] = E;
                r.H[i] = H;
                r.dw[i] = dw;
                r.df[i] = df;
                r.Dw[i] = Dw;
            }
            if (p.rollers) {
                double c_c = 0.5 * (r.C[i] + at_ray(rm1.C, up, low, y_i_dy));
                double u_c = 0.5 * (r.U[i] + at_ray(rm1.U, up, low, y_i_dy));
                double v_c = 0.5 * (r.V[i] + at_ray(rm1.V, up, low, y_i_dy));
                double du_dx = (r.U[i] - rm1.U[i]) * inv_dx;
                double dv_dy = (r.V[up] - r.V[low]) * inv_2dy;
                double wc_r_c = 0.5 * (r.WC_r[i] + at_ray(rm1.WC_r, up, low, y_i_dy));
                double dcxu_dx = (r.Cx[i] - rm1.Cx[i]) * inv_dx + du_dx;
                double dcyv_dy_c = 0.5 * (r.Cy[up] + rm1.Cy[up] - r.Cy[low] - rm1.Cy[low]) * inv_2dy + dv_dy;
                double div_cv_c = dcxu_dx + dcyv_dy_c;
                double der_dy_c = (rm1.E_r[up] - rm1.E_r[low]) * inv_2dy;
                const double d_r = d_r_div_c / r.C[i];
                r.d_r[i] = d_r;
                double d_r_c = 0.5 * (d_r + at_ray(rm1.d_r, up, low, y_i_dy));
                double Dw_c = 0.5 * (r.Dw[i] + at_ray(rm1.Dw, up, low, y_i_dy));
                double er_i = at_ray(rm1.E_r, up, low, y_i_dy);
                double upv = c_c * inv_s_i * er_i + u_c * rm1.E_r[i] * inv_dx - v_c * der_dy_c + 0.5 * alpha_r * Dw_c;
                double lowv = c_c * inv_s_i + u_c * inv_dx + div_cv_c + 0.5 * wc_r_c + 0.5 * d_r_c;
                r.E_r[i] = upv / lowv;
            }
            if (store) store_node(a.f, index + i, r, i);
        }
        __syncthreads();
        prev ^= 1;
        x0 = x0 + dx_row;
        if (store) xIndex = xIndex + 1;
        if (++steps > max_steps) {
            err = 1;
            break;
        }
    }
    if (t == 0) {
        a.status[0] = steps;
        a.status[1] = err;
    }
}

__global__ void __launch_bounds__(256) k_wave_set_uv(double *U, double *V, double *U_w, const double *Usrc,
                                                     const double *Vsrc, const double *cosT, const double *sinT, long n)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const double u = Usrc[q], v = Vsrc[q];
    U[q] = u;
    V[q] = v;
    U_w[q] = cosT[q] * u + sinT[q] * v;
}

hipError_t launch_wave_field(hipStream_t st, const WaveArgs &a, double hOff, double thetaOff, double T)
{
    // 512 threads at most: the node body needs about 250 VGPRs, two waves per SIMD
    if (a.ny >= 512) hipLaunchKernelGGL(k_wave_field<512>, dim3(1), dim3(512), 0, st, a, hOff, thetaOff, T);
    else if (a.ny >= 256) hipLaunchKernelGGL(k_wave_field<256>, dim3(1), dim3(256), 0, st, a, hOff, thetaOff, T);
    else hipLaunchKernelGGL(k_wave_field<64>, dim3(1), dim3(64), 0, st, a, hOff, thetaOff, T);
    return hipGetLastError();
}

hipError_t launch_wave_set_uv(hipStream_t st, const WaveArgs &a, double *U, double *V, double *U_w,
                              const double *Usrc, const double *Vsrc, long n)
{
    hipLaunchKernelGGL(k_wave_set_uv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, U, V, U_w, Usrc, Vsrc,
                       a.f.CosTheta, a.f.SinTheta, n);
    return hipGetLastError();
}

}  // namespace m70
