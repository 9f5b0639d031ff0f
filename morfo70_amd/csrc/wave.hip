// wave.hip -- the wave and rollers driver (class Wave, class WaveRow: src/Wave.cpp,
// src/WaveRow.cpp, src/WaveDir.cpp, src/WaveEnergy.cpp, src/WaveParametrizations.cpp,
// src/WaveRollers.cpp) for gfx950.
//
// The field marches offshore to onshore, one row after the other: a row needs the whole
// previous row (the step length takes the largest |tan(theta)| over it). A marching launch
// computes up to WKMAX rows over workgroups of one wavefront each: lane t holds the same
// node for every row, and a node needs the previous row at its neighbours and the
// neighbours' new directions, so the values near a workgroup's edges go stale one lane per
// row and WHALO halo lanes each side keep the owned lanes exact (the same arithmetic as the
// neighbouring workgroup's owner, so the values agree bit for bit). A launch's first row
// follows the reference's step rule; its later rows are taken to be mesh rows, and the next
// launch checks that against the rule from each row's largest |tan(theta)| (reduced across
// the wavefront by lane shuffles as soon as the direction is known and folded by one atomic
// maximum on the bit patterns of non-negative doubles into the row's accumulator slot),
// keeping the rows that hold. The control state (x0, mesh row, step count, done, rows
// per launch) is double-buffered in device memory, so that launch L reads what launch L-1
// wrote. A mesh row stores the node values only the march computes (Theta, H, E, Dw, E_r);
// the arrays Wave::storageRow derives from them and from the mesh node's dispersion are
// formed for the whole field by one launch after the march (k_wave_field_terms), with the
// same arithmetic. Each marching kernel is instantiated per breaking model and friction
// switch, so the energy iteration is one branch-free block. Control flow (mesh rows,
// intermediate rows, the step count) follows src/Wave.cpp:80-172, with the two row
// buffers swapped after every step (the reference's brace-less `else` at :158-163 swaps
// them only after intermediate rows, aliasing the buffers otherwise). Each node's
// Newton-Raphson, friction terms included, is its own (src/WaveEnergy.cpp:62-115 with the
// per-node indexing; friction is off by default). Arithmetic keeps the reference's
// operation order (-ffp-contract=off); transcendental functions are the device library's.
#include "wave.h"

#include <algorithm>

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
// (:206-251), regular waves (:254-302). BM is the model, or 0 to select it from p.
template <int BM>
__device__ __forceinline__ void w_breaking(const m70w_params &p, double h, double sigma, double D, double &dw,
                                           double &ddw_dh)
{
    const int model = BM ? BM : p.breaking_model;
    if (model == M70W_BREAK_CT) {
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
    } else if (model == M70W_BREAK_TG) {
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
        const double ff3 = 1 / (ff * ff * ff * sqrt(ff));   // ff^-3.5
        const double ff1 = ff3 * ff;                        // ff^-2.5
        const double ff2 = (1 - ff1);
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
__device__ __forceinline__ void w_friction(const m70w_params &p, double h, double sigma0, double urms_h, double &df,
                                           double &ddf_dh)
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

// Value at the foot of the incident ray: linear interpolation between the neighbours
// j+1 (up) and j-1 (low) of the previous row (src/WaveEnergy.cpp:188-193).
__device__ __forceinline__ double at_ray(double f_up, double f_low, double y_i_dy)
{
    return (f_up + f_low) * 0.5 + (f_up - f_low) * y_i_dy;
}

// One node of a row, in registers.
struct WNode {
    double D, U, V, U_w, lK, Sigma, C, Cg, Urms_H;
    double Y_i, inv_s_i, Theta, CosTheta, SinTheta, TanTheta, Cx, Cgx, Cy, Cgy, sxx, syy, sxy, sxx_r, syy_r, sxy_r;
    double lE, E, H, dw, df, Dw, WC, WC_r, E_r, d_r;
};

// Direction and radiation-stress factors (src/WaveRow.cpp:171-190, src/WaveDir.cpp:74-99).
__device__ __forceinline__ void node_dir_terms(WNode &n, double theta, double cosTheta, double sinTheta,
                                               double cossinTheta)
{
    const double c = n.C, cg = n.Cg;
    n.Cx = c * cosTheta;
    n.Cgx = cg * cosTheta;
    n.Cy = c * sinTheta;
    n.Cgy = cg * sinTheta;
    const double cg_c = cg / c;
    const double cosTheta2 = cosTheta * cosTheta;
    n.sxx = ((1 + cosTheta2) * cg_c - 0.5);
    n.syy = ((2 - cosTheta2) * cg_c - 0.5);
    n.sxy = (cossinTheta * cg_c);
    n.sxx_r = 2 * cosTheta2;
    n.syy_r = 2 * (1 - cosTheta2);
    n.sxy_r = 2 * cossinTheta;
    n.Theta = theta;
    n.CosTheta = cosTheta;
    n.SinTheta = sinTheta;
    n.TanTheta = sinTheta / cosTheta;
}

// The arrays of a row that the next step reads as the previous row (row buffer r).
__device__ __forceinline__ void write_row(double *r, int ny, int j, const WNode &n, bool rollers, bool friction)
{
    auto at = [&](int q) -> double & { return r[(long)q * ny + j]; };
    at(WR_H) = n.H;
    at(WR_lE) = n.lE;
    at(WR_TanTheta) = n.TanTheta;
    at(WR_lK) = n.lK;
    at(WR_Cg) = n.Cg;
    at(WR_U) = n.U;
    at(WR_V) = n.V;
    at(WR_dw) = n.dw;
    at(WR_WC) = n.WC;
    at(WR_Cgx) = n.Cgx;
    at(WR_Cgy) = n.Cgy;
    if (friction) at(WR_df) = n.df;
    if (rollers) {
        at(WR_C) = n.C;
        at(WR_WC_r) = n.WC_r;
        at(WR_Cx) = n.Cx;
        at(WR_Cy) = n.Cy;
        at(WR_E_r) = n.E_r;
        at(WR_d_r) = n.d_r;
        at(WR_Dw) = n.Dw;
    }
}

// The node values of Wave::storageRow (src/Wave.cpp:175-264) that only the march computes,
// into element o of the field (E_r is zero without rollers, as allocated).
__device__ __forceinline__ void store_node(const WField &f, long o, const WNode &n, bool rollers)
{
    f.Theta[o] = n.Theta;
    f.H[o] = n.H;
    f.E[o] = n.E;
    f.Dw[o] = n.Dw;
    if (rollers) f.E_r[o] = n.E_r;
}

// Wavefront maximum of v over its lanes (v >= 0 for a valid |tan(theta)|, -1 for a NaN or
// a lane without an owned node), folded by lane 0 into acc as 1 + its bit pattern: non-
// negative doubles order as their bit patterns, and 0 stays "no valid value". Every lane
// of the wavefront calls it.
__device__ __forceinline__ void fold_tan_max(double v, unsigned long long *acc)
{
    for (int o = WB / 2; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, WB);
        v = (w > v) ? w : v;
    }
    if (threadIdx.x == 0 && v >= 0) atomicMax(acc, (unsigned long long)__double_as_longlong(v) + 1ull);
}

// The largest |tan(theta)| of the row a launch reads its step length from (-999 when the
// row had no valid value, the reference's initial value, src/Wave.cpp:267-290).
__device__ __forceinline__ double tan_max_of(unsigned long long v)
{
    return v ? __longlong_as_double((long long)(v - 1ull)) : -999.0;
}

// WaveRow::initFirstRow (src/WaveRow.cpp:129-232) and storageRow(0): one thread per
// node of the workgroup's WNI nodes. This is launch 0 of the march: its one row goes to
// row buffer set 0, slot 1, its largest |tan(theta)| to accumulator set 0, slot 1.
__global__ void __launch_bounds__(WB) k_wave_first(WaveArgs a, double hOff, double thetaOff, double T, WaveCtl *ctl,
                                                   unsigned long long *acc)
{
    __shared__ double s_dw0, s_df0;
    const int t = threadIdx.x, k = blockIdx.x;
    const int ny = a.ny;
    const m70w_params &p = a.p;
    const bool rollers = p.rollers != 0, friction = p.bottom_friction != 0;
    const double w = 2 * WPI / T;
    if (t == 0) {   // node 0's depth, the absolute frequency and the offshore height (:143-146)
        const double Uh = a.disp.Urms_H[0];
        double dw0 = 0, ddw_dh0 = 0, df0 = 0, ddf_dh0 = 0;
        w_breaking<0>(p, hOff, w, a.D[0], dw0, ddw_dh0);
        if (friction) w_friction(p, hOff, w, Uh, df0, ddf_dh0);
        s_dw0 = dw0;
        s_df0 = df0;
    }
    __syncthreads();
    const int j0 = k * WNI;
    const int nint = min(WNI, ny - j0);
    double m = -1;
    if (t < nint) {
        const int j = j0 + t;
        WNode n{};
        n.D = a.D[j];
        n.U = a.U[j];
        n.V = a.V[j];
        n.U_w = a.U_w[j];
        n.lK = a.disp.lK[j];
        n.Sigma = a.disp.Sigma[j];
        n.C = a.disp.C[j];
        n.Cg = a.disp.Cg[j];
        n.Urms_H = a.disp.Urms_H[j];
        const double cosTheta = cos(thetaOff);
        const double sinTheta = sin(thetaOff);
        node_dir_terms(n, thetaOff, cosTheta, sinTheta, sinTheta * cosTheta);
        const double EOff = hOff * hOff * CTE_ENERGY;
        n.H = hOff;
        n.E = EOff;
        n.Dw = s_dw0 * EOff;
        n.lE = log(EOff);
        n.dw = s_dw0;
        n.df = s_df0;
        write_row(a.rows[0], ny, j, n, rollers, friction);
        store_node(a.f, j, n, rollers);
        const double tt = fabs(n.TanTheta);
        if (tt >= 0) m = tt;
    }
    fold_tan_max(m, &acc[1]);
    if (k == 0 && t == 0) {
        WaveCtl c{};
        c.x0 = a.x[0];
        c.done = (a.nx - 1 <= 0);
        c.K = 1;
        ctl[0] = c;
    }
}

// The previous row's values one node's step reads (up = j+1, low = j-1, periodic).
struct PrevRow {
    double tt_p, tt_m, lk_j, lk_p, lk_m;                                       // direction
    double cg_p, cg_m, u_p, u_m, u_j, v_p, v_m, v_j, dw_p, dw_m, df_p, df_m;  // energy
    double le_p, le_m, le_j, wc_p, wc_m, cgx_j, cgy_p, cgy_m, h_j;
    double c_p, c_m, wcr_p, wcr_m, cx_j, cy_p, cy_m, er_p, er_m, er_j, dr_p, dr_m, Dw_p, Dw_m;   // rollers
};

// A node's own previous-row values (the rest of PrevRow is read from the neighbours).
struct PrevOwn {
    double lk, u, v, le, cgx, h, cx, er;
};

// The previous row's values the neighbours read, one LDS array per value, indexed by lane.
enum PrevLds { PL_TT, PL_LK, PL_CG, PL_U, PL_V, PL_DW, PL_DF, PL_LE, PL_WC, PL_CGY,   // always
               PL_C, PL_WCR, PL_CY, PL_ER, PL_DR, PL_DWE,                          // rollers
               PL_N };

// A mesh node's inputs: depth, currents and the dispersion solved before the march.
struct MeshNode {
    double D, U, V, U_w, lK, Sigma, C, Cg, Urms_H;
};

__device__ __forceinline__ MeshNode load_mesh(const WaveArgs &a, long o)
{
    MeshNode m;
    m.D = a.D[o];
    m.U = a.U[o];
    m.V = a.V[o];
    m.U_w = a.U_w[o];
    m.lK = a.disp.lK[o];
    m.Sigma = a.disp.Sigma[o];
    m.C = a.disp.C[o];
    m.Cg = a.disp.Cg[o];
    m.Urms_H = a.disp.Urms_H[o];
    return m;
}

// Wave::getDx (src/Wave.cpp:267-290): the step length from the previous row's largest
// |tan(theta)|; store = 1 when the step reaches the next mesh row.
__device__ __forceinline__ void get_dx(const WaveArgs &a, const WaveCtl &c, double tanThetaMax, double &dx_row,
                                       int &store)
{
    const double dx_remain = a.x[c.xIndex + 1] - c.x0;
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
}

// The marching state after a step of length dx_row (src/Wave.cpp:150-170).
__device__ __forceinline__ WaveCtl advance(const WaveCtl &c, double dx_row, int store, int nx)
{
    WaveCtl nc = c;
    nc.x0 = c.x0 + dx_row;
    if (store) nc.xIndex = c.xIndex + 1;
    nc.steps = c.steps + 1;
    nc.done = (nc.xIndex >= nx - 1);
    if (nc.steps > 1000L * nx + 1000) {
        nc.done = 1;
        nc.err = 1;
    }
    return nc;
}

// Marching launch L: Wave::getDx, WaveRow::initRow (src/WaveRow.cpp:235-296), dispersion
// (src/WaveParametrizations.cpp:14-60), solveDir (src/WaveDir.cpp:12-126),
// energyPreprocessor and solveEnergy (src/WaveEnergy.cpp), solveRollers
// (src/WaveRollers.cpp) and storageRow of mesh rows, for up to WKMAX rows.
//
// Every workgroup first replays the previous launch's rows 2..K under the step rule, from
// the largest |tan(theta)| of each of its rows: the rows up to the first one the rule
// would have made an intermediate row hold (they were computed as mesh rows), and the
// march goes on from the last of them, v. Row 1 of this launch then takes the rule's own
// step from row v, and rows 2..K are again taken to be mesh rows (K doubles after a launch
// whose rows all held, up to WKMAX, and is 1 after one whose rows did not). Each row is
// folded into its own accumulator slot as soon as its direction is known. A row's values
// go to row buffer set L&1, slot r, so that any of them can be the next launch's previous
// row; the fields of mesh rows go to the field arrays (a mesh row computed past the
// point that holds is computed again later, and stored again).
//
// Lane t holds the same node through the launch; the previous row of row 1 is loaded from
// row v, each later one is the lanes' own values of the row before, exchanged through
// LDS. All lanes compute every row; values near the workgroup's edges go stale one lane
// per row (their neighbours are missing), and the owned lanes stay exact through WKMAX
// rows. Only owned lanes store or fold. BM is the breaking model, FRIC the bottom friction
// switch.
template <int BM, bool FRIC>
__global__ void __launch_bounds__(WB) k_wave_block(WaveArgs a, double T, int L, WaveCtl *ctl, unsigned long long *acc)
{
    const int t = threadIdx.x, k = blockIdx.x;
    const int ny = a.ny, nx = a.nx;
    const int j0 = k * WNB;
    const int nint = min(WNB, ny - j0);
    const bool owned = t >= WHALO && t < WHALO + nint;
    int j = (j0 + t - WHALO) % ny;
    if (j < 0) j += ny;
    const int tm = (t == 0) ? 0 : t - 1, tp = (t == WB - 1) ? WB - 1 : t + 1;
    const bool rollers = a.p.rollers != 0;
    const m70w_params &p = a.p;
    const WaveCtl cp = ctl[(L - 1) & 1];
    if (cp.done) {   // pass the final state on to the launches that follow
        if (k == 0 && t == 0) {
            WaveCtl nc = cp;
            nc.K = 0;
            ctl[L & 1] = nc;
        }
        return;
    }
    // the previous launch's rows 2..K under the step rule
    const unsigned long long *accp = acc + ((L - 1) % 3) * WACC_SLOTS;
    WaveCtl c = cp;
    int v = 1;
    while (v < cp.K && !c.done) {
        double dxr;
        int st;
        get_dx(a, c, tan_max_of(accp[v]), dxr, st);
        if (!st) break;   // row v+1 should have been an intermediate row
        c = advance(c, dxr, 1, nx);
        ++v;
    }
    if (c.done) {   // the previous launch's rows reached the last mesh row
        if (k == 0 && t == 0) {
            WaveCtl nc = c;
            nc.K = 0;
            ctl[L & 1] = nc;
        }
        return;
    }
    const bool all_held = (v == cp.K);
    // this launch's row 1
    double dx_row;
    int store;
    get_dx(a, c, tan_max_of(accp[v]), dx_row, store);
    const WaveCtl c1 = advance(c, dx_row, store, nx);
    int K = all_held ? min(2 * cp.K, WKMAX) : 1;
    K = c1.done ? 1 : min(K, 1 + (nx - 1 - c1.xIndex));
    if (k == 0 && t == 0) {
        WaveCtl nc = c1;
        nc.K = K;
        ctl[L & 1] = nc;
        unsigned long long *accn = acc + ((L + 1) % 3) * WACC_SLOTS;   // the set launch L+1 folds into
        for (int r = 0; r < WACC_SLOTS; ++r) accn[r] = 0;
    }
    unsigned long long *const accl = acc + (L % 3) * WACC_SLOTS;
    const long rstride = (long)WROW_ARRAYS * ny;
    double *const rows_out = a.rows[L & 1];

    __shared__ double sp[PL_N][WB];
    __shared__ double sl_lK[WB], sl_U[WB], sl_V[WB], sl_Cgy[WB], sl_Cy[WB];
    __shared__ double s_sigma0;
    const double inv_2dy = a.inv_2dy;
    const double w = 2 * WPI / T;

    // the previous row of row 1: row v of the previous launch
    PrevOwn own;
    {
        const double *rb = a.rows[(L - 1) & 1] + (long)(v - 1) * rstride;
        auto r = [&](int q) { return rb + (long)q * ny; };
        own.lk = r(WR_lK)[j];
        own.u = r(WR_U)[j];
        own.v = r(WR_V)[j];
        own.le = r(WR_lE)[j];
        own.cgx = r(WR_Cgx)[j];
        own.h = r(WR_H)[j];
        sp[PL_TT][t] = r(WR_TanTheta)[j];
        sp[PL_LK][t] = own.lk;
        sp[PL_CG][t] = r(WR_Cg)[j];
        sp[PL_U][t] = own.u;
        sp[PL_V][t] = own.v;
        sp[PL_DW][t] = r(WR_dw)[j];
        sp[PL_DF][t] = FRIC ? r(WR_df)[j] : 0.0;   // df is zero on every row without friction
        sp[PL_LE][t] = own.le;
        sp[PL_WC][t] = r(WR_WC)[j];
        sp[PL_CGY][t] = r(WR_Cgy)[j];
        if (rollers) {
            own.cx = r(WR_Cx)[j];
            own.er = r(WR_E_r)[j];
            sp[PL_C][t] = r(WR_C)[j];
            sp[PL_WCR][t] = r(WR_WC_r)[j];
            sp[PL_CY][t] = r(WR_Cy)[j];
            sp[PL_ER][t] = own.er;
            sp[PL_DR][t] = r(WR_d_r)[j];
            sp[PL_DWE][t] = r(WR_Dw)[j];
        } else {
            own.cx = own.er = 0.0;
        }
    }
    MeshNode next_m{};
    if (store) next_m = load_mesh(a, (long)(c.xIndex + 1) * ny + j);
    __syncthreads();

    for (int row = 1; row <= K; ++row) {
        if (row > 1) {   // taken to be a mesh row
            dx_row = a.x[c.xIndex + 1] - c.x0;
            store = 1;
        }
        const WaveCtl cn = advance(c, dx_row, store, nx);
        const long index = (long)(c.xIndex + 1) * ny, index_m1 = (long)c.xIndex * ny;
        const MeshNode m = next_m;
        if (row < K) next_m = load_mesh(a, (long)(cn.xIndex + 1) * ny + j);   // the next row's, early
        WNode n{};
        // initRow + dispersion
        if (store) {   // a mesh row: solved before the march, from the same inputs
            n.D = m.D;
            n.U = m.U;
            n.V = m.V;
            n.U_w = m.U_w;
            n.lK = m.lK;
            n.Sigma = m.Sigma;
            n.C = m.C;
            n.Cg = m.Cg;
            n.Urms_H = m.Urms_H;
        } else {
            const double dx_dx = dx_row / a.dx[c.xIndex];
            auto interp = [&](const double *f, long q) { return f[index_m1 + q] + (f[index + q] - f[index_m1 + q]) * dx_dx; };
            n.D = interp(a.D, j);
            n.U = interp(a.U, j);
            n.V = interp(a.V, j);
            n.U_w = interp(a.U_w, j);
            w_dispersion(n.D, n.U_w, w, p.z0, n.lK, n.Sigma, n.C, n.Cg, n.Urms_H);
        }
        sl_lK[t] = n.lK;
        sl_U[t] = n.U;
        sl_V[t] = n.V;
        if (FRIC && t == 0 && p.friction_model == M70W_FRIC_SOULSBY97) {
            double sg0;   // Sigma of the row's first node
            if (store) {
                sg0 = a.disp.Sigma[index];
            } else {
                const double dx_dx = dx_row / a.dx[c.xIndex];
                auto interp = [&](const double *f) { return f[index_m1] + (f[index] - f[index_m1]) * dx_dx; };
                double lK0, C0, Cg0, Uh0;
                w_dispersion(interp(a.D), interp(a.U_w), w, p.z0, lK0, sg0, C0, Cg0, Uh0);
            }
            s_sigma0 = sg0;
        }
        __syncthreads();
        PrevRow P;
        P.tt_p = sp[PL_TT][tp];
        P.tt_m = sp[PL_TT][tm];
        P.lk_j = own.lk;
        P.lk_p = sp[PL_LK][tp];
        P.lk_m = sp[PL_LK][tm];
        // solveDir
        const double dx = dx_row;
        const double dx2 = dx * dx;
        const double inv_dx = 1 / dx;
        {
            double tt_p1 = P.tt_p;
            double tt_m1 = P.tt_m;
            double dttheta_dy = (tt_p1 - tt_m1) * inv_2dy;
            double ttheta_med = (tt_p1 + tt_m1) * 0.5;
            double y_i = -ttheta_med / (dttheta_dy + inv_dx);
            double s_i = sqrt(dx2 + y_i * y_i);
            double inv_s_i = 1 / s_i;
            double sinTheta_i = -y_i * inv_s_i;
            double cosTheta_i = dx * inv_s_i;
            double theta_i = asin(sinTheta_i);
            n.Y_i = y_i;
            n.inv_s_i = inv_s_i;
            double dlk_dx = (n.lK - P.lk_j) * inv_dx;
            double dlk_dy_c = 0.5 * inv_2dy * (sl_lK[tp] + P.lk_p - sl_lK[tm] - P.lk_m);
            double dlk_ds = dlk_dx * cosTheta_i + dlk_dy_c * sinTheta_i;
            double dlk_dst = -dlk_dx * sinTheta_i + dlk_dy_c * cosTheta_i;
            double theta = theta_i + dlk_dst / (inv_s_i + dlk_ds * 0.5);
            theta = smin_(theta, p.theta_max);
            theta = smax_(theta, -p.theta_max);
            const double sinTheta = sin(theta);
            const double cosTheta = cos(theta);
            node_dir_terms(n, theta, cosTheta, sinTheta, cosTheta * sinTheta);
            sl_Cgy[t] = n.Cgy;
            sl_Cy[t] = n.Cy;
        }
        __syncthreads();
        {   // the row's largest |tan(theta)| over the owned nodes, folded while the energy runs
            double mv = -1;
            if (owned) {
                const double tt = fabs(n.TanTheta);
                if (tt >= 0) mv = tt;
            }
            fold_tan_max(mv, &accl[row]);
        }
        P.cg_p = sp[PL_CG][tp]; P.cg_m = sp[PL_CG][tm];
        P.u_p = sp[PL_U][tp]; P.u_m = sp[PL_U][tm]; P.u_j = own.u;
        P.v_p = sp[PL_V][tp]; P.v_m = sp[PL_V][tm]; P.v_j = own.v;
        P.dw_p = sp[PL_DW][tp]; P.dw_m = sp[PL_DW][tm];
        P.df_p = sp[PL_DF][tp]; P.df_m = sp[PL_DF][tm];
        P.le_p = sp[PL_LE][tp]; P.le_m = sp[PL_LE][tm]; P.le_j = own.le;
        P.wc_p = sp[PL_WC][tp]; P.wc_m = sp[PL_WC][tm];
        P.cgx_j = own.cgx; P.cgy_p = sp[PL_CGY][tp]; P.cgy_m = sp[PL_CGY][tm];
        P.h_j = own.h;
        if (rollers) {
            P.c_p = sp[PL_C][tp]; P.c_m = sp[PL_C][tm];
            P.wcr_p = sp[PL_WCR][tp]; P.wcr_m = sp[PL_WCR][tm];
            P.cx_j = own.cx; P.cy_p = sp[PL_CY][tp]; P.cy_m = sp[PL_CY][tm];
            P.er_p = sp[PL_ER][tp]; P.er_m = sp[PL_ER][tm]; P.er_j = own.er;
            P.dr_p = sp[PL_DR][tp]; P.dr_m = sp[PL_DR][tm];
            P.Dw_p = sp[PL_DWE][tp]; P.Dw_m = sp[PL_DWE][tm];
        }
        {
            const double y_i_dy = n.Y_i * inv_2dy;
            const double inv_s_i = n.inv_s_i;
            const double U_up = sl_U[tp], U_low = sl_U[tm], V_up = sl_V[tp], V_low = sl_V[tm];
            double G, Gp;
            {   // energyPreprocessor (src/WaveEnergy.cpp:181-219)
                double cg_c = 0.5 * (n.Cg + at_ray(P.cg_p, P.cg_m, y_i_dy));
                double u_c = 0.5 * (n.U + at_ray(P.u_p, P.u_m, y_i_dy));
                double v_c = 0.5 * (n.V + at_ray(P.v_p, P.v_m, y_i_dy));
                double dw_i = at_ray(P.dw_p, P.dw_m, y_i_dy);
                double df_i = at_ray(P.df_p, P.df_m, y_i_dy);
                double le_i = at_ray(P.le_p, P.le_m, y_i_dy);
                double du_dx = (n.U - P.u_j) * inv_dx;
                double du_dy = (U_up - U_low) * inv_2dy;
                double dv_dx = (n.V - P.v_j) * inv_dx;
                double dv_dy = (V_up - V_low) * inv_2dy;
                double wc = 2 * n.sxx * du_dx + n.sxy * (du_dy + dv_dx) + 2 * n.syy * dv_dy;
                double wc_c = 0.5 * (wc + at_ray(P.wc_p, P.wc_m, y_i_dy));
                n.WC = wc;
                n.WC_r = n.sxx_r * du_dx + n.sxy_r * (du_dy + dv_dx) + n.syy_r * dv_dy;
                double dcgxu_dx = (n.Cgx - P.cgx_j) * inv_dx + du_dx;
                double dcgyv_dy_c = 0.5 * (sl_Cgy[tp] + P.cgy_p - sl_Cgy[tm] - P.cgy_m) * inv_2dy + dv_dy;
                double div_cgv_c = dcgxu_dx + dcgyv_dy_c;
                double dle_dy_c = (P.le_p - P.le_m) * inv_2dy;
                G = v_c * dle_dy_c + wc_c + div_cgv_c + 0.5 * (dw_i + df_i) - cg_c * inv_s_i * le_i -
                    u_c * P.le_j * inv_dx;
                Gp = cg_c * inv_s_i + u_c * inv_dx;
            }
            {   // solveEnergy: Newton-Raphson on log(E); the Jacobian's doubled Gp is the
                // reference's (src/WaveEnergy.cpp:97)
                const double logCteEnergy = log(CTE_ENERGY);
                double le_old = P.le_j, H = P.h_j, E = 0.0, Dw = 0.0;
                double dw = 0.0, ddw_dh = 0.0, df = 0.0, ddf_dh = 0.0;
                for (int it = 0; it < p.max_iterations; it++) {
                    w_breaking<BM>(p, H, n.Sigma, n.D, dw, ddw_dh);
                    if (FRIC) w_friction(p, H, s_sigma0, n.Urms_H, df, ddf_dh);
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
                n.lE = le_old;
                n.E = E;
                n.H = H;
                n.dw = dw;
                n.df = df;
                n.Dw = Dw;
            }
            if (rollers) {   // solveRollers (src/WaveRollers.cpp:57-100)
                const double d_r_div_c = 2 * WG * p.sinBeta_r;
                double c_c = 0.5 * (n.C + at_ray(P.c_p, P.c_m, y_i_dy));
                double u_c = 0.5 * (n.U + at_ray(P.u_p, P.u_m, y_i_dy));
                double v_c = 0.5 * (n.V + at_ray(P.v_p, P.v_m, y_i_dy));
                double du_dx = (n.U - P.u_j) * inv_dx;
                double dv_dy = (V_up - V_low) * inv_2dy;
                double wc_r_c = 0.5 * (n.WC_r + at_ray(P.wcr_p, P.wcr_m, y_i_dy));
                double dcxu_dx = (n.Cx - P.cx_j) * inv_dx + du_dx;
                double dcyv_dy_c = 0.5 * (sl_Cy[tp] + P.cy_p - sl_Cy[tm] - P.cy_m) * inv_2dy + dv_dy;
                double div_cv_c = dcxu_dx + dcyv_dy_c;
                double der_dy_c = (P.er_p - P.er_m) * inv_2dy;
                const double d_r = d_r_div_c / n.C;
                n.d_r = d_r;
                double d_r_c = 0.5 * (d_r + at_ray(P.dr_p, P.dr_m, y_i_dy));
                double Dw_c = 0.5 * (n.Dw + at_ray(P.Dw_p, P.Dw_m, y_i_dy));
                double er_i = at_ray(P.er_p, P.er_m, y_i_dy);
                double upv = c_c * inv_s_i * er_i + u_c * P.er_j * inv_dx - v_c * der_dy_c + 0.5 * p.alpha_r * Dw_c;
                double lowv = c_c * inv_s_i + u_c * inv_dx + div_cv_c + 0.5 * wc_r_c + 0.5 * d_r_c;
                n.E_r = upv / lowv;
            }
        }
        if (owned) {
            write_row(rows_out + (long)(row - 1) * rstride, ny, j, n, rollers, FRIC);
            if (store) store_node(a.f, index + j, n, rollers);
        }
        __syncthreads();   // every lane has read the previous row
        own.lk = n.lK;
        own.u = n.U;
        own.v = n.V;
        own.le = n.lE;
        own.cgx = n.Cgx;
        own.h = n.H;
        sp[PL_TT][t] = n.TanTheta;
        sp[PL_LK][t] = n.lK;
        sp[PL_CG][t] = n.Cg;
        sp[PL_U][t] = n.U;
        sp[PL_V][t] = n.V;
        sp[PL_DW][t] = n.dw;
        sp[PL_DF][t] = n.df;
        sp[PL_LE][t] = n.lE;
        sp[PL_WC][t] = n.WC;
        sp[PL_CGY][t] = n.Cgy;
        if (rollers) {
            own.cx = n.Cx;
            own.er = n.E_r;
            sp[PL_C][t] = n.C;
            sp[PL_WCR][t] = n.WC_r;
            sp[PL_CY][t] = n.Cy;
            sp[PL_ER][t] = n.E_r;
            sp[PL_DR][t] = n.d_r;
            sp[PL_DWE][t] = n.Dw;
        }
        __syncthreads();
        c = cn;
    }
}

// The arrays Wave::storageRow (src/Wave.cpp:175-264) derives, for every node of the field:
// the direction's cosine and sine, the mesh node's Cg and log k, the radiation stresses
// from E and E_r, the rollers' dissipation and Urms. The node terms are node_dir_terms',
// on the stored Theta and the mesh node's C and Cg (a mesh row's dispersion is a.disp's),
// so each value is the one storageRow forms from the marching node; d_r is zero on the
// first row and without rollers.
__global__ void __launch_bounds__(256) k_wave_field_terms(WaveArgs a, long n)
{
    const long o = (long)blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    const WField &f = a.f;
    WNode nd{};
    nd.C = a.disp.C[o];
    nd.Cg = a.disp.Cg[o];
    const double theta = f.Theta[o];
    const double cosTheta = cos(theta), sinTheta = sin(theta);
    node_dir_terms(nd, theta, cosTheta, sinTheta, cosTheta * sinTheta);
    double d_r = 0.0;
    if (a.p.rollers && o >= a.ny) {   // solveRollers (src/WaveRollers.cpp:57-100)
        const double d_r_div_c = 2 * WG * a.p.sinBeta_r;
        d_r = d_r_div_c / nd.C;
    }
    const double e = f.E[o], e_r = f.E_r[o];
    const double sxx_w = nd.sxx * e, sxy_w = nd.sxy * e, syy_w = nd.syy * e;
    const double sxx_r = nd.sxx_r * e_r, sxy_r = nd.sxy_r * e_r, syy_r = nd.syy_r * e_r;
    f.CosTheta[o] = cosTheta;
    f.SinTheta[o] = sinTheta;
    f.Cg[o] = nd.Cg;
    f.K[o] = a.disp.lK[o];
    f.Sxx_w[o] = sxx_w;
    f.Sxy_w[o] = sxy_w;
    f.Syy_w[o] = syy_w;
    f.D_r[o] = d_r * e_r;
    f.Sxx_r[o] = sxx_r;
    f.Sxy_r[o] = sxy_r;
    f.Syy_r[o] = syy_r;
    f.Sxx[o] = sxx_w + sxx_r;
    f.Sxy[o] = sxy_w + sxy_r;
    f.Syy[o] = syy_w + syy_r;
    f.Urms[o] = a.disp.Urms_H[o] * f.H[o];
}

__global__ void __launch_bounds__(256) k_wave_dispersion(WaveArgs a, double w, long n)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    w_dispersion(a.D[q], a.U_w[q], w, a.p.z0, a.disp.lK[q], a.disp.Sigma[q], a.disp.C[q], a.disp.Cg[q],
                 a.disp.Urms_H[q]);
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

int wave_blocks(int ny) { return (ny + WNI - 1) / WNI; }
int wave_blocks_march(int ny) { return (ny + WNB - 1) / WNB; }

int wave_march_launches(int nx)
{
    int n = 0, K = 1;
    for (long rows = 0; rows < nx - 1; ++n) {
        K = std::min(2 * K, WKMAX);
        rows += K;
    }
    return n + 1;
}

hipError_t launch_wave_first(hipStream_t st, const WaveArgs &a, double hOff, double thetaOff, double T,
                             const WaveCtlBufs &b)
{
    hipLaunchKernelGGL(k_wave_first, dim3((unsigned)wave_blocks(a.ny)), dim3(WB), 0, st, a, hOff, thetaOff, T, b.ctl,
                       b.acc);
    return hipGetLastError();
}

hipError_t launch_wave_dispersion(hipStream_t st, const WaveArgs &a, double T)
{
    const long n = (long)a.nx * a.ny;
    hipLaunchKernelGGL(k_wave_dispersion, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, 2 * WPI / T, n);
    return hipGetLastError();
}

using WaveStepKernel = void (*)(WaveArgs, double, int, WaveCtl *, unsigned long long *);

template <bool FRIC>
WaveStepKernel wave_step_kernel(int breaking_model)
{
    switch (breaking_model) {
    case M70W_BREAK_CT: return k_wave_block<M70W_BREAK_CT, FRIC>;
    case M70W_BREAK_TG: return k_wave_block<M70W_BREAK_TG, FRIC>;
    default: return k_wave_block<M70W_BREAK_REGULAR, FRIC>;
    }
}

hipError_t launch_wave_steps(hipStream_t st, const WaveArgs &a, double T, int s0, int count, const WaveCtlBufs &b)
{
    const WaveStepKernel kern =
        a.p.bottom_friction ? wave_step_kernel<true>(a.p.breaking_model) : wave_step_kernel<false>(a.p.breaking_model);
    for (int L = s0; L < s0 + count; ++L)
        hipLaunchKernelGGL(kern, dim3((unsigned)wave_blocks_march(a.ny)), dim3(WB), 0, st, a, T, L, b.ctl, b.acc);
    return hipGetLastError();
}

hipError_t launch_wave_field_terms(hipStream_t st, const WaveArgs &a)
{
    const long n = (long)a.nx * a.ny;
    hipLaunchKernelGGL(k_wave_field_terms, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
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
