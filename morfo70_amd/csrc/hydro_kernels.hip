// hydro_kernels.hip -- stencil kernels of the Hydro ADI step for gfx950.
//
// One thread per cell, blocks of 256 threads along j (alongshore, contiguous), one
// grid row per i: every global access of a wavefront is a contiguous 512-byte run.
// Values that the reference computes once in one row and reads from neighbouring rows
// (IsWater, Di/inv_Di, Ui_j, guess_D, U, V) are recomputed per thread from the same
// inputs with the same expression, which reproduces them bit for bit without a grid-wide
// barrier between producer and consumer.
#include "hydro_device.h"
#include "hydro_kernels.h"

namespace m70 {

#define CELL_IJ(nrows)                                   \
    const int j = blockIdx.x * blockDim.x + threadIdx.x; \
    const int i = blockIdx.y;                            \
    if (j >= k.ny || i >= (nrows)) return;               \
    const int ny = k.ny;                                 \
    const long r = (long)i * ny;

// ------------------------------------------------------------------------------------
// Per-Run fields, src/Hydro.cpp:175-192
// ------------------------------------------------------------------------------------

// Hydro::calculateDragAndViscosity with the model functions,
// src/HydroParametrizations.cpp:30-113. Rows 0..nx-1.
__global__ void __launch_bounds__(256) k_drag_viscosity(Consts k, const double *__restrict__ D,
                                                        Forcing fw, RunFields rf)
{
    CELL_IJ(k.nx)
    const long q = r + j;
    const double d = D[q];
    double cd;
    if (k.drag_model == 0) {                       // dragCoefficient_Constant :49-51
        cd = k.cD0;
    } else if (k.drag_model == 1) {                // dragCoefficient_LogDepth :53-63
        double dd = d;
        dd = (dd < k.z0 + 0.001) ? (k.z0 + 0.001) : dd;   // std::max(d, z0+0.001)
        const double c = 0.4 / (log(dd / k.z0) - 1.0);
        cd = c * c;
    } else {                                       // dragCoefficient_ManningStrickler :65-71
        cd = 0.015 * cbrt(k.ka / d);
    }
    double nud;
    if (k.visc_model == 0) {                       // viscosity_Constant :73-77
        nud = d * k.nu0;
    } else if (k.visc_model == 1) {                // viscosity_BreakingRollers :79-90
        nud = d * (k.nu0 + k.M * d * cbrt(((1 - k.alpha_r) * fw.Dw[q] + fw.D_r[q]) * k.inv_rho));
    } else if (k.visc_model == 2) {                // viscosity_BreakingDepth :92-101
        nud = d * (k.nu0 + k.M * d * cbrt(fw.Dw[q] * k.inv_rho));
    } else {                                       // viscosity_BreakingHrms :103-113 (float 1/rho)
        nud = d * (k.nu0 + k.M * fw.H[q] * cbrt(fw.Dw[q] * (double)k.inv_rho_f));
    }
    rf.cD[q] = cd;
    rf.nuD[q] = nud;
}

// Forcing, face averages and viscosity derivatives. Rows 0..nx-1 (x-face fields only
// for rows 0..nx-2).
__global__ void __launch_bounds__(256) k_run_fields(Consts k, Metrics m, Forcing fw, RunFields rf)
{
    CELL_IJ(k.nx)
    const int nx = k.nx;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    const double *nuD = rf.nuD;
    if (i < nx - 1) {
        const double *Sxy = fw.Sxy + r, *Sxy_up = Sxy + ny, *Sxx = fw.Sxx + r, *Sxx_up = Sxx + ny;
        // getForcing_x, src/HydroQx.cpp:371-386 (j = 0 and ny-1 wrap, :378-379)
        rf.Fx[r + j] = -(k.inv_4dy_rho * (Sxy[jp] - Sxy[jm] + Sxy_up[jp] - Sxy_up[jm]) +
                         (Sxx_up[j] - Sxx[j]) * m.inv_rhodx[i]);
        // C2X (include/MeshOperations.h:18-27) of Urms, cD, nuD: src/Hydro.cpp:177,182,186
        rf.Urms_x[r + j] = 0.5 * (fw.Urms[r + j] + fw.Urms[r + ny + j]);
        rf.cD_x[r + j] = 0.5 * (rf.cD[r + j] + rf.cD[r + ny + j]);
        const double nux = 0.5 * (nuD[r + j] + nuD[r + ny + j]);
        rf.nuD_x[r + j] = nux;
        // difC2X(dnuD_dx, nuD, dx), src/HydroMeshOperations.cpp:99-114
        rf.dnuD_dx[r + j] = m.inv_dx[i] * (nuD[r + ny + j] - nuD[r + j]);
        // difC2C_y(dnuDx_dy, nuD_x), src/HydroMeshOperations.cpp:202-221
        const double nux_p = 0.5 * (nuD[r + jp] + nuD[r + ny + jp]);
        const double nux_m = 0.5 * (nuD[r + jm] + nuD[r + ny + jm]);
        rf.dnuDx_dy[r + j] = k.inv_2dy_c2c * (nux_p - nux_m);
    }
    {   // getForcing_y, src/HydroQy.cpp:367-424. Row 0 uses Sxy rows 0,1,2 with the row-0
        // coefficients (:385-393); the last row has no "up" term (:413-421).
        const double d_dx_d = m.difx_c_d[i] * k.inv_rho;
        const double d_dx_low = m.difx_c_low[i] * k.inv_rho;
        const double *Syy = fw.Syy + r;
        double val;
        if (i < nx - 1) {
            const double d_dx_up = m.difx_c_up[i] * k.inv_rho;
            const long base = (i == 0) ? 0 : (long)(i - 1) * ny;
            const double *Sxy_low = fw.Sxy + base, *Sxy = Sxy_low + ny, *Sxy_up = Sxy + ny;
            val = -(k.inv_dy_rho * (Syy[jp] - Syy[j]) +
                    0.5 * ((Sxy_up[j] + Sxy_up[jp]) * d_dx_up + (Sxy[j] + Sxy[jp]) * d_dx_d +
                           (Sxy_low[j] + Sxy_low[jp]) * d_dx_low));
        } else {
            const double *Sxy_low = fw.Sxy + (long)(i - 1) * ny, *Sxy = Sxy_low + ny;
            val = -(k.inv_dy_rho * (Syy[jp] - Syy[j]) +
                    0.5 * ((Sxy[j] + Sxy[jp]) * d_dx_d + (Sxy_low[j] + Sxy_low[jp]) * d_dx_low));
        }
        rf.Fy[r + j] = val;
    }
    // C2Y (src/HydroMeshOperations.cpp:47-61) of Urms, cD, nuD: src/Hydro.cpp:180,184,187
    rf.Urms_y[r + j] = 0.5 * (fw.Urms[r + jp] + fw.Urms[r + j]);
    rf.cD_y[r + j] = 0.5 * (rf.cD[r + jp] + rf.cD[r + j]);
    const double nuy = 0.5 * (nuD[r + jp] + nuD[r + j]);
    rf.nuD_y[r + j] = nuy;
    // difC2Y(dnuD_dy, nuD), src/HydroMeshOperations.cpp:116-130
    rf.dnuD_dy[r + j] = (nuD[r + jp] - nuD[r + j]) * k.inv_dy;
    // difC2C_x_addBoundary(dnuDy_dx, nuD_y), src/HydroMeshOperations.cpp:161-200
    {
        double out;
        if (i == 0) {
            const double vu = 0.5 * (nuD[r + ny + jp] + nuD[r + ny + j]);
            out = m.difx_c_low[0] * (2 * nuy - vu) + m.difx_c_d[0] * nuy + vu * m.difx_c_up[0];
        } else if (i < nx - 1) {
            const double vl = 0.5 * (nuD[r - ny + jp] + nuD[r - ny + j]);
            const double vu = 0.5 * (nuD[r + ny + jp] + nuD[r + ny + j]);
            out = m.difx_c_low[i] * vl + m.difx_c_d[i] * nuy + vu * m.difx_c_up[i];
        } else {
            const double vl = 0.5 * (nuD[r - ny + jp] + nuD[r - ny + j]);
            out = m.difx_c_low[nx - 1] * vl + m.difx_c_d[nx - 1] * nuy;
        }
        rf.dnuDy_dx[r + j] = out;
    }
}

// ------------------------------------------------------------------------------------
// Qx predictor assembly: src/HydroQx.cpp:23-48 with the coefficient functions of
// :99-256 and setSolverQxx_BC (src/HydroBoundaryConditions.cpp:25-39). Rows 0..nx-2.
// Writes a (d), b (low), c (up), f plus IsWater and DragTerm for the projector.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_qxx_assemble(Consts k, Metrics m, State s, RunFields rf,
                                                      Work w, double f0)
{
    CELL_IJ(k.nx - 1)
    const int nx = k.nx;
    const int jm = wrap_dec(j, ny);
    const double dt = k.dt;
    const double *D = s.D, *Zs = s.Zs, *dQx = s.dQx;
    const double Di = D[r + j], Di1 = D[r + ny + j];
    // getD_x(i): Di row i+1 (:403-418)
    const double Dx1 = 0.5 * (Di + Di1);
    const double invDx1 = 1 / Dx1;
    // setWaterNodes_x(i) (:231-256)
    const int wi = wet_rule(Di, Di1, Zs[r + j], Zs[r + ny + j], dQx[r + j], dQx[r + ny + j], k.Dmask);
    // Ui_j = C2X(V_c) (:23, src/HydroMeshOperations.cpp:15-28) and drag (:32,42)
    const double vx = 0.5 * (s.Vc[r + j] + s.Vc[r + ny + j]);
    const double u1 = s.U[r + ny + j];
    const double drag = drag_term(rf.cD_x[r + j], rf.Urms_x[r + j], u1, vx, invDx1, k);
    // setSolverQxx_f (:99-133)
    double fv;
    {
        const double inv_dx = m.inv_dx[i];
        const double *zs = Zs + r, *dqy_dy = s.dQy + r, *v = s.V + r + ny;
        const double dzs_dx = inv_dx * (zs[ny + j] - zs[j]);
        const double d2qy_dxy = inv_dx * (dqy_dy[ny + j] - dqy_dy[j]);
        const double dV_dx = 0.5 * inv_dx * ((v[ny + j] + v[ny + jm]) - (v[j] + v[jm]));
        const double d2V_dxy = inv_dx * k.inv_dy * ((v[ny + j] - v[ny + jm]) - (v[j] - v[jm]));
        fv = wi * (rf.Fx[r + j] + s.Qx[r + ny + j] * k.inv_dt + G * Dx1 * (dt * d2qy_dxy - dzs_dx) +
                   rf.dnuDx_dy[r + j] * dV_dx + rf.nuD_x[r + j] * d2V_dxy);
    }
    const double dnud_dx = rf.dnuD_dx[r + j], nud_x = rf.nuD_x[r + j];
    // setSolverQxx_d (:135-162)
    const int isl = !wi;
    double dv = (k.inv_dt + drag + m.difx_x_d[i] * (u1 - 2 * dnud_dx * invDx1) -
                 m.dif2x_x_d[i] * (G * Dx1 * dt + 2 * nud_x * invDx1)) * wi + isl;
    // setSolverQxx0_low (:214-229) / setSolverQxx_low (:190-212)
    double lv;
    if (i == 0) {
        const double Dx0 = 1.5 * Di - 0.5 * Di1;     // getD_x0 (:390-401)
        const double invDx0 = 1.0 / Dx0;
        lv = (m.difx_x_low[0] * (s.U[j] - 2 * dnud_dx * invDx0) -
              m.dif2x_x_low[0] * (G * Dx1 * dt + 2 * nud_x * invDx0));
    } else {
        const double Dm = D[r - ny + j];
        const double Dxi = 0.5 * (Dm + Di);
        const double invDxi = 1 / Dxi;
        const int wm = wet_rule(Dm, Di, Zs[r - ny + j], Zs[r + j], dQx[r - ny + j], dQx[r + j], k.Dmask);
        const double ww = wi && wm;
        lv = (m.difx_x_low[i] * (s.U[r + j] - 2 * dnud_dx * invDxi) -
              m.dif2x_x_low[i] * (G * Dx1 * dt + 2 * nud_x * invDxi)) * ww;
    }
    // setSolverQxx_up (:164-188). Row nx-2's up is never read by the Thomas solve
    // (src/Math.cpp:38-44 uses c rows 0..n-2), so it is left zero.
    double uv = 0.0;
    if (i < nx - 2) {
        const double Dp = D[r + 2L * ny + j];
        const double Dx2 = 0.5 * (Di1 + Dp);
        const double invDx2 = 1 / Dx2;
        const int wp = wet_rule(Di1, Dp, Zs[r + ny + j], Zs[r + 2L * ny + j], dQx[r + ny + j],
                                dQx[r + 2L * ny + j], k.Dmask);
        const double ww = wi && wp;
        uv = (m.difx_x_up[i] * (s.U[r + 2L * ny + j] - 2 * dnud_dx * invDx2) -
              m.dif2x_x_up[i] * (G * Dx1 * dt + 2 * nud_x * invDx2)) * ww;
    }
    if (i == 0) {   // setSolverQxx_BC
        dv += k.f1 * lv;
        fv -= f0 * lv;
    }
    w.a[r + j] = dv;
    w.b[r + j] = lv;
    w.c[r + j] = uv;
    w.f[r + j] = fv;
    w.w[r + j] = (uint8_t)wi;
    w.drag[r + j] = drag;
}

// ------------------------------------------------------------------------------------
// Qx projector assembly: setSolverQxy_f/_d/_up/_low, src/HydroQx.cpp:262-349. Rows
// 0..nx-2. Reads the predictor's Qx* (Qx rows 1..nx-1), IsWater and DragTerm.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_qxy_assemble(Consts k, State s, RunFields rf, Work w)
{
    CELL_IJ(k.nx - 1)
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    const double dt = k.dt;
    const double *D = s.D;
    const int wj = w.w[r + j], wp = w.w[r + jp], wm = w.w[r + jm];
    const double drag = w.drag[r + j];
    const double nud_x = rf.nuD_x[r + j], dnudx_dy = rf.dnuDx_dy[r + j];
    const double inv_j = 1 / (0.5 * (D[r + j] + D[r + ny + j]));
    const double inv_p = 1 / (0.5 * (D[r + jp] + D[r + ny + jp]));
    const double inv_m = 1 / (0.5 * (D[r + jm] + D[r + ny + jm]));
    const double v_p = 0.5 * (s.Vc[r + jp] + s.Vc[r + ny + jp]);
    const double v_m = 0.5 * (s.Vc[r + jm] + s.Vc[r + ny + jm]);
    const double qx = s.Qx[r + ny + j];
    w.f[r + j] = wj * qx * (1 + dt * drag);
    const int isl = !wj;
    w.a[r + j] = wj * ((1 + dt * drag) + dt * inv_j * nud_x * k.two_inv_dy2) + isl;
    const double wwp = wj && wp;
    w.c[r + j] = dt * (v_p * k.inv_2dy - inv_p * (nud_x * k.inv_dy2_xy + dnudx_dy * k.inv_2dy)) * wwp;
    const double wwm = wj && wm;
    w.b[r + j] = dt * (-v_m * k.inv_2dy - inv_m * (nud_x * k.inv_dy2_xy - dnudx_dy * k.inv_2dy)) * wwm;
}

// ------------------------------------------------------------------------------------
// After the Qx projector: setQx_BC (src/HydroBoundaryConditions.cpp:13-23), difC2X of Qx
// (src/HydroQx.cpp:80), solveZs_guess (src/Hydro.cpp:230-250), getU0/getU
// (src/HydroQx.cpp:420-451), X2C (include/MeshOperations.h:36-47). Rows 0..nx-1.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double qx_at(const State &s, const Consts &k, long rowoff, int row, int j,
                                        double f0)
{
    return row == 0 ? k.f1 * s.Qx[k.ny + j] + f0 : s.Qx[rowoff + j];
}

__global__ void __launch_bounds__(256) k_qx_tail(Consts k, Metrics m, State s, Work w, double f0)
{
    CELL_IJ(k.nx)
    const int nx = k.nx;
    const double dt = k.dt;
    // Qx rows i-1..i+2 (row 0 through the BC, row nx is the wall row held in memory)
    double q[4];
    for (int t = 0; t < 4; ++t) {
        const int row = i - 1 + t;
        q[t] = (row >= 0 && row <= nx) ? qx_at(s, k, (long)row * ny, row, j, f0) : 0.0;
    }
    // dQx_dx and guess_D for rows i-1, i, i+1
    double gd[3];
    double dq_i = 0.0, inc_i = 0.0;
    for (int t = 0; t < 3; ++t) {
        const int row = i - 1 + t;
        if (row < 0 || row > nx - 1) { gd[t] = 0.0; continue; }
        const double dq = m.inv_dxx[row] * (q[t + 1] - q[t]);
        const double incrZs = -dt * (dq + s.dQy[(long)row * ny + j]);
        gd[t] = s.D[(long)row * ny + j] + incrZs;
        if (t == 1) { dq_i = dq; inc_i = incrZs; }
    }
    // U rows i and i+1
    double ui;
    if (i == 0) {
        const double d_x0 = 1.5 * gd[1] - 0.5 * gd[2];
        ui = q[1] / d_x0;
    } else {
        const double d_x = 0.5 * (gd[0] + gd[1]);
        ui = q[1] / d_x;
    }
    double ui1;
    if (i + 1 <= nx - 1) {
        const double d_x = 0.5 * (gd[1] + gd[2]);
        ui1 = q[2] / d_x;
    } else {
        ui1 = s.U[(long)nx * ny + j];   // U row nx is never written by the reference
    }
    s.dQx[r + j] = dq_i;
    w.gZs[r + j] = s.Zs[r + j] + inc_i;
    w.gD[r + j] = gd[1];
    s.U[r + j] = ui;
    s.Uc[r + j] = m.x2c_d[i] * ui + m.x2c_up[i] * ui1;
    if (i == 0) s.Qx[j] = q[1];
}

// ------------------------------------------------------------------------------------
// Qy predictor assembly: src/HydroQy.cpp:26-41 with setWaterNodes_y (:92-122), getD_y
// (:442-456), setSolverQyy_f/_d/_up/_low (:124-230), C2Y of U_c (:34) and the drag
// (:35). Rows 0..nx-2. f and d use IsWater row 0 (src/HydroQy.cpp:140,170).
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_qyy_assemble(Consts k, Metrics m, State s, RunFields rf,
                                                      Work w)
{
    CELL_IJ(k.nx - 1)
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny), jpp = wrap_inc(jp, ny);
    const double dt = k.dt;
    const double *gD = w.gD + r, *gZs = w.gZs + r, *dqy = s.dQy + r;
    // masks of row i at faces jm, j, jp
    const double g_m = gD[jm], g_j = gD[j], g_p = gD[jp], g_pp = gD[jpp];
    const int wy_j = wet_rule(g_j, g_p, gZs[j], gZs[jp], dqy[j], dqy[jp], k.Dmask);
    const int wy_m = wet_rule(g_m, g_j, gZs[jm], gZs[j], dqy[jm], dqy[j], k.Dmask);
    const int wy_p = wet_rule(g_p, g_pp, gZs[jp], gZs[jpp], dqy[jp], dqy[jpp], k.Dmask);
    // mask of row 0 at face j (Q1)
    const int w0 = wet_rule(w.gD[j], w.gD[jp], w.gZs[j], w.gZs[jp], s.dQy[j], s.dQy[jp], k.Dmask);
    // Di row i+1 (getD_y(i)) at j, jm, jp
    const double Dy = 0.5 * (g_p + g_j);
    const double invDy = 1 / Dy;
    const double invDy_m = 1 / (0.5 * (g_j + g_m));
    const double invDy_p = 1 / (0.5 * (g_pp + g_p));
    const double nud_y = rf.nuD_y[r + j];
    double fv;
    {   // setSolverQyy_f (:124-158)
        const double inv_dx = m.inv_dxx[i];
        const double x2c_d = m.x2c_d[i], x2c_up = m.x2c_up[i];
        const double *zs = s.Zs + r, *dqx_dx = s.dQx + r, *u = s.U + r;
        const double dzs_dy = k.inv_dy * (zs[jp] - zs[j]);
        const double d2qx_dxy = k.inv_dy * (dqx_dx[jp] - dqx_dx[j]);
        const double du_dy = k.inv_dy * (x2c_up * (u[ny + jp] - u[ny + j]) + x2c_d * (u[jp] - u[j]));
        const double d2u_dxy = inv_dx * k.inv_dy * ((u[ny + jp] - u[ny + j]) - (u[jp] - u[j]));
        fv = w0 * (rf.Fy[r + j] + s.Qy[r + ny + j] * k.inv_dt + G * Dy * (dt * d2qx_dxy - dzs_dy) +
                   rf.dnuDy_dx[r + j] * du_dy + nud_y * d2u_dxy);
    }
    // Ui_j = C2Y(U_c row i) and the drag with V row i+1 (old)
    const double uy = 0.5 * (s.Uc[r + jp] + s.Uc[r + j]);
    const double *v = s.V + r + ny;
    const double drag = drag_term(rf.cD_y[r + j], rf.Urms_y[r + j], uy, v[j], invDy, k);
    // setSolverQyy_d (:160-181)
    const int isl = !w0;
    const double dv = (k.inv_dt + drag + G * Dy * dt * k.two_inv_dy2 + 2 * nud_y * invDy * k.two_inv_dy2) * w0 + isl;
    const double dnud_dy = rf.dnuD_dy[r + j];
    // setSolverQyy_low (:206-230)
    const double wwm = wy_j && wy_m;
    const double lv = (0.5 * k.inv_dy * (-v[jm] + 2 * dnud_dy * invDy_m) -
                       k.inv_dy2_yy * (G * Dy * dt + 2 * nud_y * invDy_m)) * wwm;
    // setSolverQyy_up (:183-204)
    const double wwp = wy_j && wy_p;
    const double uv = (0.5 * k.inv_dy * (v[jp] - 2 * dnud_dy * invDy_p) -
                       k.inv_dy2_yy * (G * Dy * dt + 2 * nud_y * invDy_p)) * wwp;
    w.a[r + j] = dv;
    w.b[r + j] = lv;
    w.c[r + j] = uv;
    w.f[r + j] = fv;
    w.w[r + j] = (uint8_t)wy_j;
    w.drag[r + j] = drag;
}

// ------------------------------------------------------------------------------------
// Qy projector assembly: setSolverQyx_f/_d/_up/_low/0_low (src/HydroQy.cpp:235-362) and
// setSolverQyx_BC (src/HydroBoundaryConditions.cpp:41-60). Rows 0..nx-2.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_qyx_assemble(Consts k, Metrics m, State s, RunFields rf,
                                                      Work w)
{
    CELL_IJ(k.nx - 1)
    const int nx = k.nx;
    const int jp = wrap_inc(j, ny);
    const double dt = k.dt;
    const double *gD = w.gD;
    const int wi = w.w[r + j];
    const double drag = w.drag[r + j];
    const double nud = rf.nuD_y[r + j], dnud_dx = rf.dnuDy_dx[r + j];
    // setSolverQyx_f (:235-248)
    const double fv = wi * s.Qy[r + ny + j] * (1 + dt * drag);
    // setSolverQyx_d (:250-285)
    double dv;
    {
        const double x2c_d = m.x2c_d[i], x2c_up = m.x2c_up[i];
        const double *u = s.U + r;
        const double dx_d = m.difx_c_d[i], d2x_d = m.dif2x_c_d[i];
        const double inv_dx = m.inv_dxx[i];
        const double inv_d = 1 / (0.5 * (gD[r + jp] + gD[r + j]));   // inv_Di row i+1
        const int isl = !wi;
        const double du_dx = 0.5 * inv_dx * ((u[ny + jp] + u[ny + j]) - (u[jp] + u[j]));
        const double u_y = 0.5 * (x2c_up * (u[ny + jp] + u[ny + j]) + x2c_d * (u[jp] + u[j]));
        dv = wi * ((1 + dt * drag) + dt * (u_y * dx_d + du_dx) - dt * (inv_d * nud * d2x_d + inv_d * dnud_dx * dx_d)) + isl;
    }
    const double uy = 0.5 * (s.Uc[r + jp] + s.Uc[r + j]);   // Ui_j row i (C2Y of U_c)
    // setSolverQyx_up (:287-311); inv_Di row i+2 = getD_y(i+1); IsWater row nx-1 is false
    double uv;
    {
        const double inv_up = 1 / (0.5 * (gD[r + ny + jp] + gD[r + ny + j]));
        const int wu = (i + 1 <= nx - 2) ? w.w[r + ny + j] : 0;
        const double ww = wi && wu;
        const double dx_up = m.difx_c_up[i], d2x_up = m.dif2x_c_up[i];
        uv = ww * dt * (uy * dx_up - nud * d2x_up * inv_up - dnud_dx * dx_up * inv_up);
    }
    // setSolverQyx0_low (:340-362) / setSolverQyx_low (:313-338)
    double lv;
    {
        const double dx_low = m.difx_c_low[i], d2x_low = m.dif2x_c_low[i];
        if (i == 0) {
            // getD_y0 (:426-440); the wrap element reads d[0] where the interior reads
            // row 1 (:438)
            double Dy0;
            if (j < ny - 1)
                Dy0 = 1 * (gD[j + 1] + gD[j]) - 0.5 * (gD[ny + j + 1] + gD[ny + j]);
            else
                Dy0 = 1 * (gD[0] + gD[ny - 1]) - 0.5 * (gD[0] + gD[2 * ny - 1]);
            const double inv_low = 1 / Dy0;
            lv = wi * dt * (uy * dx_low - nud * d2x_low * inv_low - dnud_dx * dx_low * inv_low);
        } else {
            const double inv_low = 1 / (0.5 * (gD[r - ny + jp] + gD[r - ny + j]));
            const int wl = w.w[r - ny + j];
            const double ww = wi && wl;
            lv = ww * dt * (uy * dx_low - nud * d2x_low * inv_low - dnud_dx * dx_low * inv_low);
        }
    }
    if (i == 0) dv += k.bcQy * lv;
    if (i == nx - 2) dv += uv * k.inv_3;
    w.a[r + j] = dv;
    w.b[r + j] = lv;
    w.c[r + j] = uv;
    w.f[r + j] = fv;
}

// ------------------------------------------------------------------------------------
// After the Qy projector: setQy_BC (src/HydroBoundaryConditions.cpp:62-79), difY2C
// (src/HydroMeshOperations.cpp:148-159), solveZs (src/Hydro.cpp:253-271), getV/getV0
// (src/HydroQy.cpp:458-487), Y2C (src/HydroMeshOperations.cpp:88-97). Rows 0..nx-1.
// Reads D, writes the new depth to Dn.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double qy_row(const State &s, const Consts &k, int row, int j)
{   // Qy row `row` in 1..nx after the BC: row nx is Qy[nx-1]*inv_3
    return row < k.nx ? s.Qy[(long)row * k.ny + j] : s.Qy[(long)(k.nx - 1) * k.ny + j] * k.inv_3;
}

__device__ __forceinline__ double dqy_at(const State &s, const Consts &k, int i, int j)
{
    const int jm = wrap_dec(j, k.ny);
    return (qy_row(s, k, i + 1, j) - qy_row(s, k, i + 1, jm)) * k.inv_dy;
}

__device__ __forceinline__ double new_depth(const State &s, const Consts &k, int i, int j)
{
    const long q = (long)i * k.ny + j;
    const double incrZs = -k.dt * (s.dQx[q] + dqy_at(s, k, i, j));
    return s.D[q] + incrZs;
}

__global__ void __launch_bounds__(256) k_qy_tail(Consts k, State s)
{
    CELL_IJ(k.nx)
    const int nx = k.nx;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    const double dq = dqy_at(s, k, i, j);
    const double incrZs = -k.dt * (s.dQx[r + j] + dq);
    const double dn = s.D[r + j] + incrZs;
    const double dn_p = new_depth(s, k, i, jp);
    const double dn_m = new_depth(s, k, i, jm);
    const double qy_j = qy_row(s, k, i + 1, j), qy_m = qy_row(s, k, i + 1, jm);
    const double v_j = qy_j / (0.5 * (dn + dn_p));
    const double v_m = qy_m / (0.5 * (dn_m + dn));
    s.dQy[r + j] = dq;
    s.Zs[r + j] = s.Zs[r + j] + incrZs;
    s.Dn[r + j] = dn;
    s.V[r + ny + j] = v_j;
    s.Vc[r + j] = 0.5 * (v_j + v_m);
    if (i == 0) {
        const double qy0 = k.bcQy * s.Qy[ny + j];
        const double d_y0 = 2 * dn - new_depth(s, k, 1, j);
        s.Qy[j] = qy0;
        s.V[j] = qy0 / d_y0;
    }
    if (i == nx - 1) s.Qy[(long)nx * ny + j] = s.Qy[(long)(nx - 1) * ny + j] * k.inv_3;
}

// ------------------------------------------------------------------------------------
// Non-finite scan of a field; sets *flag to 1 if any element is NaN or Inf.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_nonfinite(const double *__restrict__ x, long n, int *flag)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    int bad = 0;
    for (; idx < n; idx += stride) {
        const double v = x[idx];
        bad |= !(v - v == 0.0);
    }
    if (bad) atomicOr(flag, 1);
}

// Zs = D - Z + h, src/Hydro.cpp:35-37.
__global__ void __launch_bounds__(256) k_init_zs(double *Zs, const double *D, const double *Z, const double *hb,
                                                 long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) Zs[q] = D[q] - Z[q] + hb[q];
}

// ------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------
static inline dim3 cell_grid(int ny, int rows) { return dim3((unsigned)((ny + 255) / 256), (unsigned)rows); }

hipError_t launch_drag_viscosity(hipStream_t st, const Consts &k, const double *D, const Forcing &fw,
                                 const RunFields &rf)
{
    hipLaunchKernelGGL(k_drag_viscosity, cell_grid(k.ny, k.nx), dim3(256), 0, st, k, D, fw, rf);
    return hipGetLastError();
}
hipError_t launch_run_fields(hipStream_t st, const Consts &k, const Metrics &m, const Forcing &fw,
                             const RunFields &rf)
{
    hipLaunchKernelGGL(k_run_fields, cell_grid(k.ny, k.nx), dim3(256), 0, st, k, m, fw, rf);
    return hipGetLastError();
}
hipError_t launch_qxx_assemble(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                               const RunFields &rf, const Work &w, double f0)
{
    hipLaunchKernelGGL(k_qxx_assemble, cell_grid(k.ny, k.nx - 1), dim3(256), 0, st, k, m, s, rf, w, f0);
    return hipGetLastError();
}
hipError_t launch_qxy_assemble(hipStream_t st, const Consts &k, const State &s, const RunFields &rf,
                               const Work &w)
{
    hipLaunchKernelGGL(k_qxy_assemble, cell_grid(k.ny, k.nx - 1), dim3(256), 0, st, k, s, rf, w);
    return hipGetLastError();
}
hipError_t launch_qx_tail(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const Work &w,
                          double f0)
{
    hipLaunchKernelGGL(k_qx_tail, cell_grid(k.ny, k.nx), dim3(256), 0, st, k, m, s, w, f0);
    return hipGetLastError();
}
hipError_t launch_qyy_assemble(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                               const RunFields &rf, const Work &w)
{
    hipLaunchKernelGGL(k_qyy_assemble, cell_grid(k.ny, k.nx - 1), dim3(256), 0, st, k, m, s, rf, w);
    return hipGetLastError();
}
hipError_t launch_qyx_assemble(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                               const RunFields &rf, const Work &w)
{
    hipLaunchKernelGGL(k_qyx_assemble, cell_grid(k.ny, k.nx - 1), dim3(256), 0, st, k, m, s, rf, w);
    return hipGetLastError();
}
hipError_t launch_qy_tail(hipStream_t st, const Consts &k, const State &s)
{
    hipLaunchKernelGGL(k_qy_tail, cell_grid(k.ny, k.nx), dim3(256), 0, st, k, s);
    return hipGetLastError();
}
hipError_t launch_nonfinite(hipStream_t st, const double *x, long n, int *flag)
{
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_nonfinite, dim3((unsigned)blocks), dim3(256), 0, st, x, n, flag);
    return hipGetLastError();
}

hipError_t launch_init_zs(hipStream_t st, double *Zs, const double *D, const double *Z, const double *hb, long n)
{
    hipLaunchKernelGGL(k_init_zs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Zs, D, Z, hb, n);
    return hipGetLastError();
}

}  // namespace m70
