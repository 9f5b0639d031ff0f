// assembly.h -- per-cell coefficient assembly of the four ADI sweeps (device code).
//
// Each function computes row i, column j of one sweep's tridiagonal system (a = diagonal,
// b = lower, c = upper, f = right-hand side) plus the side outputs the next sweep reads.
// The standalone assembly kernels (hydro_kernels.hip) and the fused partition solvers
// (partition_solver.hip) both call them, so both paths see identical coefficients.
// Values the reference keeps in per-row scratch (IsWater, Di/inv_Di, Ui_j) are recomputed
// from the same inputs with the same expressions, which reproduces them exactly.
#pragma once
#include "hydro_device.h"

namespace m70 {

// Qx predictor, src/HydroQx.cpp:23-48 and :99-256 with setSolverQxx_BC
// (src/HydroBoundaryConditions.cpp:25-39), in two parts: the values element (i, j) reads
// (QxxIn, qxx_load) and the arithmetic on them (asm_qxx_vals). The column solver keeps a
// window of rows in registers and fills QxxIn without reloading shared rows; every caller
// runs the same arithmetic.
struct QxxIn {
    double Dm, D0, D1, D2;       // D rows i-1..i+2 (Dm only for i > 0, D2 only for i < nx-2)
    double Zm, Z0, Z1, Z2;       // Zs rows i-1..i+2 (same rows as D)
    double Qm, Q0, Q1, Q2;       // dQx_dx rows i-1..i+2 (same rows as D)
    double Vc0, Vc1;             // V_c rows i, i+1
    double U0, U1, U2;           // U rows i, i+1, i+2 (U2 only for i < nx-2)
    double dqy0, dqy1;           // dQy_dy rows i, i+1
    double v1j, v1m, v2j, v2m;   // V rows i+1, i+2 at columns j and j-1 (periodic)
    double Qx1;                  // Qx row i+1
    double cD, Urms, Fx, dnuDx_dy, nuD_x, dnuD_dx;   // run fields, row i
};

__host__ __device__ __forceinline__ QxxIn qxx_load(const Consts &k, const State &s, const RunFields &rf, int i, int j)
{
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;
    const int jm = wrap_dec(j, ny);
    QxxIn q{};
    q.D0 = s.D[r + j]; q.D1 = s.D[r + ny + j];
    q.Z0 = s.Zs[r + j]; q.Z1 = s.Zs[r + ny + j];
    q.Q0 = s.dQx[r + j]; q.Q1 = s.dQx[r + ny + j];
    if (i > 0) {
        q.Dm = s.D[r - ny + j]; q.Zm = s.Zs[r - ny + j]; q.Qm = s.dQx[r - ny + j];
    }
    if (i < nx - 2) {
        q.D2 = s.D[r + 2L * ny + j]; q.Z2 = s.Zs[r + 2L * ny + j]; q.Q2 = s.dQx[r + 2L * ny + j];
        q.U2 = s.U[r + 2L * ny + j];
    }
    q.Vc0 = s.Vc[r + j]; q.Vc1 = s.Vc[r + ny + j];
    q.U0 = s.U[r + j]; q.U1 = s.U[r + ny + j];
    q.dqy0 = s.dQy[r + j]; q.dqy1 = s.dQy[r + ny + j];
    q.v1j = s.V[r + ny + j]; q.v1m = s.V[r + ny + jm];
    q.v2j = s.V[r + 2L * ny + j]; q.v2m = s.V[r + 2L * ny + jm];
    q.Qx1 = s.Qx[r + ny + j];
    q.cD = rf.cD_x[r + j]; q.Urms = rf.Urms_x[r + j]; q.Fx = rf.Fx[r + j];
    q.dnuDx_dy = rf.dnuDx_dy[r + j]; q.nuD_x = rf.nuD_x[r + j]; q.dnuD_dx = rf.dnuD_dx[r + j];
    return q;
}

// Side outputs: IsWater (wi) and DragTerm (drag) of row i.
__host__ __device__ __forceinline__ void asm_qxx_vals(const Consts &k, const Metrics &m, const QxxIn &q, double f0,
                                                      int i, double &a_out, double &b_out, double &c_out,
                                                      double &f_out, int &wi_out, double &drag_out)
{
    const int nx = k.nx;
    const double dt = k.dt;
    const double Di = q.D0, Di1 = q.D1;
    // getD_x(i): Di row i+1 (:403-418)
    const double Dx1 = 0.5 * (Di + Di1);
    const double invDx1 = 1 / Dx1;
    // setWaterNodes_x(i) (:231-256)
    const int wi = wet_rule(Di, Di1, q.Z0, q.Z1, q.Q0, q.Q1, k.Dmask);
    // Ui_j = C2X(V_c) (:23, src/HydroMeshOperations.cpp:15-28) and drag (:32,42)
    const double vx = 0.5 * (q.Vc0 + q.Vc1);
    const double u1 = q.U1;
    const double drag = drag_term(q.cD, q.Urms, u1, vx, invDx1, k);
    // setSolverQxx_f (:99-133)
    double fv;
    {
        const double inv_dx = m.inv_dx[i];
        const double dzs_dx = inv_dx * (q.Z1 - q.Z0);
        const double d2qy_dxy = inv_dx * (q.dqy1 - q.dqy0);
        const double dV_dx = 0.5 * inv_dx * ((q.v2j + q.v2m) - (q.v1j + q.v1m));
        const double d2V_dxy = inv_dx * k.inv_dy * ((q.v2j - q.v2m) - (q.v1j - q.v1m));
        fv = wi * (q.Fx + q.Qx1 * k.inv_dt + G * Dx1 * (dt * d2qy_dxy - dzs_dx) + q.dnuDx_dy * dV_dx +
                   q.nuD_x * d2V_dxy);
    }
    const double dnud_dx = q.dnuD_dx, nud_x = q.nuD_x;
    // setSolverQxx_d (:135-162)
    const int isl = !wi;
    double dv = (k.inv_dt + drag + m.difx_x_d[i] * (u1 - 2 * dnud_dx * invDx1) -
                 m.dif2x_x_d[i] * (G * Dx1 * dt + 2 * nud_x * invDx1)) * wi + isl;
    // setSolverQxx0_low (:214-229) / setSolverQxx_low (:190-212)
    double lv;
    if (i == 0) {
        const double Dx0 = 1.5 * Di - 0.5 * Di1;     // getD_x0 (:390-401)
        const double invDx0 = 1.0 / Dx0;
        lv = (m.difx_x_low[0] * (q.U0 - 2 * dnud_dx * invDx0) -
              m.dif2x_x_low[0] * (G * Dx1 * dt + 2 * nud_x * invDx0));
    } else {
        const double Dm = q.Dm;
        const double Dxi = 0.5 * (Dm + Di);
        const double invDxi = 1 / Dxi;
        const int wm = wet_rule(Dm, Di, q.Zm, q.Z0, q.Qm, q.Q0, k.Dmask);
        const double ww = wi && wm;
        lv = (m.difx_x_low[i] * (q.U0 - 2 * dnud_dx * invDxi) -
              m.dif2x_x_low[i] * (G * Dx1 * dt + 2 * nud_x * invDxi)) * ww;
    }
    // setSolverQxx_up (:164-188). Row nx-2's up is never read by the Thomas solve
    // (src/Math.cpp:38-44 uses c rows 0..n-2), so it is left zero.
    double uv = 0.0;
    if (i < nx - 2) {
        const double Dp = q.D2;
        const double Dx2 = 0.5 * (Di1 + Dp);
        const double invDx2 = 1 / Dx2;
        const int wp = wet_rule(Di1, Dp, q.Z1, q.Z2, q.Q1, q.Q2, k.Dmask);
        const double ww = wi && wp;
        uv = (m.difx_x_up[i] * (q.U2 - 2 * dnud_dx * invDx2) -
              m.dif2x_x_up[i] * (G * Dx1 * dt + 2 * nud_x * invDx2)) * ww;
    }
    if (i == 0) {   // setSolverQxx_BC
        dv += k.f1 * lv;
        fv -= f0 * lv;
    }
    a_out = dv;
    b_out = lv;
    c_out = uv;
    f_out = fv;
    wi_out = wi;
    drag_out = drag;
}

__host__ __device__ __forceinline__ void asm_qxx(const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                                        const Work &w, double f0, int i, int j, double &a_out, double &b_out,
                                        double &c_out, double &f_out)
{
    int wi;
    double drag;
    asm_qxx_vals(k, m, qxx_load(k, s, rf, i, j), f0, i, a_out, b_out, c_out, f_out, wi, drag);
    const long e = (long)i * k.ny + j;
    w.w[e] = (uint8_t)wi;
    w.drag[e] = drag;
}

// Qx projector, setSolverQxy_f/_d/_up/_low, src/HydroQx.cpp:262-349.
__host__ __device__ __forceinline__ void asm_qxy(const Consts &k, const State &s, const RunFields &rf, const Work &w, int i,
                                        int j, double &a_out, double &b_out, double &c_out, double &f_out)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
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
    f_out = wj * qx * (1 + dt * drag);
    const int isl = !wj;
    a_out = wj * ((1 + dt * drag) + dt * inv_j * nud_x * k.two_inv_dy2) + isl;
    const double wwp = wj && wp;
    c_out = dt * (v_p * k.inv_2dy - inv_p * (nud_x * k.inv_dy2_xy + dnudx_dy * k.inv_2dy)) * wwp;
    const double wwm = wj && wm;
    b_out = dt * (-v_m * k.inv_2dy - inv_m * (nud_x * k.inv_dy2_xy - dnudx_dy * k.inv_2dy)) * wwm;
}

// Qy predictor, src/HydroQy.cpp:26-41 with :92-230; f and d use IsWater row 0
// (src/HydroQy.cpp:140,170). IsWater (setWaterNodes_y) of every face is in w.w, written by
// the Qx tail from guess_D and guess_Zs. Side output: DragTerm of row i.
__host__ __device__ __forceinline__ void asm_qyy(const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                                        const Work &w, int i, int j, double &a_out, double &b_out, double &c_out,
                                        double &f_out)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny), jpp = wrap_inc(jp, ny);
    const double dt = k.dt;
    const double *gD = w.gD + r;
    const double g_m = gD[jm], g_j = gD[j], g_p = gD[jp], g_pp = gD[jpp];
    // masks of row i at faces jm, j, jp
    const uint8_t *wq = w.w + r;
    const int wy_j = wq[j], wy_m = wq[jm], wy_p = wq[jp];
    // mask of row 0 at face j (Q1)
    const int w0 = w.w[j];
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
    a_out = dv;
    b_out = lv;
    c_out = uv;
    f_out = fv;
    w.drag[r + j] = drag;
}

// Qy projector, setSolverQyx_* (src/HydroQy.cpp:235-362) and setSolverQyx_BC
// (src/HydroBoundaryConditions.cpp:41-60).
__host__ __device__ __forceinline__ void asm_qyx(const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                                        const Work &w, int i, int j, double &a_out, double &b_out, double &c_out,
                                        double &f_out)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
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
    a_out = dv;
    b_out = lv;
    c_out = uv;
    f_out = fv;
}

}  // namespace m70
