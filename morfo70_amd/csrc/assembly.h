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

// Loads of two adjacent doubles (an even column of a row whose length is even: 16-byte
// aligned) and of two adjacent masks.
template <bool NT = false>
__host__ __device__ __forceinline__ void ld_pair(const double *p, long e, double &a, double &b)
{
    ldn2<NT>(p + e, a, b);
}
template <bool NT = false>
__host__ __device__ __forceinline__ void ld_mask_pair(const uint8_t *p, long e, int &a, int &b)
{
    uint16_t v;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (NT) v = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(p + e));
    else v = *reinterpret_cast<const uint16_t *>(p + e);
#else
    v = *reinterpret_cast<const uint16_t *>(p + e);
#endif
    a = v & 0xff;
    b = v >> 8;
}

// Unknown kk of column j of the three-pass column solve, formed as the back-substitution
// pass forms it (k_cols_backsub): the segment's boundary values come from the reduced
// solution X, interior ones from the eliminated segment E0..E2. n unknowns per column.
__host__ __device__ __forceinline__ void cols_segment_of(const ColsWork &cw, int n, int kk, int &sg, int &q, int &m)
{
    const int nseg = cw.nseg;
    sg = (int)(((long)kk * nseg) / n);
    if ((int)(((long)(sg + 1) * n) / nseg) <= kk) ++sg;
    const int k0 = (int)(((long)sg * n) / nseg), k1 = (int)(((long)(sg + 1) * n) / nseg);
    q = kk - k0;
    m = k1 - k0;
}
template <bool STR = false>
__host__ __device__ __forceinline__ void cols_value2(const ColsWork &cw, int n, int ny, int kk, int j, double &x0, double &x1)
{
    constexpr bool NTL = Nt<NT_QXY_ROWS, STR>::L;
    constexpr bool NTF = NTL || Nt<NT_QXY_STREAM, STR>::L;   // single-use fields
    int sg, q, m;
    cols_segment_of(cw, n, kk, sg, q, m);
    if (q == 0) {
        ld_pair<NTL>(cw.X, (long)(2 * sg) * ny + j, x0, x1);
    } else if (q == m - 1) {
        ld_pair<NTL>(cw.X, (long)(2 * sg + 1) * ny + j, x0, x1);
    } else {
        double F0, F1, L0, L1, e00, e01, e10, e11, e20, e21;
        ld_pair<NTL>(cw.X, (long)(2 * sg) * ny + j, F0, F1);
        ld_pair<NTL>(cw.X, (long)(2 * sg + 1) * ny + j, L0, L1);
        const long o = (long)kk * ny + j;
        ld_pair<NTF>(cw.E0, o, e00, e01);
        ld_pair<NTF>(cw.E1, o, e10, e11);
        ld_pair<NTF>(cw.E2, o, e20, e21);
        x0 = e20 - e00 * F0 - e10 * L0;
        x1 = e21 - e01 * F1 - e11 * L1;
    }
}
__host__ __device__ __forceinline__ double cols_value(const ColsWork &cw, int n, int ny, int kk, int j)
{
    int sg, q, m;
    cols_segment_of(cw, n, kk, sg, q, m);
    const double xF = cw.X[(long)(2 * sg) * ny + j];
    const double xL = cw.X[(long)(2 * sg + 1) * ny + j];
    if (q == 0) return xF;
    if (q == m - 1) return xL;
    const long o = (long)kk * ny + j;
    return cw.E2[o] - cw.E0[o] * xF - cw.E1[o] * xL;
}

// Qx projector, setSolverQxy_f/_d/_up/_low, src/HydroQx.cpp:262-349, in two parts: the
// values element (i, j) reads (QxyIn) and the arithmetic on them (asm_qxy_vals).
struct QxyIn {
    int wj, wp, wm;                  // IsWater of the predictor at columns j, j+1, j-1
    double drag, nud_x, dnudx_dy;    // DragTerm of the predictor, nuD_x, dnuDx_dy
    double Dj0, Dj1, Dp0, Dp1, Dm0, Dm1;          // D rows i, i+1 at columns j, j+1, j-1
    double Vcp0, Vcp1, Vcm0, Vcm1;               // V_c rows i, i+1 at columns j+1, j-1
    double qx;                       // Qx* row i+1 (filled by the caller)
};

__host__ __device__ __forceinline__ QxyIn qxy_load(const Consts &k, const State &s, const RunFields &rf, const Work &w,
                                                   int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    QxyIn q;
    q.wj = w.w[r + j]; q.wp = w.w[r + jp]; q.wm = w.w[r + jm];
    q.drag = w.drag[r + j];
    q.nud_x = rf.nuD_x[r + j]; q.dnudx_dy = rf.dnuDx_dy[r + j];
    q.Dj0 = s.D[r + j]; q.Dj1 = s.D[r + ny + j];
    q.Dp0 = s.D[r + jp]; q.Dp1 = s.D[r + ny + jp];
    q.Dm0 = s.D[r + jm]; q.Dm1 = s.D[r + ny + jm];
    q.Vcp0 = s.Vc[r + jp]; q.Vcp1 = s.Vc[r + ny + jp];
    q.Vcm0 = s.Vc[r + jm]; q.Vcm1 = s.Vc[r + ny + jm];
    q.qx = 0.0;
    return q;
}

// Elements j = e and e + 1 of row i (e even, ny even), from loads of column pairs.
template <bool STR = false>
__host__ __device__ __forceinline__ void qxy_load2(const Consts &k, const State &s, const RunFields &rf, const Work &w,
                                                   int i, int e, QxyIn &a, QxyIn &b)
{
    constexpr bool NTL = Nt<NT_QXY_ROWS, STR>::L;
    constexpr bool NTF = NTL || Nt<NT_QXY_STREAM, STR>::L;   // single-use fields
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int em = (e == 0) ? ny - 2 : e - 2, ep = (e + 2 == ny) ? 0 : e + 2;
    int wmm, wm, w0, w1, wp, wpp;
    ld_mask_pair<NTL>(w.w, r + em, wmm, wm);
    ld_mask_pair<NTL>(w.w, r + e, w0, w1);
    ld_mask_pair<NTL>(w.w, r + ep, wp, wpp);
    a.wj = w0; a.wp = w1; a.wm = wm;
    b.wj = w1; b.wp = wp; b.wm = w0;
    ld_pair<NTF>(w.drag, r + e, a.drag, b.drag);
    ld_pair<NTF>(rf.nuD_x, r + e, a.nud_x, b.nud_x);
    ld_pair<NTF>(rf.dnuDx_dy, r + e, a.dnudx_dy, b.dnudx_dy);
    double d0m, d0, d0n, d0p, d1m, d1, d1n, d1p, x;
    ld_pair<NTL>(s.D, r + em, x, d0m);
    ld_pair<NTL>(s.D, r + e, d0, d0n);
    ld_pair<NTL>(s.D, r + ep, d0p, x);
    ld_pair<NTL>(s.D, r + ny + em, x, d1m);
    ld_pair<NTL>(s.D, r + ny + e, d1, d1n);
    ld_pair<NTL>(s.D, r + ny + ep, d1p, x);
    a.Dj0 = d0; a.Dj1 = d1; a.Dp0 = d0n; a.Dp1 = d1n; a.Dm0 = d0m; a.Dm1 = d1m;
    b.Dj0 = d0n; b.Dj1 = d1n; b.Dp0 = d0p; b.Dp1 = d1p; b.Dm0 = d0; b.Dm1 = d1;
    double v0m, v0, v0n, v0p, v1m, v1, v1n, v1p;
    ld_pair<NTL>(s.Vc, r + em, x, v0m);
    ld_pair<NTL>(s.Vc, r + e, v0, v0n);
    ld_pair<NTL>(s.Vc, r + ep, v0p, x);
    ld_pair<NTL>(s.Vc, r + ny + em, x, v1m);
    ld_pair<NTL>(s.Vc, r + ny + e, v1, v1n);
    ld_pair<NTL>(s.Vc, r + ny + ep, v1p, x);
    a.Vcp0 = v0n; a.Vcp1 = v1n; a.Vcm0 = v0m; a.Vcm1 = v1m;
    b.Vcp0 = v0p; b.Vcp1 = v1p; b.Vcm0 = v0; b.Vcm1 = v1;
    a.qx = b.qx = 0.0;
}

__host__ __device__ __forceinline__ void asm_qxy_vals(const Consts &k, const QxyIn &q, double &a_out, double &b_out,
                                                      double &c_out, double &f_out)
{
    const double dt = k.dt;
    const int wj = q.wj, wp = q.wp, wm = q.wm;
    const double drag = q.drag;
    const double nud_x = q.nud_x, dnudx_dy = q.dnudx_dy;
    const double inv_j = 1 / (0.5 * (q.Dj0 + q.Dj1));
    const double inv_p = 1 / (0.5 * (q.Dp0 + q.Dp1));
    const double inv_m = 1 / (0.5 * (q.Dm0 + q.Dm1));
    const double v_p = 0.5 * (q.Vcp0 + q.Vcp1);
    const double v_m = 0.5 * (q.Vcm0 + q.Vcm1);
    const double qx = q.qx;
    f_out = wj * qx * (1 + dt * drag);
    const int isl = !wj;
    a_out = wj * ((1 + dt * drag) + dt * inv_j * nud_x * k.two_inv_dy2) + isl;
    const double wwp = wj && wp;
    c_out = dt * (v_p * k.inv_2dy - inv_p * (nud_x * k.inv_dy2_xy + dnudx_dy * k.inv_2dy)) * wwp;
    const double wwm = wj && wm;
    b_out = dt * (-v_m * k.inv_2dy - inv_m * (nud_x * k.inv_dy2_xy - dnudx_dy * k.inv_2dy)) * wwm;
}

__host__ __device__ __forceinline__ void asm_qxy(const Consts &k, const State &s, const RunFields &rf, const Work &w, int i,
                                        int j, double &a_out, double &b_out, double &c_out, double &f_out)
{
    QxyIn q = qxy_load(k, s, rf, w, i, j);
    q.qx = s.Qx[(long)(i + 1) * k.ny + j];
    asm_qxy_vals(k, q, a_out, b_out, c_out, f_out);
}

// Qy predictor, src/HydroQy.cpp:26-41 with :92-230; f and d use IsWater row 0
// (src/HydroQy.cpp:140,170). IsWater (setWaterNodes_y) of every face is in w.w, written by
// the Qx tail from guess_D and guess_Zs. In two parts: the values element (i, j) reads
// (QyyIn) and the arithmetic on them (asm_qyy_vals). Side output: DragTerm of row i.
struct QyyIn {
    double g_m, g_j, g_p, g_pp;      // guess_D row i at columns j-1, j, j+1, j+2
    int wy_m, wy_j, wy_p, w0;        // IsWater row i at faces j-1, j, j+1; row 0 at face j (Q1)
    double zs0, zs1, dqx0, dqx1;     // Zs, dQx_dx row i at columns j, j+1
    double u0j, u0p, u1j, u1p;       // U rows i, i+1 at columns j, j+1 (U_c row i is formed from them)
    double vm, vj, vp;               // V row i+1 at columns j-1, j, j+1
    double Fy, Qy1, dnuDy_dx, nuD_y, cD, Urms, dnuD_dy;
};

__host__ __device__ __forceinline__ QyyIn qyy_load(const Consts &k, const State &s, const RunFields &rf, const Work &w,
                                                   int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny), jpp = wrap_inc(jp, ny);
    QyyIn q;
    const double *gD = w.gD + r;
    q.g_m = gD[jm]; q.g_j = gD[j]; q.g_p = gD[jp]; q.g_pp = gD[jpp];
    const uint8_t *wq = w.w + r;
    q.wy_j = wq[j]; q.wy_m = wq[jm]; q.wy_p = wq[jp];
    q.w0 = w.w[j];
    q.zs0 = s.Zs[r + j]; q.zs1 = s.Zs[r + jp];
    q.dqx0 = s.dQx[r + j]; q.dqx1 = s.dQx[r + jp];
    q.u0j = s.U[r + j]; q.u0p = s.U[r + jp]; q.u1j = s.U[r + ny + j]; q.u1p = s.U[r + ny + jp];
    q.vm = s.V[r + ny + jm]; q.vj = s.V[r + ny + j]; q.vp = s.V[r + ny + jp];
    q.Fy = rf.Fy[r + j]; q.Qy1 = s.Qy[r + ny + j]; q.dnuDy_dx = rf.dnuDy_dx[r + j]; q.nuD_y = rf.nuD_y[r + j];
    q.cD = rf.cD_y[r + j]; q.Urms = rf.Urms_y[r + j]; q.dnuD_dy = rf.dnuD_dy[r + j];
    return q;
}

// Elements j = e and e + 1 of row i (e even, ny even), from loads of column pairs.
template <bool STR = false>
__host__ __device__ __forceinline__ void qyy_load2(const Consts &k, const State &s, const RunFields &rf, const Work &w,
                                                   int i, int e, QyyIn &a, QyyIn &b,
                                                   bool load_cd = true)
{
    constexpr bool NTL = Nt<NT_QYY_ROWS, STR>::L;
    constexpr bool NTF = NTL || Nt<NT_QYY_STREAM, STR>::L;   // single-use fields
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int em = (e == 0) ? ny - 2 : e - 2, ep = (e + 2 == ny) ? 0 : e + 2;
    double x, gm, g0, g1, gp, gpp;
    ld_pair<NTL>(w.gD, r + em, x, gm);
    ld_pair<NTL>(w.gD, r + e, g0, g1);
    ld_pair<NTL>(w.gD, r + ep, gp, gpp);
    a.g_m = gm; a.g_j = g0; a.g_p = g1; a.g_pp = gp;
    b.g_m = g0; b.g_j = g1; b.g_p = gp; b.g_pp = gpp;
    int wmm, wm, w0, w1, wp, wpp;
    ld_mask_pair<NTL>(w.w, r + em, wmm, wm);
    ld_mask_pair<NTL>(w.w, r + e, w0, w1);
    ld_mask_pair<NTL>(w.w, r + ep, wp, wpp);
    a.wy_m = wm; a.wy_j = w0; a.wy_p = w1;
    b.wy_m = w0; b.wy_j = w1; b.wy_p = wp;
    ld_mask_pair<NTL>(w.w, e, a.w0, b.w0);
    double z0, z1, z2, q0, q1, q2;
    ld_pair<NTL>(s.Zs, r + e, z0, z1);
    ld_pair<NTL>(s.Zs, r + ep, z2, x);
    ld_pair<NTL>(s.dQx, r + e, q0, q1);
    ld_pair<NTL>(s.dQx, r + ep, q2, x);
    a.zs0 = z0; a.zs1 = z1; b.zs0 = z1; b.zs1 = z2;
    a.dqx0 = q0; a.dqx1 = q1; b.dqx0 = q1; b.dqx1 = q2;
    double u00, u01, u02, u10, u11, u12;
    ld_pair<NTL>(s.U, r + e, u00, u01);
    ld_pair<NTL>(s.U, r + ep, u02, x);
    ld_pair<NTL>(s.U, r + ny + e, u10, u11);
    ld_pair<NTL>(s.U, r + ny + ep, u12, x);
    a.u0j = u00; a.u0p = u01; a.u1j = u10; a.u1p = u11;
    b.u0j = u01; b.u0p = u02; b.u1j = u11; b.u1p = u12;
    double vm, v0, v1, vp;
    ld_pair<NTL>(s.V, r + ny + em, x, vm);
    ld_pair<NTL>(s.V, r + ny + e, v0, v1);
    ld_pair<NTL>(s.V, r + ny + ep, vp, x);
    a.vm = vm; a.vj = v0; a.vp = v1;
    b.vm = v0; b.vj = v1; b.vp = vp;
    ld_pair<NTF>(rf.Fy, r + e, a.Fy, b.Fy);
    ld_pair<NTF>(s.Qy, r + ny + e, a.Qy1, b.Qy1);
    ld_pair<NTF>(rf.dnuDy_dx, r + e, a.dnuDy_dx, b.dnuDy_dx);
    ld_pair<NTF>(rf.nuD_y, r + e, a.nuD_y, b.nuD_y);
    if (load_cd) ld_pair<NTF>(rf.cD_y, r + e, a.cD, b.cD);
    else a.cD = b.cD = k.cD0;   // constant drag model: 0.5*(cD0+cD0) is cD0 exactly
    ld_pair<NTF>(rf.Urms_y, r + e, a.Urms, b.Urms);
    ld_pair<NTF>(rf.dnuD_dy, r + e, a.dnuD_dy, b.dnuD_dy);
}

__host__ __device__ __forceinline__ void asm_qyy_vals(const Consts &k, const Metrics &m, const QyyIn &q, int i,
                                                      double &a_out, double &b_out, double &c_out, double &f_out,
                                                      double &drag_out)
{
    const double dt = k.dt;
    const double g_m = q.g_m, g_j = q.g_j, g_p = q.g_p, g_pp = q.g_pp;
    const int wy_j = q.wy_j, wy_m = q.wy_m, wy_p = q.wy_p;
    const int w0 = q.w0;
    // Di row i+1 (getD_y(i)) at j, jm, jp
    const double Dy = 0.5 * (g_p + g_j);
    const double invDy = 1 / Dy;
    const double invDy_m = 1 / (0.5 * (g_j + g_m));
    const double invDy_p = 1 / (0.5 * (g_pp + g_p));
    const double nud_y = q.nuD_y;
    double fv;
    {   // setSolverQyy_f (:124-158)
        const double inv_dx = m.inv_dxx[i];
        const double x2c_d = m.x2c_d[i], x2c_up = m.x2c_up[i];
        const double dzs_dy = k.inv_dy * (q.zs1 - q.zs0);
        const double d2qx_dxy = k.inv_dy * (q.dqx1 - q.dqx0);
        const double du_dy = k.inv_dy * (x2c_up * (q.u1p - q.u1j) + x2c_d * (q.u0p - q.u0j));
        const double d2u_dxy = inv_dx * k.inv_dy * ((q.u1p - q.u1j) - (q.u0p - q.u0j));
        fv = w0 * (q.Fy + q.Qy1 * k.inv_dt + G * Dy * (dt * d2qx_dxy - dzs_dy) + q.dnuDy_dx * du_dy +
                   nud_y * d2u_dxy);
    }
    // Ui_j = C2Y(U_c row i) and the drag with V row i+1 (old)
    // U_c row i at columns j, j+1: X2C of U, the Qx tail's own expression for U_c, from the
    // U rows loaded here; the same bits as the stored U_c, which a step need not write.
    const double uc0 = m.x2c_d[i] * q.u0j + m.x2c_up[i] * q.u1j;
    const double uc1 = m.x2c_d[i] * q.u0p + m.x2c_up[i] * q.u1p;
    const double uy = 0.5 * (uc1 + uc0);
    const double drag = drag_term(q.cD, q.Urms, uy, q.vj, invDy, k);
    // setSolverQyy_d (:160-181)
    const int isl = !w0;
    const double dv = (k.inv_dt + drag + G * Dy * dt * k.two_inv_dy2 + 2 * nud_y * invDy * k.two_inv_dy2) * w0 + isl;
    const double dnud_dy = q.dnuD_dy;
    // setSolverQyy_low (:206-230)
    const double wwm = wy_j && wy_m;
    const double lv = (0.5 * k.inv_dy * (-q.vm + 2 * dnud_dy * invDy_m) -
                       k.inv_dy2_yy * (G * Dy * dt + 2 * nud_y * invDy_m)) * wwm;
    // setSolverQyy_up (:183-204)
    const double wwp = wy_j && wy_p;
    const double uv = (0.5 * k.inv_dy * (q.vp - 2 * dnud_dy * invDy_p) -
                       k.inv_dy2_yy * (G * Dy * dt + 2 * nud_y * invDy_p)) * wwp;
    a_out = dv;
    b_out = lv;
    c_out = uv;
    f_out = fv;
    drag_out = drag;
}

__host__ __device__ __forceinline__ void asm_qyy(const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                                        const Work &w, int i, int j, double &a_out, double &b_out, double &c_out,
                                        double &f_out)
{
    double drag;
    asm_qyy_vals(k, m, qyy_load(k, s, rf, w, i, j), i, a_out, b_out, c_out, f_out, drag);
    w.drag[(long)i * k.ny + j] = drag;
}

// Qy projector, setSolverQyx_* (src/HydroQy.cpp:235-362) and setSolverQyx_BC
// (src/HydroBoundaryConditions.cpp:41-60).
template <bool STR = false>
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
    constexpr bool NTL = Nt<NT_QYX, STR>::L;
    constexpr bool NTF = NTL || Nt<NT_QYX_STREAM, STR>::L;   // single-use fields
    const int wi = ldn<NTL>(w.w + r + j);
    const double drag = ldn<NTF>(w.drag + r + j);
    const double nud = ldn<NTF>(rf.nuD_y + r + j), dnud_dx = ldn<NTF>(rf.dnuDy_dx + r + j);
    // setSolverQyx_f (:235-248)
    const double fv = wi * ldn<NTF>(s.Qy + r + ny + j) * (1 + dt * drag);
    // setSolverQyx_d (:250-285)
    double dv;
    {
        const double x2c_d = m.x2c_d[i], x2c_up = m.x2c_up[i];
        const double *u = s.U + r;
        const double dx_d = m.difx_c_d[i], d2x_d = m.dif2x_c_d[i];
        const double inv_dx = m.inv_dxx[i];
        const double inv_d = 1 / (0.5 * (ldn<NTL>(gD + r + jp) + ldn<NTL>(gD + r + j)));   // inv_Di row i+1
        const int isl = !wi;
        const double u1p = ldn<NTL>(u + ny + jp), u1j = ldn<NTL>(u + ny + j), u0p = ldn<NTL>(u + jp), u0j = ldn<NTL>(u + j);
        const double du_dx = 0.5 * inv_dx * ((u1p + u1j) - (u0p + u0j));
        const double u_y = 0.5 * (x2c_up * (u1p + u1j) + x2c_d * (u0p + u0j));
        dv = wi * ((1 + dt * drag) + dt * (u_y * dx_d + du_dx) - dt * (inv_d * nud * d2x_d + inv_d * dnud_dx * dx_d)) + isl;
    }
    // Ui_j row i (C2Y of U_c), U_c formed from U rows i, i+1 as the Qx tail forms it
    const double *uu = s.U + r;
    const double uc_j = m.x2c_d[i] * ldn<NTL>(uu + j) + m.x2c_up[i] * ldn<NTL>(uu + ny + j);
    const double uc_p = m.x2c_d[i] * ldn<NTL>(uu + jp) + m.x2c_up[i] * ldn<NTL>(uu + ny + jp);
    const double uy = 0.5 * (uc_p + uc_j);
    // setSolverQyx_up (:287-311); inv_Di row i+2 = getD_y(i+1); IsWater row nx-1 is false
    double uv;
    {
        const double inv_up = 1 / (0.5 * (ldn<NTL>(gD + r + ny + jp) + ldn<NTL>(gD + r + ny + j)));
        const int wu = (i + 1 <= nx - 2) ? ldn<NTL>(w.w + r + ny + j) : 0;
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
            const double inv_low = 1 / (0.5 * (ldn<NTL>(gD + r - ny + jp) + ldn<NTL>(gD + r - ny + j)));
            const int wl = ldn<NTL>(w.w + r - ny + j);
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
