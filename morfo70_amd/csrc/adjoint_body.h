// adjoint_body.h -- hand-written reverse (adjoint) of one Hydro ADI step, per cell.
//
// The forward step (solveQx then solveQy, src/HydroQx.cpp:18-93, src/HydroQy.cpp:17-85,
// as assembled in assembly.h and hydro_kernels.hip) is a composition of
//   X1 Qx predictor assembly -> X2 Thomas along x -> X3 Qx projector assembly ->
//   X4 periodic solve along y -> X5 tail (BC, dQx_dx, guess, U, U_c) ->
//   Y1 Qy predictor assembly -> Y2 periodic solve along y -> Y3 Qy projector assembly ->
//   Y4 Thomas along x -> Y5 tail (BC, dQy_dy, Zs, D, V, V_c).
// Its adjoint runs the stages backwards. For a solve A x = f the adjoint solves
// A^T lam = xbar and gives fbar = lam, Abar_kl = -lam_k x_l on the tridiagonal pattern;
// each assembly reverse reads lam and the taped solution to form its coefficient
// adjoints and scatters their contributions onto the assembly's inputs. Wet/dry masks
// are piecewise constant (zero derivative) and are read from the tape.
//
// Notation: S0 = state before the step, S1 = after; T = taped intermediates; Ain =
// adjoint of S1 (input), A = adjoint of S0 (accumulated); W = adjoints of intermediates;
// R = adjoints of the per-Run fields, accumulated over all steps of a Run.
//
// Contributions to arrays that several threads touch go through acc(): a relaxed
// hardware FP64 atomic add on the device, a plain add on the host. Fields owned by one
// thread in a kernel (coefficient adjoints, per-Run field adjoints at (i, j)) are
// updated in place.
#pragma once
#include "hydro_device.h"

namespace m70 {

// Intermediates of one step, written by the taped forward step.
struct StepTape {
    double *a1, *b1, *c1, *qxs;   // Qx predictor: diagonal, lower, upper; solution Qx* (rows 0..nx-2)
    double *a2, *b2, *c2;         // Qx projector
    double *a3, *b3, *c3, *qys;   // Qy predictor; solution Qy*
    double *a4, *b4, *c4;         // Qy projector
    double *f;                    // right-hand side scratch of the taped solves
    double *dragx, *dragy;        // DragTerm of the x and y half-steps
    uint8_t *wx, *wy;             // IsWater of the x and y half-steps
    double *gD;                   // guess_D
};

// Adjoints of the ten state fields (same shapes as State).
struct AdjFields {
    double *Zs, *D, *Qx, *Qy, *U, *V, *Uc, *Vc, *dQx, *dQy;
};

struct AdjWork {
    double *rhs, *lam;            // transposed solve: right-hand side and solution (rows 0..nx-2)
    double *dragx, *dragy;        // adjoints of the DragTerms
    double *UN, *UcN, *dQxN, *gD; // adjoints of U, U_c, dQx_dx after the x half-step, and of guess_D
    double *dQyN, *qyv, *qy0v, *qU;
};

struct AdjRun {
    double *cD_x, *nuD_x, *dnuD_dx, *dnuDx_dy, *cD_y, *nuD_y, *dnuD_dy, *dnuDy_dx;
    double *cD, *nuD;
};

__host__ __device__ __forceinline__ void acc(double *p, double v)
{
    if (v == 0.0) return;
#if defined(__HIP_DEVICE_COMPILE__)
    unsafeAtomicAdd(p, v);
#else
    *p += v;
#endif
}

// own at column j, left at j-1 and right at j+1 (periodic) of one row. In a kernel whose
// wavefronts are 64 consecutive columns of a row (one lane per column), the parts that
// lanes l-1 and l+1 send to column j join lane l's own part through lane shuffles, so one
// atomic adds all three. A lane adds a part itself where the receiving lane is not in its
// wavefront: lane 0 or column 0 (left), lane 63 or column ny-1 (right). Must be reached
// by every lane of the wavefront that has a column; a lane's parts may be zero.
__host__ __device__ __forceinline__ void scatter_row(double *row, int j, int ny, double own, double left,
                                                     double right)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const int lane = (int)(threadIdx.x & 63);
    const double from_r = __shfl_down(left, 1);   // lane l+1's part for column j
    const double from_l = __shfl_up(right, 1);    // lane l-1's part for column j
    const bool has_r = lane != 63 && j + 1 < ny;
    const bool has_l = lane != 0 && j > 0;
    double v = own;
    if (has_r) v += from_r;
    if (has_l) v += from_l;
    acc(row + j, v);
    if (!has_l) acc(row + wrap_dec(j, ny), left);
    if (!has_r) acc(row + wrap_inc(j, ny), right);
#else
    acc(row + j, own);
    acc(row + wrap_dec(j, ny), left);
    acc(row + wrap_inc(j, ny), right);
#endif
}

// Reverse of drag_term (src/HydroParametrizations.cpp:19-28):
// drag = cd*urms/sqrt2*sqrt(af^2 + 2(v^2+u^2)/urms^2)*inv_d; g = adjoint of drag.
__host__ __device__ __forceinline__ void drag_term_rev(double g, double cd, double urms, double u, double v,
                                                       double inv_d, const Consts &k, double &cd_b, double &u_b,
                                                       double &v_b, double &invd_b)
{
    if (g == 0.0) return;
    const double r2 = urms * urms;
    const double S = sqrt(k.alpha_f2 + 2 * (v * v + u * u) / r2);
    const double base = cd * urms * k.inv_sqrt2;
    cd_b += g * urms * k.inv_sqrt2 * S * inv_d;
    invd_b += g * base * S;
    const double q = g * base * inv_d * 2 / (r2 * S);
    u_b += q * u;
    v_b += q * v;
}

// Coefficient adjoints of a Thomas (non-periodic) system along x at row i of column j:
// x = taped solution (rows 0..n-1, stride ny). lam = adjoint solution.
struct CoefBar {
    double a, b, c, f;
};
__host__ __device__ __forceinline__ CoefBar coef_bar_cols(const double *lam, const double *x, int i, int j, int n,
                                                          int ny)
{
    const long q = (long)i * ny + j;
    const double l = lam[q];
    CoefBar cb;
    cb.a = -l * x[q];
    cb.b = (i >= 1) ? -l * x[q - ny] : 0.0;
    cb.c = (i <= n - 2) ? -l * x[q + ny] : 0.0;
    cb.f = l;
    return cb;
}
// Same for a periodic system along y at row i (corners couple j = 0 and ny-1).
__host__ __device__ __forceinline__ CoefBar coef_bar_rows(const double *lam, const double *x, int i, int j, int ny)
{
    const long r = (long)i * ny;
    const double l = lam[r + j];
    CoefBar cb;
    cb.a = -l * x[r + j];
    cb.b = -l * x[r + wrap_dec(j, ny)];
    cb.c = -l * x[r + wrap_inc(j, ny)];
    cb.f = l;
    return cb;
}

// ----------------------------------------------------------------------------------
// Y5: tail after the Qy projector (qy_tail_body). Pass a, rows 0..nx-1.
// ----------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void adj_qy_tail_a(const Consts &k, const State &S1, const AdjFields &Ain,
                                                       const AdjFields &A, const AdjWork &W, int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    const double dt = k.dt;
    const double *Dn = S1.D;
    // V[i+1][j] = Qyr[i+1][j] / (0.5 (Dn[i][j] + Dn[i][jp])), Vc[i][j] = 0.5 (V[i+1][j] + V[i+1][jm])
    const double vbar_j = Ain.V[r + ny + j] + 0.5 * (Ain.Vc[r + j] + Ain.Vc[r + jp]);
    const double vbar_m = Ain.V[r + ny + jm] + 0.5 * (Ain.Vc[r + jm] + Ain.Vc[r + j]);
    const double dav_j = 0.5 * (Dn[r + j] + Dn[r + jp]);
    const double dav_m = 0.5 * (Dn[r + jm] + Dn[r + j]);
    // Qy row i+1 after the BC (row nx holds Qy[nx-1]/3, as qy_tail stores it)
    const double q_j = S1.Qy[r + ny + j], q_m = S1.Qy[r + ny + jm];
    double dn_b = Ain.D[r + j] - 0.5 * vbar_j * q_j / (dav_j * dav_j) - 0.5 * vbar_m * q_m / (dav_m * dav_m);
    if (i <= 1) {   // V[0][j] = Qy[0][j] / (2 Dn[0][j] - Dn[1][j])
        const double den = 2 * Dn[j] - Dn[ny + j];
        const double den_b = -Ain.V[j] * S1.Qy[j] / (den * den);
        if (i == 0) {
            dn_b += 2 * den_b;
            W.qy0v[j] = Ain.V[j] / den;
        } else {
            dn_b -= den_b;
        }
    }
    A.D[r + j] += dn_b;
    A.Zs[r + j] += Ain.Zs[r + j];
    const double inc_b = dn_b + Ain.Zs[r + j];
    W.dQxN[r + j] = Ain.dQx[r + j] - dt * inc_b;
    W.dQyN[r + j] = Ain.dQy[r + j] - dt * inc_b;
    W.qyv[r + j] = vbar_j / dav_j;
}

// Pass b: right-hand side of the transposed Qy projector solve (Qy rows 1..nx-1), i = row-1.
__host__ __device__ __forceinline__ void adj_qy_tail_b(const Consts &k, const AdjFields &Ain, const AdjWork &W, int i,
                                                       int j)
{
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;   // Qy row kq = i + 1
    const int jp = wrap_inc(j, ny);
    double qb = W.qyv[r + j] + k.inv_dy * (W.dQyN[r + j] - W.dQyN[r + jp]) + Ain.Qy[r + ny + j];
    if (i == 0) qb += k.bcQy * (Ain.Qy[j] + W.qy0v[j]);   // Qy[0] = bcQy * Qy[1]
    if (i == nx - 2) {                                     // Qy[nx] = Qy[nx-1] / 3
        const long rl = (long)(nx - 1) * ny;
        qb += k.inv_3 * (Ain.Qy[rl + ny + j] + W.qyv[rl + j] + k.inv_dy * (W.dQyN[rl + j] - W.dQyN[rl + jp]));
    }
    W.rhs[r + j] = qb;
}

// ----------------------------------------------------------------------------------
// Y3: Qy projector assembly (asm_qyx), rows 0..nx-2. lam = transposed Thomas solution.
// ----------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void adj_qyx(const Consts &k, const Metrics &m, const State &S1,
                                                 const RunFields &rf, const StepTape &T, const AdjWork &W,
                                                 const AdjRun &R, int i, int j)
{
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny);
    const double dt = k.dt;
    const CoefBar cb = coef_bar_cols(W.lam, S1.Qy + ny, i, j, nx - 1, ny);
    // scatter values, applied at the end through scatter_row (zero where a branch is not taken)
    double s_cu = 0, s_cd = 0, s_da = 0, s_du = 0, s_dl = 0;
    double lv_b = cb.b, uv_b = cb.c;
    const double dv_b = cb.a;
    if (i == 0) lv_b += k.bcQy * cb.a;
    if (i == nx - 2) uv_b += k.inv_3 * cb.a;
    const int wi = T.wy[r + j];
    const double drag = T.dragy[r + j];
    const double nud = rf.nuD_y[r + j], dnud_dx = rf.dnuDy_dx[r + j];
    const double *gD = T.gD;
    double qb = 0, drag_b = 0, nud_b = 0, dnud_b = 0, uy_b = 0;
    if (wi) {
        qb = cb.f * (1 + dt * drag);
        drag_b = cb.f * T.qys[r + j] * dt + dv_b * dt;
        const double x2c_d = m.x2c_d[i], x2c_up = m.x2c_up[i];
        const double dx_d = m.difx_c_d[i], d2x_d = m.dif2x_c_d[i];
        const double inv_dx = m.inv_dxx[i];
        const double inv_d = 1 / (0.5 * (gD[r + jp] + gD[r + j]));
        const double uyl_b = dv_b * dt * dx_d;   // adjoint of u_y
        const double dudx_b = dv_b * dt;
        const double invd_b = -dv_b * dt * (nud * d2x_d + dnud_dx * dx_d);
        nud_b -= dv_b * dt * inv_d * d2x_d;
        dnud_b -= dv_b * dt * inv_d * dx_d;
        const double cu = 0.5 * x2c_up * uyl_b + 0.5 * inv_dx * dudx_b;
        const double cd = 0.5 * x2c_d * uyl_b - 0.5 * inv_dx * dudx_b;
        s_cu = cu;
        s_cd = cd;
        const double da_b = -invd_b * inv_d * inv_d;
        s_da = 0.5 * da_b;
    }
    {   // up: inv_Di row i+2; IsWater row nx-1 is false
        const int wu = (i + 1 <= nx - 2) ? T.wy[r + ny + j] : 0;
        if (wi && wu && uv_b != 0.0) {
            const double e = uv_b * dt;
            const double inv_up = 1 / (0.5 * (gD[r + ny + jp] + gD[r + ny + j]));
            const double dx_up = m.difx_c_up[i], d2x_up = m.dif2x_c_up[i];
            uy_b += e * dx_up;
            nud_b -= e * d2x_up * inv_up;
            dnud_b -= e * dx_up * inv_up;
            const double iu_b = -e * (nud * d2x_up + dnud_dx * dx_up);
            const double du_b = -iu_b * inv_up * inv_up;
            s_du = 0.5 * du_b;
        }
    }
    {   // low
        const double dx_low = m.difx_c_low[i], d2x_low = m.dif2x_c_low[i];
        const int mask = (i == 0) ? wi : (wi && T.wy[r - ny + j]);
        if (mask && lv_b != 0.0) {
            const double e = lv_b * dt;
            double inv_low;
            if (i == 0) {
                double Dy0;
                if (j < ny - 1)
                    Dy0 = 1 * (gD[j + 1] + gD[j]) - 0.5 * (gD[ny + j + 1] + gD[ny + j]);
                else
                    Dy0 = 1 * (gD[0] + gD[ny - 1]) - 0.5 * (gD[0] + gD[2 * ny - 1]);
                inv_low = 1 / Dy0;
            } else {
                inv_low = 1 / (0.5 * (gD[r - ny + jp] + gD[r - ny + j]));
            }
            uy_b += e * dx_low;
            nud_b -= e * d2x_low * inv_low;
            dnud_b -= e * dx_low * inv_low;
            const double il_b = -e * (nud * d2x_low + dnud_dx * dx_low);
            const double dl_b = -il_b * inv_low * inv_low;
            if (i == 0) {   // getD_y0 with its wrap element (src/HydroQy.cpp:426-440)
                if (j < ny - 1) {
                    acc(&W.gD[j + 1], dl_b);
                    acc(&W.gD[j], dl_b);
                    acc(&W.gD[ny + j + 1], -0.5 * dl_b);
                    acc(&W.gD[ny + j], -0.5 * dl_b);
                } else {
                    acc(&W.gD[0], 0.5 * dl_b);
                    acc(&W.gD[ny - 1], dl_b);
                    acc(&W.gD[2 * ny - 1], -0.5 * dl_b);
                }
            } else {
                s_dl = 0.5 * dl_b;
            }
        }
    }
    W.rhs[r + j] = qb;
    W.dragy[r + j] = drag_b;
    R.nuD_y[r + j] += nud_b;
    R.dnuDy_dx[r + j] += dnud_b;
    scatter_row(W.UN + r + ny, j, ny, s_cu, 0.0, s_cu);
    scatter_row(W.UN + r, j, ny, s_cd, 0.0, s_cd);
    scatter_row(W.gD + r, j, ny, s_da, 0.0, s_da);
    if (i + 1 <= nx - 2) scatter_row(W.gD + r + ny, j, ny, s_du, 0.0, s_du);
    if (i > 0) scatter_row(W.gD + r - ny, j, ny, s_dl, 0.0, s_dl);
    scatter_row(W.UcN + r, j, ny, 0.5 * uy_b, 0.0, 0.5 * uy_b);
}

// ----------------------------------------------------------------------------------
// Y1: Qy predictor assembly (asm_qyy), rows 0..nx-2. lam = transposed periodic solution.
// ----------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void adj_qyy(const Consts &k, const Metrics &m, const State &S0, const State &S1,
                                                 const RunFields &rf, const StepTape &T, const AdjFields &A,
                                                 const AdjWork &W, const AdjRun &R, int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny), jpp = wrap_inc(jp, ny);
    const double dt = k.dt;
    const CoefBar cb = coef_bar_rows(W.lam, T.qys, i, j, ny);
    const double *gD = T.gD + r;
    const double g_m = gD[jm], g_j = gD[j], g_p = gD[jp], g_pp = gD[jpp];
    const int wy_j = T.wy[r + j], wy_m = T.wy[r + jm], wy_p = T.wy[r + jp];
    const int w0 = T.wy[j];
    const double Dy = 0.5 * (g_p + g_j);
    const double invDy = 1 / Dy;
    const double invDy_m = 1 / (0.5 * (g_j + g_m));
    const double invDy_p = 1 / (0.5 * (g_pp + g_p));
    const double nud_y = rf.nuD_y[r + j];
    const double dnud_dy = rf.dnuD_dy[r + j];
    double Dy_b = 0, invDy_b = 0, invm_b = 0, invp_b = 0, nud_b = 0, dnud_b = 0;
    double drag_b = W.dragy[r + j];
    // scatter values, applied after the last load (an atomic ahead of a load holds the load)
    double s_qy = 0, s_dq = 0, s_dz = 0, r_dy = 0, s_u1p = 0, s_u1j = 0, s_u0p = 0, s_u0j = 0, s_vm = 0, s_vp = 0;
    if (w0) {
        const double g = cb.f;
        const double inv_dx = m.inv_dxx[i];
        const double x2c_d = m.x2c_d[i], x2c_up = m.x2c_up[i];
        const double *zs = S0.Zs + r, *dqx = S1.dQx + r, *u = S1.U + r;
        const double dzs_dy = k.inv_dy * (zs[jp] - zs[j]);
        const double d2qx_dxy = k.inv_dy * (dqx[jp] - dqx[j]);
        const double du_dy = k.inv_dy * (x2c_up * (u[ny + jp] - u[ny + j]) + x2c_d * (u[jp] - u[j]));
        const double d2u_dxy = inv_dx * k.inv_dy * ((u[ny + jp] - u[ny + j]) - (u[jp] - u[j]));
        s_qy = g * k.inv_dt;
        Dy_b += g * G * (dt * d2qx_dxy - dzs_dy);
        const double t = g * G * Dy;
        s_dq = t * dt * k.inv_dy;
        s_dz = -t * k.inv_dy;
        r_dy = g * du_dy;
        nud_b += g * d2u_dxy;
        const double du_b = g * rf.dnuDy_dx[r + j] * k.inv_dy;
        const double c2 = g * nud_y * inv_dx * k.inv_dy;
        s_u1p = x2c_up * du_b + c2;
        s_u1j = -x2c_up * du_b - c2;
        s_u0p = x2c_d * du_b - c2;
        s_u0j = -x2c_d * du_b + c2;
        const double h = cb.a;
        drag_b += h;
        Dy_b += h * G * dt * k.two_inv_dy2;
        nud_b += h * 2 * invDy * k.two_inv_dy2;
        invDy_b += h * 2 * nud_y * k.two_inv_dy2;
    }
    const double *v = S0.V + r + ny;
    if (wy_j && wy_m && cb.b != 0.0) {
        const double e = cb.b;
        s_vm = -0.5 * k.inv_dy * e;
        dnud_b += e * k.inv_dy * invDy_m;
        invm_b += e * (k.inv_dy * dnud_dy - k.inv_dy2_yy * 2 * nud_y);
        Dy_b -= e * k.inv_dy2_yy * G * dt;
        nud_b -= e * k.inv_dy2_yy * 2 * invDy_m;
    }
    if (wy_j && wy_p && cb.c != 0.0) {
        const double e = cb.c;
        s_vp = 0.5 * k.inv_dy * e;
        dnud_b -= e * k.inv_dy * invDy_p;
        invp_b += e * (-k.inv_dy * dnud_dy - k.inv_dy2_yy * 2 * nud_y);
        Dy_b -= e * k.inv_dy2_yy * G * dt;
        nud_b -= e * k.inv_dy2_yy * 2 * invDy_p;
    }
    {   // drag with Ui_j = C2Y(U_c) and V row i+1
        const double uy = 0.5 * (S1.Uc[r + jp] + S1.Uc[r + j]);
        double cd_b = 0, uy_b = 0, v_b = 0;
        drag_term_rev(drag_b, rf.cD_y[r + j], rf.Urms_y[r + j], uy, v[j], invDy, k, cd_b, uy_b, v_b, invDy_b);
        Dy_b -= invDy_b * invDy * invDy;
        const double m_b = -invm_b * invDy_m * invDy_m;
        const double p_b = -invp_b * invDy_p * invDy_p;
        if (w0) R.dnuDy_dx[r + j] += r_dy;
        R.cD_y[r + j] += cd_b;
        R.nuD_y[r + j] += nud_b;
        R.dnuD_dy[r + j] += dnud_b;
        acc(&A.Qy[r + ny + j], s_qy);
        scatter_row(W.dQxN + r, j, ny, -s_dq, 0.0, s_dq);
        scatter_row(A.Zs + r, j, ny, -s_dz, 0.0, s_dz);
        scatter_row(W.UN + r + ny, j, ny, s_u1j, 0.0, s_u1p);
        scatter_row(W.UN + r, j, ny, s_u0j, 0.0, s_u0p);
        scatter_row(A.V + r + ny, j, ny, v_b, s_vm, s_vp);
        scatter_row(W.UcN + r, j, ny, 0.5 * uy_b, 0.0, 0.5 * uy_b);
        scatter_row(W.gD + r, j, ny, 0.5 * Dy_b + 0.5 * m_b, 0.5 * m_b, 0.5 * Dy_b + 0.5 * p_b);
        acc(&W.gD[r + jpp], 0.5 * p_b);
    }
}

// ----------------------------------------------------------------------------------
// X5: tail after the Qx projector (qx_tail_body). Pass a, rows 0..nx-1.
// ----------------------------------------------------------------------------------
__host__ __device__ __forceinline__ double x5_ubar(const Consts &k, const Metrics &m, const AdjFields &Ain,
                                                   const AdjWork &W, int i, int j)
{   // total adjoint of U_new row i (0..nx-1): direct, through later stages, through U_c
    const int ny = k.ny;
    const long r = (long)i * ny;
    double u = Ain.U[r + j] + W.UN[r + j] + m.x2c_d[i] * (Ain.Uc[r + j] + W.UcN[r + j]);
    if (i >= 1) u += m.x2c_up[i - 1] * (Ain.Uc[r - ny + j] + W.UcN[r - ny + j]);
    return u;
}

__host__ __device__ __forceinline__ void adj_qx_tail_a(const Consts &k, const Metrics &m, const State &S1,
                                                       const StepTape &T, const AdjFields &Ain, const AdjFields &A,
                                                       const AdjWork &W, int i, int j)
{
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;
    const double dt = k.dt;
    const double *gD = T.gD, *q = S1.Qx;   // S1.Qx row 0 holds the BC value
    double gd_b = W.gD[r + j];
    double qU;
    const double ub = x5_ubar(k, m, Ain, W, i, j);
    if (i == 0) {
        const double dx0 = 1.5 * gD[j] - 0.5 * gD[ny + j];
        qU = ub / dx0;
        gd_b += 1.5 * (-ub * q[j] / (dx0 * dx0));
    } else {
        const double dav = 0.5 * (gD[r - ny + j] + gD[r + j]);
        qU = ub / dav;
        gd_b += 0.5 * (-ub * q[r + j] / (dav * dav));
    }
    if (i + 1 <= nx - 1) {
        const double ub1 = x5_ubar(k, m, Ain, W, i + 1, j);
        const double dav1 = 0.5 * (gD[r + j] + gD[r + ny + j]);
        gd_b += 0.5 * (-ub1 * q[r + ny + j] / (dav1 * dav1));
    }
    if (i == 1) {
        const double ub0 = x5_ubar(k, m, Ain, W, 0, j);
        const double dx0 = 1.5 * gD[j] - 0.5 * gD[ny + j];
        gd_b -= 0.5 * (-ub0 * q[j] / (dx0 * dx0));
    }
    A.D[r + j] += gd_b;
    W.dQxN[r + j] -= dt * gd_b;
    A.dQy[r + j] -= dt * gd_b;
    W.qU[r + j] = qU;
    if (i == nx - 1) {   // U row nx is carried over unchanged; U_c row nx-1 reads it
        const long rn = (long)nx * ny;
        A.U[rn + j] += Ain.U[rn + j] + W.UN[rn + j] + m.x2c_up[nx - 1] * (Ain.Uc[r + j] + W.UcN[r + j]);
    }
}

// Pass b: right-hand side of the transposed Qx projector solve (Qx rows 1..nx-1), i = row-1.
__host__ __device__ __forceinline__ void adj_qx_tail_b(const Consts &k, const Metrics &m, const AdjFields &Ain,
                                                       const AdjFields &A, const AdjWork &W, int i, int j)
{
    const int ny = k.ny, nx = k.nx;
    const int kq = i + 1;
    const long rq = (long)kq * ny;
    double qb = Ain.Qx[rq + j] + W.qU[rq + j] + m.inv_dxx[kq - 1] * W.dQxN[rq - ny + j] - m.inv_dxx[kq] * W.dQxN[rq + j];
    if (kq == 1) {   // Qx[0] = f1 * Qx[1] + f0
        const double qb0 = Ain.Qx[j] + W.qU[j] - m.inv_dxx[0] * W.dQxN[j];
        qb += k.f1 * qb0;
    }
    if (kq == nx - 1) {   // Qx row nx is carried over unchanged; dQx_dx row nx-1 reads it
        const long rn = (long)nx * ny;
        A.Qx[rn + j] += Ain.Qx[rn + j] + m.inv_dxx[nx - 1] * W.dQxN[rn - ny + j];
    }
    W.rhs[(long)i * ny + j] = qb;
}

// ----------------------------------------------------------------------------------
// X3: Qx projector assembly (asm_qxy), rows 0..nx-2. lam = transposed periodic solution.
// ----------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void adj_qxy(const Consts &k, const State &S0, const State &S1,
                                                 const RunFields &rf, const StepTape &T, const AdjFields &A,
                                                 const AdjWork &W, const AdjRun &R, int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    const double dt = k.dt;
    const CoefBar cb = coef_bar_rows(W.lam, S1.Qx + ny, i, j, ny);
    const double *D = S0.D;
    const int wj = T.wx[r + j], wp = T.wx[r + jp], wm = T.wx[r + jm];
    const double drag = T.dragx[r + j];
    const double nud_x = rf.nuD_x[r + j], dnudx_dy = rf.dnuDx_dy[r + j];
    const double inv_j = 1 / (0.5 * (D[r + j] + D[r + ny + j]));
    const double inv_p = 1 / (0.5 * (D[r + jp] + D[r + ny + jp]));
    const double inv_m = 1 / (0.5 * (D[r + jm] + D[r + ny + jm]));
    double qb = 0, drag_b = 0, ij_b = 0, ip_b = 0, im_b = 0, vp_b = 0, vm_b = 0, nud_b = 0, dnud_b = 0;
    if (wj) {
        qb = cb.f * (1 + dt * drag);
        drag_b = cb.f * T.qxs[r + j] * dt + cb.a * dt;
        ij_b = cb.a * dt * nud_x * k.two_inv_dy2;
        nud_b += cb.a * dt * inv_j * k.two_inv_dy2;
        if (wp) {
            const double e = cb.c * dt;
            vp_b = e * k.inv_2dy;
            ip_b = -e * (nud_x * k.inv_dy2_xy + dnudx_dy * k.inv_2dy);
            nud_b -= e * inv_p * k.inv_dy2_xy;
            dnud_b -= e * inv_p * k.inv_2dy;
        }
        if (wm) {
            const double e = cb.b * dt;
            vm_b = -e * k.inv_2dy;
            im_b = -e * (nud_x * k.inv_dy2_xy - dnudx_dy * k.inv_2dy);
            nud_b -= e * inv_m * k.inv_dy2_xy;
            dnud_b += e * inv_m * k.inv_2dy;
        }
    }
    const double dj = -0.5 * ij_b * inv_j * inv_j, dp = -0.5 * ip_b * inv_p * inv_p, dm = -0.5 * im_b * inv_m * inv_m;
    scatter_row(A.D + r, j, ny, dj, dm, dp);
    scatter_row(A.D + r + ny, j, ny, dj, dm, dp);
    scatter_row(A.Vc + r, j, ny, 0.0, 0.5 * vm_b, 0.5 * vp_b);
    scatter_row(A.Vc + r + ny, j, ny, 0.0, 0.5 * vm_b, 0.5 * vp_b);
    W.rhs[r + j] = qb;
    W.dragx[r + j] = drag_b;
    R.nuD_x[r + j] += nud_b;
    R.dnuDx_dy[r + j] += dnud_b;
}

// ----------------------------------------------------------------------------------
// X1: Qx predictor assembly (asm_qxx), rows 0..nx-2. lam = transposed Thomas solution.
// ----------------------------------------------------------------------------------
// The adjoint parts element (i, j) of the Qx predictor sends to the state before the step;
// the per-Run field adjoints it owns (R) are updated in place.
struct QxxParts {
    double qx;                  // A.Qx row i+1
    double dq, dz;              // A.dQy and A.Zs: +dq/+dz at row i+1, -dq/-dz at row i
    double v1j, v1m, v0j, v0m;  // A.V row i+2 (v1) and row i+1 (v0) at columns j and j-1
    double ul, u1, uu;          // A.U rows i, i+1, i+2
    double vc;                  // A.Vc rows i and i+1, each
    double dm, d0, d1, d2;      // A.D rows i-1, i, i+1, i+2
};
__host__ __device__ __forceinline__ QxxParts adj_qxx_parts(const Consts &k, const Metrics &m, const State &S0,
                                                           const RunFields &rf, const StepTape &T, const AdjWork &W,
                                                           const AdjRun &R, double f0, int i, int j)
{
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;
    const int jm = wrap_dec(j, ny);
    const double dt = k.dt;
    const CoefBar cb = coef_bar_cols(W.lam, T.qxs, i, j, nx - 1, ny);
    const int wi = T.wx[r + j];
    const double *D = S0.D;
    const double Di = D[r + j], Di1 = D[r + ny + j];
    const double Dx1 = 0.5 * (Di + Di1);
    const double invDx1 = 1 / Dx1;
    const double vx = 0.5 * (S0.Vc[r + j] + S0.Vc[r + ny + j]);
    const double u1 = S0.U[r + ny + j];
    const double dnud_dx = rf.dnuD_dx[r + j], nud_x = rf.nuD_x[r + j];
    double lv_b = cb.b;
    if (i == 0) lv_b += k.f1 * cb.a - f0 * cb.f;   // setSolverQxx_BC
    double Dx1_b = 0, invDx1_b = 0, u1_b = 0, dnud_b = 0, nud_b = 0, drag_b = W.dragx[r + j];
    double dD_m = 0, dD_0 = 0, dD_1 = 0, dD_2 = 0;
    // scatter values, applied after the last load (an atomic ahead of a load holds the load)
    double s_qx = 0, s_dq = 0, s_dz = 0, r_dy = 0, s_v1j = 0, s_v1m = 0, s_v0j = 0, s_v0m = 0, s_ul = 0, s_uu = 0;
    if (wi) {
        const double g = cb.f;
        const double inv_dx = m.inv_dx[i];
        const double *zs = S0.Zs + r, *dqy = S0.dQy + r, *v = S0.V + r + ny;
        const double dzs_dx = inv_dx * (zs[ny + j] - zs[j]);
        const double d2qy_dxy = inv_dx * (dqy[ny + j] - dqy[j]);
        const double dV_dx = 0.5 * inv_dx * ((v[ny + j] + v[ny + jm]) - (v[j] + v[jm]));
        const double d2V_dxy = inv_dx * k.inv_dy * ((v[ny + j] - v[ny + jm]) - (v[j] - v[jm]));
        s_qx = g * k.inv_dt;
        Dx1_b += g * G * (dt * d2qy_dxy - dzs_dx);
        const double t = g * G * Dx1;
        s_dq = t * dt * inv_dx;
        s_dz = -t * inv_dx;
        r_dy = g * dV_dx;
        nud_b += g * d2V_dxy;
        const double dv_b = g * rf.dnuDx_dy[r + j] * 0.5 * inv_dx;
        const double c2 = g * nud_x * inv_dx * k.inv_dy;
        s_v1j = dv_b + c2;
        s_v1m = dv_b - c2;
        s_v0j = -dv_b - c2;
        s_v0m = -dv_b + c2;
        const double h = cb.a;
        drag_b += h;
        u1_b += h * m.difx_x_d[i];
        dnud_b += h * m.difx_x_d[i] * (-2 * invDx1);
        invDx1_b += h * m.difx_x_d[i] * (-2 * dnud_dx);
        Dx1_b -= h * m.dif2x_x_d[i] * G * dt;
        nud_b -= h * m.dif2x_x_d[i] * 2 * invDx1;
        invDx1_b -= h * m.dif2x_x_d[i] * 2 * nud_x;
    }
    if (i == 0) {   // row-0 low is unmasked (src/HydroQx.cpp:214-229)
        const double e = lv_b;
        const double Dx0 = 1.5 * Di - 0.5 * Di1, invDx0 = 1.0 / Dx0;
        s_ul = e * m.difx_x_low[0];   // A.U row 0
        dnud_b += e * m.difx_x_low[0] * (-2 * invDx0);
        const double i0_b = e * m.difx_x_low[0] * (-2 * dnud_dx) - e * m.dif2x_x_low[0] * 2 * nud_x;
        Dx1_b -= e * m.dif2x_x_low[0] * G * dt;
        nud_b -= e * m.dif2x_x_low[0] * 2 * invDx0;
        const double d0_b = -i0_b * invDx0 * invDx0;
        dD_0 += 1.5 * d0_b;
        dD_1 -= 0.5 * d0_b;
    } else if (wi && T.wx[r - ny + j]) {
        const double e = lv_b;
        const double Dxi = 0.5 * (D[r - ny + j] + Di), invDxi = 1 / Dxi;
        s_ul = e * m.difx_x_low[i];
        dnud_b += e * m.difx_x_low[i] * (-2 * invDxi);
        const double ii_b = e * m.difx_x_low[i] * (-2 * dnud_dx) - e * m.dif2x_x_low[i] * 2 * nud_x;
        Dx1_b -= e * m.dif2x_x_low[i] * G * dt;
        nud_b -= e * m.dif2x_x_low[i] * 2 * invDxi;
        const double di_b = -ii_b * invDxi * invDxi;
        dD_m += 0.5 * di_b;
        dD_0 += 0.5 * di_b;
    }
    if (i < nx - 2 && wi && T.wx[r + ny + j]) {
        const double e = cb.c;
        const double Dx2 = 0.5 * (Di1 + D[r + 2L * ny + j]), invDx2 = 1 / Dx2;
        s_uu = e * m.difx_x_up[i];
        dnud_b += e * m.difx_x_up[i] * (-2 * invDx2);
        const double i2_b = e * m.difx_x_up[i] * (-2 * dnud_dx) - e * m.dif2x_x_up[i] * 2 * nud_x;
        Dx1_b -= e * m.dif2x_x_up[i] * G * dt;
        nud_b -= e * m.dif2x_x_up[i] * 2 * invDx2;
        const double d2_b = -i2_b * invDx2 * invDx2;
        dD_1 += 0.5 * d2_b;
        dD_2 += 0.5 * d2_b;
    }
    double cd_b = 0, vx_b = 0;
    drag_term_rev(drag_b, rf.cD_x[r + j], rf.Urms_x[r + j], u1, vx, invDx1, k, cd_b, u1_b, vx_b, invDx1_b);
    Dx1_b -= invDx1_b * invDx1 * invDx1;
    dD_0 += 0.5 * Dx1_b;
    dD_1 += 0.5 * Dx1_b;
    if (wi) R.dnuDx_dy[r + j] += r_dy;
    R.cD_x[r + j] += cd_b;
    R.dnuD_dx[r + j] += dnud_b;
    R.nuD_x[r + j] += nud_b;
    QxxParts p;
    p.qx = s_qx;
    p.dq = s_dq;
    p.dz = s_dz;
    p.v1j = s_v1j; p.v1m = s_v1m; p.v0j = s_v0j; p.v0m = s_v0m;
    p.ul = s_ul;
    p.u1 = u1_b;
    p.uu = (i < nx - 2) ? s_uu : 0.0;
    p.vc = 0.5 * vx_b;
    p.dm = (i > 0) ? dD_m : 0.0;
    p.d0 = dD_0;
    p.d1 = dD_1;
    p.d2 = (i < nx - 2) ? dD_2 : 0.0;
    return p;
}
// Rows per run of the Qx predictor's reverse (adj_qxx_rows): 2 and 8 took 0.5% and 3%
// longer than 4 over the C4 gradient.
constexpr int ADJ_RUN = 4;

// The Qx predictor's reverse over rows k0..k1-1 of column j (one lane per column, a
// wavefront 64 consecutive columns). Element i sends parts to rows i-1..i+2; the row no
// later element of the run reaches is added once, with every part the run has for it, so
// a row takes one atomic per field instead of up to four (rows the run shares with the
// runs above and below it take one from each). Every lane of the wavefront that has a
// column must call it with the same rows.
__host__ __device__ __forceinline__ void adj_qxx_rows(const Consts &k, const Metrics &m, const State &S0,
                                                      const RunFields &rf, const StepTape &T, const AdjFields &A,
                                                      const AdjWork &W, const AdjRun &R, double f0, int k0, int k1,
                                                      int j)
{
    const int ny = k.ny, nx = k.nx;
    auto at = [&](int row) { return (long)row * ny + j; };
    // pending parts by row: dQy, Zs, Vc rows i, i+1; U rows i..i+2; D rows i-1..i+2;
    // V rows i+1, i+2 (own column and the one to the left)
    double dq0 = 0, dq1 = 0, dz0 = 0, dz1 = 0, vc0 = 0, vc1 = 0, u0 = 0, u1 = 0, u2 = 0;
    double dm = 0, d0 = 0, d1 = 0, d2 = 0, v0 = 0, v0l = 0, v1 = 0, v1l = 0;
    for (int i = k0; i < k1; ++i) {
        const QxxParts p = adj_qxx_parts(k, m, S0, rf, T, W, R, f0, i, j);
        acc(&A.Qx[at(i + 1)], p.qx);
        dq0 -= p.dq; dq1 += p.dq;
        dz0 -= p.dz; dz1 += p.dz;
        vc0 += p.vc; vc1 += p.vc;
        u0 += p.ul; u1 += p.u1; u2 += p.uu;
        dm += p.dm; d0 += p.d0; d1 += p.d1; d2 += p.d2;
        v0 += p.v0j; v0l += p.v0m; v1 += p.v1j; v1l += p.v1m;
        // rows element i+1 no longer reaches: i for dQy, Zs, Vc and U; i-1 for D; i+1 for V
        acc(&A.dQy[at(i)], dq0);
        acc(&A.Zs[at(i)], dz0);
        acc(&A.Vc[at(i)], vc0);
        acc(&A.U[at(i)], u0);
        if (i > 0) acc(&A.D[at(i - 1)], dm);
        scatter_row(A.V + at(i + 1) - j, j, ny, v0, v0l, 0.0);
        dq0 = dq1; dq1 = 0; dz0 = dz1; dz1 = 0; vc0 = vc1; vc1 = 0;
        u0 = u1; u1 = u2; u2 = 0;
        dm = d0; d0 = d1; d1 = d2; d2 = 0;
        v0 = v1; v0l = v1l; v1 = 0; v1l = 0;
    }
    // the rows past the run: k1 for dQy, Zs, Vc; k1, k1+1 for U; k1-1..k1+1 for D; k1+1 for V
    acc(&A.dQy[at(k1)], dq0);
    acc(&A.Zs[at(k1)], dz0);
    acc(&A.Vc[at(k1)], vc0);
    acc(&A.U[at(k1)], u0);
    if (k1 + 1 <= nx) acc(&A.U[at(k1 + 1)], u1);
    if (k1 > 0) acc(&A.D[at(k1 - 1)], dm);
    if (k1 < nx) acc(&A.D[at(k1)], d0);
    if (k1 + 1 < nx) acc(&A.D[at(k1 + 1)], d1);
    scatter_row(A.V + at(k1 + 1) - j, j, ny, v0, v0l, 0.0);
}

// ----------------------------------------------------------------------------------
// Per-Run fields (run_fields_body, drag_viscosity_body), once at the end of the sweep.
// ----------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void adj_run_fields(const Consts &k, const Metrics &m, const AdjRun &R, int i,
                                                        int j)
{
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    double *cD = R.cD, *nuD = R.nuD;
    if (i < nx - 1) {
        const double cx = 0.5 * R.cD_x[r + j];
        acc(&cD[r + j], cx);
        acc(&cD[r + ny + j], cx);
        const double nx_b = 0.5 * R.nuD_x[r + j];
        const double dx_b = m.inv_dx[i] * R.dnuD_dx[r + j];
        acc(&nuD[r + j], nx_b - dx_b);
        acc(&nuD[r + ny + j], nx_b + dx_b);
        const double dy_b = 0.5 * k.inv_2dy_c2c * R.dnuDx_dy[r + j];
        acc(&nuD[r + jp], dy_b);
        acc(&nuD[r + ny + jp], dy_b);
        acc(&nuD[r + jm], -dy_b);
        acc(&nuD[r + ny + jm], -dy_b);
    }
    const double cy = 0.5 * R.cD_y[r + j];
    acc(&cD[r + jp], cy);
    acc(&cD[r + j], cy);
    double nuy_b = R.nuD_y[r + j];
    const double ddy = R.dnuD_dy[r + j] * k.inv_dy;
    acc(&nuD[r + jp], ddy);
    acc(&nuD[r + j], -ddy);
    const double g = R.dnuDy_dx[r + j];
    double vl_b = 0, vu_b = 0;
    if (i == 0) {
        nuy_b += g * (2 * m.difx_c_low[0] + m.difx_c_d[0]);
        vu_b = g * (m.difx_c_up[0] - m.difx_c_low[0]);
    } else if (i < nx - 1) {
        nuy_b += g * m.difx_c_d[i];
        vl_b = g * m.difx_c_low[i];
        vu_b = g * m.difx_c_up[i];
    } else {
        nuy_b += g * m.difx_c_d[nx - 1];
        vl_b = g * m.difx_c_low[nx - 1];
    }
    acc(&nuD[r + jp], 0.5 * nuy_b);
    acc(&nuD[r + j], 0.5 * nuy_b);
    if (vl_b != 0.0) {
        acc(&nuD[r - ny + jp], 0.5 * vl_b);
        acc(&nuD[r - ny + j], 0.5 * vl_b);
    }
    if (vu_b != 0.0) {
        acc(&nuD[r + ny + jp], 0.5 * vu_b);
        acc(&nuD[r + ny + j], 0.5 * vu_b);
    }
}

// Depth and parameter adjoints of the drag and viscosity models at one cell; adds to
// Dbar and returns the cell's contributions to (cD0, nu0, M).
__host__ __device__ __forceinline__ void adj_drag_visc(const Consts &k, const double *D, const Forcing &fw,
                                                       const AdjRun &R, double *Dbar, long q, double &cD0_b,
                                                       double &nu0_b, double &M_b)
{
    const double d = D[q];
    const double cb = R.cD[q], nb = R.nuD[q];
    double db = 0.0;
    if (k.drag_model == 0) {
        cD0_b += cb;
    } else if (k.drag_model == 1) {
        if (!(d < k.z0 + 0.001)) {
            const double L = log(d / k.z0) - 1.0;
            const double c = 0.4 / L;
            db += cb * 2 * c * (-0.4 / (L * L)) / d;
        }
    } else {
        db += cb * (-(0.015 * cbrt(k.ka / d)) / (3 * d));
    }
    nu0_b += nb * d;
    if (k.visc_model == 0) {
        db += nb * k.nu0;
    } else if (k.visc_model == 1 || k.visc_model == 2) {
        const double X = (k.visc_model == 1) ? ((1 - k.alpha_r) * fw.Dw[q] + fw.D_r[q]) * k.inv_rho : fw.Dw[q] * k.inv_rho;
        const double c = cbrt(X);
        db += nb * (k.nu0 + 2 * k.M * d * c);
        M_b += nb * d * d * c;
    } else {
        const double c = cbrt(fw.Dw[q] * (double)k.inv_rho_f);
        db += nb * (k.nu0 + k.M * fw.H[q] * c);
        M_b += nb * d * fw.H[q] * c;
    }
    Dbar[q] += db;
}

}  // namespace m70
