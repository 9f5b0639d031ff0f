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
#include "assembly.h"
#include "tails.h"

namespace m70 {

#define CELL_IJ(nrows)                                   \
    const int j = blockIdx.x * blockDim.x + threadIdx.x; \
    const int i = blockIdx.y;                            \
    if (j >= k.ny || i >= (nrows)) return;               \
    const int ny = k.ny;                                 \
    const long r = (long)i * ny;

#define CELL_IJ_ONLY(nrows)                              \
    const int j = blockIdx.x * blockDim.x + threadIdx.x; \
    const int i = blockIdx.y;                            \
    if (j >= k.ny || i >= (nrows)) return;

// ------------------------------------------------------------------------------------
// Per-Run fields, src/Hydro.cpp:175-192
// ------------------------------------------------------------------------------------

// Hydro::calculateDragAndViscosity with the model functions,
// src/HydroParametrizations.cpp:30-113. Rows 0..nx-1.
__host__ __device__ __forceinline__ void drag_viscosity_body(const Consts &k, const double *__restrict__ D, const Forcing &fw,
                                                    const RunFields &rf, int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    (void)r;
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

__global__ void __launch_bounds__(256) k_drag_viscosity(Consts k, const double *__restrict__ D, Forcing fw, RunFields rf)
{
    CELL_IJ_ONLY(k.nx)
    drag_viscosity_body(k, D, fw, rf, i, j);
}

// Forcing, face averages and viscosity derivatives. Rows 0..nx-1 (x-face fields only
// for rows 0..nx-2).
__host__ __device__ __forceinline__ void run_fields_body(const Consts &k, const Metrics &m, const Forcing &fw, const RunFields &rf,
                                                int i, int j)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    (void)r;
    const int nx = k.nx;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny);
    const double *nuD = rf.nuD;
    // every value is formed before the first store: a store ahead of a load would hold it
    // (the output arrays may alias the inputs as far as the compiler knows)
    const bool xrow = i < nx - 1;
    double Fx = 0.0, Urms_x = 0.0, cD_x = 0.0, nuD_x = 0.0, dnuD_dx = 0.0, dnuDx_dy = 0.0;
    if (xrow) {
        const double *Sxy = fw.Sxy + r, *Sxy_up = Sxy + ny, *Sxx = fw.Sxx + r, *Sxx_up = Sxx + ny;
        // getForcing_x, src/HydroQx.cpp:371-386 (j = 0 and ny-1 wrap, :378-379)
        Fx = -(k.inv_4dy_rho * (Sxy[jp] - Sxy[jm] + Sxy_up[jp] - Sxy_up[jm]) +
                         (Sxx_up[j] - Sxx[j]) * m.inv_rhodx[i]);
        // C2X (include/MeshOperations.h:18-27) of Urms, cD, nuD: src/Hydro.cpp:177,182,186
        Urms_x = 0.5 * (fw.Urms[r + j] + fw.Urms[r + ny + j]);
        cD_x = 0.5 * (rf.cD[r + j] + rf.cD[r + ny + j]);
        nuD_x = 0.5 * (nuD[r + j] + nuD[r + ny + j]);
        // difC2X(dnuD_dx, nuD, dx), src/HydroMeshOperations.cpp:99-114
        dnuD_dx = m.inv_dx[i] * (nuD[r + ny + j] - nuD[r + j]);
        // difC2C_y(dnuDx_dy, nuD_x), src/HydroMeshOperations.cpp:202-221
        const double nux_p = 0.5 * (nuD[r + jp] + nuD[r + ny + jp]);
        const double nux_m = 0.5 * (nuD[r + jm] + nuD[r + ny + jm]);
        dnuDx_dy = k.inv_2dy_c2c * (nux_p - nux_m);
    }
    double Fy;
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
        Fy = val;
    }
    // C2Y (src/HydroMeshOperations.cpp:47-61) of Urms, cD, nuD: src/Hydro.cpp:180,184,187
    const double Urms_y = 0.5 * (fw.Urms[r + jp] + fw.Urms[r + j]);
    const double cD_y = 0.5 * (rf.cD[r + jp] + rf.cD[r + j]);
    const double nuy = 0.5 * (nuD[r + jp] + nuD[r + j]);
    // difC2Y(dnuD_dy, nuD), src/HydroMeshOperations.cpp:116-130
    const double dnuD_dy = (nuD[r + jp] - nuD[r + j]) * k.inv_dy;
    // difC2C_x_addBoundary(dnuDy_dx, nuD_y), src/HydroMeshOperations.cpp:161-200
    double out;
    {
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
    }
    if (xrow) {
        rf.Fx[r + j] = Fx;
        rf.Urms_x[r + j] = Urms_x;
        rf.cD_x[r + j] = cD_x;
        rf.nuD_x[r + j] = nuD_x;
        rf.dnuD_dx[r + j] = dnuD_dx;
        rf.dnuDx_dy[r + j] = dnuDx_dy;
    }
    rf.Fy[r + j] = Fy;
    rf.Urms_y[r + j] = Urms_y;
    rf.cD_y[r + j] = cD_y;
    rf.nuD_y[r + j] = nuy;
    rf.dnuD_dy[r + j] = dnuD_dy;
    rf.dnuDy_dx[r + j] = out;
}

__global__ void __launch_bounds__(256) k_run_fields(Consts k, Metrics m, Forcing fw, RunFields rf)
{
    CELL_IJ_ONLY(k.nx)
    run_fields_body(k, m, fw, rf, i, j);
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
    asm_qxx(k, m, s, rf, w, f0, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
}

// ------------------------------------------------------------------------------------
// Qx projector assembly: setSolverQxy_f/_d/_up/_low, src/HydroQx.cpp:262-349. Rows
// 0..nx-2. Reads the predictor's Qx* (Qx rows 1..nx-1), IsWater and DragTerm.
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_qxy_assemble(Consts k, State s, RunFields rf, Work w)
{
    CELL_IJ(k.nx - 1)
    asm_qxy(k, s, rf, w, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
}

// ------------------------------------------------------------------------------------
// After the Qx projector: setQx_BC (src/HydroBoundaryConditions.cpp:13-23), difC2X of Qx
// (src/HydroQx.cpp:80), solveZs_guess (src/Hydro.cpp:230-250), getU0/getU
// (src/HydroQx.cpp:420-451), X2C (include/MeshOperations.h:36-47), and the Qy
// predictor's IsWater (setWaterNodes_y, src/HydroQy.cpp:92-122) into w.w. Rows 0..nx-1.
// guess_Zs enters only that mask and is not stored.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_qx_tail(Consts k, Metrics m, State s, Work w, double f0, int write_uc,
                                                 int rev)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = rev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;   // rev: rows last to first
    if (j >= k.ny || i >= k.nx) return;
    qx_tail_body(k, m, s, w, f0, i, j, write_uc != 0);
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
    asm_qyy(k, m, s, rf, w, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
}

// ------------------------------------------------------------------------------------
// Qy projector assembly: setSolverQyx_f/_d/_up/_low/0_low (src/HydroQy.cpp:235-362) and
// setSolverQyx_BC (src/HydroBoundaryConditions.cpp:41-60). Rows 0..nx-2.
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_qyx_assemble(Consts k, Metrics m, State s, RunFields rf,
                                                      Work w)
{
    CELL_IJ(k.nx - 1)
    asm_qyx(k, m, s, rf, w, i, j, w.a[r + j], w.b[r + j], w.c[r + j], w.f[r + j]);
}

// ------------------------------------------------------------------------------------
// After the Qy projector: setQy_BC (src/HydroBoundaryConditions.cpp:62-79), difY2C
// (src/HydroMeshOperations.cpp:148-159), solveZs (src/Hydro.cpp:253-271), getV/getV0
// (src/HydroQy.cpp:458-487), Y2C (src/HydroMeshOperations.cpp:88-97). Rows 0..nx-1.
// Reads D, writes the new depth to Dn.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_qy_tail(Consts k, State s, int write_vc, int rev)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = rev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;   // rev: rows last to first
    if (j >= k.ny || i >= k.nx) return;
    qy_tail_body(k, s, i, j, write_vc != 0);
}

// ------------------------------------------------------------------------------------
// The two tails on two adjacent columns per thread (even ny), with 16-byte loads of
// column pairs: the columns a thread needs either side come from the neighbouring pairs
// instead of one load per neighbour and column. Every value is formed by the same
// expressions as qx_tail_body / qy_tail_body, so the results are the same bits.
// ------------------------------------------------------------------------------------
template <bool NT = false>
__device__ __forceinline__ double2 ld2(const double *p, long e)
{
    double2 v;
    ldn2<NT>(p + e, v.x, v.y);
    return v;
}
template <bool NT = false>
__device__ __forceinline__ void st2(double *p, long e, double a, double b)
{
    stn2<NT>(p + e, a, b);
}

// Columns j0, j0 + 1 of row i (j0 even, ny even).
template <bool STR = false>
__device__ __forceinline__ void qy_tail2_body(const Consts &k, const State &s, int i, int j0, bool write_vc = true)
{
    constexpr bool NTL = Nt<NT_QY_TAIL, STR>::L, NTS = Nt<NT_QY_TAIL, STR>::S;
    const int ny = k.ny, nx = k.nx;
    const long r = (long)i * ny;
    const int cm = (j0 == 0) ? ny - 2 : j0 - 2, cp = (j0 + 2 == ny) ? 0 : j0 + 2;
    // qy_row(i+1, .) at columns cm, cm+1 | j0, j0+1 | cp (row nx is Qy[nx-1]*inv_3)
    auto qyr = [&](int c) -> double2 {
        if (i + 1 < nx) return ld2<NTL>(s.Qy, (long)(i + 1) * ny + c);
        const double2 v = ld2<NTL>(s.Qy, (long)(nx - 1) * ny + c);
        return make_double2(v.x * k.inv_3, v.y * k.inv_3);
    };
    const double2 qm = qyr(cm), q0 = qyr(j0), qp = qyr(cp);
    const double2 xm = ld2<NTL>(s.dQx, r + cm), x0 = ld2<NTL>(s.dQx, r + j0), xp = ld2<NTL>(s.dQx, r + cp);
    const double2 dm = ld2<NTL>(s.D, r + cm), d0 = ld2<NTL>(s.D, r + j0), dp = ld2<NTL>(s.D, r + cp);
    // columns j0-1 (m1), j0 (a), j0+1 (b), j0+2 (p1)
    const double dq_m1 = (qm.y - qm.x) * k.inv_dy;
    const double dq_a = (q0.x - qm.y) * k.inv_dy;
    const double dq_b = (q0.y - q0.x) * k.inv_dy;
    const double dq_p1 = (qp.x - q0.y) * k.inv_dy;
    const double inc_a = -k.dt * (x0.x + dq_a), inc_b = -k.dt * (x0.y + dq_b);
    const double dn_m1 = dm.y + -k.dt * (xm.y + dq_m1);
    const double dn_a = d0.x + inc_a, dn_b = d0.y + inc_b;
    const double dn_p1 = dp.x + -k.dt * (xp.x + dq_p1);
    const double v_m1 = qm.y / (0.5 * (dn_m1 + dn_a));
    const double v_a = q0.x / (0.5 * (dn_a + dn_b));
    const double v_b = q0.y / (0.5 * (dn_b + dn_p1));
    st2<NTS>(s.dQy, r + j0, dq_a, dq_b);
    const double2 zs = ld2<NTL>(s.Zs, r + j0);
    st2<NTS>(s.Zs, r + j0, zs.x + inc_a, zs.y + inc_b);
    st2<NTS>(s.Dn, r + j0, dn_a, dn_b);
    st2<NTS>(s.V, r + ny + j0, v_a, v_b);
    if (write_vc) st2<NTS>(s.Vc, r + j0, 0.5 * (v_a + v_m1), 0.5 * (v_b + v_a));
    if (i == 0) {
        const double2 q1 = ld2<NTL>(s.Qy, ny + j0);
        const double qa = k.bcQy * q1.x, qb = k.bcQy * q1.y;
        st2<NTS>(s.Qy, j0, qa, qb);
        st2<NTS>(s.V, j0, qa / (2 * dn_a - new_depth(s, k, 1, j0)), qb / (2 * dn_b - new_depth(s, k, 1, j0 + 1)));
    }
    if (i == nx - 1) {
        const double2 ql = ld2<NTL>(s.Qy, (long)(nx - 1) * ny + j0);
        st2<NTS>(s.Qy, (long)nx * ny + j0, ql.x * k.inv_3, ql.y * k.inv_3);
    }
}

// Rows i0..i1-1 of columns j0, j0 + 1 (j0 even, ny even): the Qx tail over a run of rows,
// with the Qx rows i-1..i+2 and the guess depths of rows i-1..i+1 held in registers from
// one row to the next, so each row of Qx, D and dQy_dy is loaded once per run instead of
// four or three times. The same expressions on the same values: the same bits. Lane l of a
// wavefront must hold the pair after lane l-1's (or lie past the row's last pair) over the
// same rows, and every lane that holds a pair must call it (cross-lane reads).
template <bool STR = false>
__device__ __forceinline__ void qx_tail2_rows(const Consts &k, const Metrics &m, const State &s, const Work &w,
                                              double f0, int i0, int i1, int j0, bool write_uc)
{
    constexpr bool NTL = Nt<NT_QX_TAIL, STR>::L, NTS = Nt<NT_QX_TAIL, STR>::S;
    const int ny = k.ny, nx = k.nx;
    const int c2 = (j0 + 2 == ny) ? 0 : j0 + 2;
    const double dt = k.dt;
    auto qrow = [&](int row) -> double2 {   // Qx row `row` at (j0, j0+1), through the row-0 BC
        if (row < 0 || row > nx) return make_double2(0.0, 0.0);
        if (row == 0) {
            const double2 v = ld2<NTL>(s.Qx, ny + j0);
            return make_double2(k.f1 * v.x + f0, k.f1 * v.y + f0);
        }
        return ld2<NTL>(s.Qx, (long)row * ny + j0);
    };
    auto qrow_c2 = [&](int row) -> double {   // the same at column j0+2
        if (row < 0 || row > nx) return 0.0;
        if (row == 0) return k.f1 * ldn<NTL>(s.Qx + ny + c2) + f0;
        return ldn<NTL>(s.Qx + (long)row * ny + c2);
    };
    struct G { double2 gd, dq, inc, dqy; };   // guess_D, dQx_dx, its depth increment, dQy_dy
    auto grow = [&](int row, double2 lo, double2 hi) -> G {
        G g;
        g.gd = g.dq = g.inc = g.dqy = make_double2(0.0, 0.0);
        if (row < 0 || row > nx - 1) return g;
        const double idx = m.inv_dxx[row];
        const double2 dqy = ld2<NTL>(s.dQy, (long)row * ny + j0), dd = ld2<NTL>(s.D, (long)row * ny + j0);
        const double dqa = idx * (hi.x - lo.x), dqb = idx * (hi.y - lo.y);
        const double ia = -dt * (dqa + dqy.x), ib = -dt * (dqb + dqy.y);
        g.gd = make_double2(dd.x + ia, dd.y + ib);
        g.dq = make_double2(dqa, dqb);
        g.inc = make_double2(ia, ib);
        g.dqy = dqy;
        return g;
    };
    double2 qm = qrow(i0 - 1), q0 = qrow(i0), q1 = qrow(i0 + 1), q2 = qrow(i0 + 2);
    G gm = grow(i0 - 1, qm, q0), g0 = grow(i0, q0, q1), g1 = grow(i0 + 1, q1, q2);
    // Column j0+2 is lane l+1's j0 (a wavefront's lanes take consecutive pairs of a row run):
    // its guess depth, guess level and dQy_dy come across lanes, the same expressions on the
    // same values; only a lane whose right neighbour is in another wavefront or across the
    // periodic seam loads what they are formed from.
    const int lane = (int)(threadIdx.x & 63);
    const bool own_p = lane < 63 && j0 + 2 < ny;
    double cqa = 0.0, cqb = 0.0;
    if (!own_p) { cqa = qrow_c2(i0); cqb = qrow_c2(i0 + 1); }
    for (int i = i0; i < i1; ++i) {
        const long r = (long)i * ny;
        double ua, ub;
        if (i == 0) {
            ua = q0.x / (1.5 * g0.gd.x - 0.5 * g1.gd.x);
            ub = q0.y / (1.5 * g0.gd.y - 0.5 * g1.gd.y);
        } else {
            ua = q0.x / (0.5 * (gm.gd.x + g0.gd.x));
            ub = q0.y / (0.5 * (gm.gd.y + g0.gd.y));
        }
        double u1a, u1b;
        if (i + 1 <= nx - 1) {
            u1a = q1.x / (0.5 * (g0.gd.x + g1.gd.x));
            u1b = q1.y / (0.5 * (g0.gd.y + g1.gd.y));
        } else {
            const double2 un = ld2<NTL>(s.U, (long)nx * ny + j0);   // U row nx is never written by the reference
            u1a = un.x;
            u1b = un.y;
        }
        const double2 zs = ld2<NTL>(s.Zs, r + j0);
        double e_dqy = 0.0, e_d = 0.0, e_zs = 0.0;
        if (!own_p) { e_dqy = ldn<NTL>(s.dQy + r + c2); e_d = ldn<NTL>(s.D + r + c2); e_zs = ldn<NTL>(s.Zs + r + c2); }
        const double gza = zs.x + g0.inc.x, gzb = zs.y + g0.inc.y;
        // guess_D / guess_Zs of column j0+2 (the right neighbour of j0+1); every lane takes
        // part in the cross-lane reads
        double gd_c2 = __shfl_down(g0.gd.x, 1), gz_c2 = __shfl_down(gza, 1), dqy_c2 = __shfl_down(g0.dqy.x, 1);
        if (!own_p) {
            const double dq = m.inv_dxx[i] * (cqb - cqa);
            dqy_c2 = e_dqy;
            const double incrZs = -dt * (dq + dqy_c2);
            gd_c2 = e_d + incrZs;
            gz_c2 = e_zs + incrZs;
        }
        const int wa = wet_rule(g0.gd.x, g0.gd.y, gza, gzb, g0.dqy.x, g0.dqy.y, k.Dmask);
        const int wb = wet_rule(g0.gd.y, gd_c2, gzb, gz_c2, g0.dqy.y, dqy_c2, k.Dmask);
        stn<NTS>(w.w + r + j0, (uint8_t)wa);
        stn<NTS>(w.w + r + j0 + 1, (uint8_t)wb);
        st2<NTS>(s.dQx, r + j0, g0.dq.x, g0.dq.y);
        st2<NTS>(w.gD, r + j0, g0.gd.x, g0.gd.y);
        st2<NTS>(s.U, r + j0, ua, ub);
        if (write_uc) st2<NTS>(s.Uc, r + j0, m.x2c_d[i] * ua + m.x2c_up[i] * u1a, m.x2c_d[i] * ub + m.x2c_up[i] * u1b);
        if (i == 0) st2<NTS>(s.Qx, j0, q0.x, q0.y);
        if (i + 1 < i1) {
            const double2 q3 = qrow(i + 3);
            gm = g0;
            g0 = g1;
            g1 = grow(i + 2, q2, q3);
            qm = q0; q0 = q1; q1 = q2; q2 = q3;
            if (!own_p) {
                cqa = cqb;
                cqb = qrow_c2(i + 2);
            }
        }
    }
}

// Row i alone (the batched step's tails, one row per thread).
template <bool STR = false>
__device__ __forceinline__ void qx_tail2_body(const Consts &k, const Metrics &m, const State &s, const Work &w,
                                              double f0, int i, int j0, bool write_uc = true)
{
    qx_tail2_rows<STR>(k, m, s, w, f0, i, i + 1, j0, write_uc);
}

// `rows` rows per thread, two columns (ny even): grid (ceil(ny/512), ceil(nx/rows)).
template <bool STR>
__global__ void __launch_bounds__(256) k_qx_tail2r(Consts k, Metrics m, State s, Work w, double f0, int write_uc,
                                                   int rows, int rev)
{
    const int j0 = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i0 = (rev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y) * rows;   // rev: row runs last to first
    if (j0 >= k.ny || i0 >= k.nx) return;
    const int i1 = i0 + rows < k.nx ? i0 + rows : k.nx;
    qx_tail2_rows<STR>(k, m, s, w, f0, i0, i1, j0, write_uc != 0);
}

// Rows per thread of k_qx_tail2r: up to 8, while the launch keeps at least 8192 workgroups
// (32 per CU); small grids take one row per thread. Measured at 2048 x 4096 (ms per step):
// 2 rows 0.977, 4 rows 0.991, 8 rows 0.986.
static int qx_tail_rows(const Consts &k)
{
    const long bx = (k.ny / 2 + 255) / 256;
    long r = (long)k.nx * bx / 8192;
    return r < 1 ? 1 : (r > 8 ? 8 : (int)r);
}


template <bool STR>
__global__ void __launch_bounds__(256) k_qy_tail2(Consts k, State s, int write_vc, int rev)
{
    const int j0 = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = rev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;   // rev: rows last to first
    if (j0 >= k.ny || i >= k.nx) return;
    qy_tail2_body<STR>(k, s, i, j0, write_vc != 0);
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

__global__ void __launch_bounds__(64) k_flag_clear(int *flag)
{
    if (threadIdx.x == 0) *flag = 0;
}

// Zs = D - Z + h, src/Hydro.cpp:35-37.
__global__ void __launch_bounds__(256) k_init_zs(double *Zs, const double *D, const double *Z, const double *hb,
                                                 long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) Zs[q] = D[q] - Z[q] + hb[q];
}

// ------------------------------------------------------------------------------------
// Grouped cell kernels: one launch covers every scenario of a batch. Block b belongs to
// scenario scen_of(start, nscen, b); within it blocks tile (ceil(ny/256) x nx).
// ------------------------------------------------------------------------------------
template <int OP>
__global__ void __launch_bounds__(256) k_cells_grouped(cScenDesc *D, const int *__restrict__ start,
                                                       int nscen)
{
    const int b = (int)blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const ScenDesc &d = *(const ScenDesc *)(D + sid);
    const int ny = d.k.ny;
    const int lb = b - start[sid];
    const int gx = (ny + 255) / 256;
    const int i = lb / gx;
    const int j = (lb - i * gx) * 256 + threadIdx.x;
    if (j >= ny || i >= d.k.nx) return;
    if (OP == 0) drag_viscosity_body(d.k, d.s.D, d.fw, d.rf, i, j);
    else run_fields_body(d.k, d.m, d.fw, d.rf, i, j);
}

int grouped_cell_blocks(int nx, int ny) { return ((ny + 255) / 256) * nx; }

hipError_t launch_cells_grouped(hipStream_t st, int op, const ScenDesc *D, const int *start, int nscen, int blocks)
{
    if (op == 0) hipLaunchKernelGGL(k_cells_grouped<0>, dim3(blocks), dim3(256), 0, st, (cScenDesc *)D, start, nscen);
    else hipLaunchKernelGGL(k_cells_grouped<1>, dim3(blocks), dim3(256), 0, st, (cScenDesc *)D, start, nscen);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Grouped tails (OP 2: Qx, OP 3: Qy). A scenario's work items are numbered in one line and
// taken 256 to a workgroup, so short rows leave no lanes idle: with an even ny an item is
// a column pair (qx_tail2_rows / qy_tail2_body), over GT_ROWS rows for the Qx tail; with an
// odd ny it is one cell. WD: the tails store U_c (OP 2) and V_c (OP 3).
// ------------------------------------------------------------------------------------
constexpr int GT_ROWS = 2;

static long grouped_tail_items(int nx, int ny, int op)
{
    if (ny % 2 == 0 && ny >= 4) return op == 2 ? (long)((nx + GT_ROWS - 1) / GT_ROWS) * (ny / 2) : (long)nx * (ny / 2);
    return (long)nx * ny;
}
int grouped_tail_blocks(int nx, int ny, int op) { return (int)((grouped_tail_items(nx, ny, op) + 255) / 256); }

template <int OP, bool WD>
__global__ void __launch_bounds__(256) k_tails_grouped(cScenDesc *D, const int *__restrict__ start,
                                                       int nscen, int rev)
{
    // rev: the last scenario's last rows first
    const int b = rev ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;
    // the Qy tail ends a step: count it for the offshore offset series (scen_f0); every
    // reader of the counter is in a later launch
    if (OP == 3 && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((int *)((const ScenDesc *)D)->step, 1);
    const int sid = scen_of(start, nscen, b);
    const ScenDesc &d = *(const ScenDesc *)(D + sid);
    const int nx = d.k.nx, ny = d.k.ny;
    const long q = (long)(b - start[sid]) * 256 + threadIdx.x;
    if (ny % 2 == 0 && ny >= 4) {
        const int np = ny / 2;
        const long r = q / np;
        const int j0 = 2 * (int)(q - r * np);
        if (OP == 2) {
            if (r * GT_ROWS >= nx) return;
            const int i0 = (int)r * GT_ROWS, i1 = i0 + GT_ROWS < nx ? i0 + GT_ROWS : nx;
            qx_tail2_rows<GROUPED_STREAM>(d.k, d.m, d.s, d.w, scen_f0(d), i0, i1, j0, WD);
        } else {
            if (r >= nx) return;
            qy_tail2_body<GROUPED_STREAM>(d.k, d.s, (int)r, j0, WD);
        }
        return;
    }
    const long i = q / ny;
    const int j = (int)(q - i * ny);
    if (i >= nx) return;
    if (OP == 2) qx_tail_body(d.k, d.m, d.s, d.w, scen_f0(d), (int)i, j, WD);
    else qy_tail_body(d.k, d.s, (int)i, j, WD);
}

hipError_t launch_tails_grouped(hipStream_t st, int op, const ScenDesc *D, const int *start, int nscen, int blocks,
                                int rev, bool write_derived)
{
    const dim3 g((unsigned)blocks), b(256);
    if (op == 2) {
        if (write_derived) hipLaunchKernelGGL((k_tails_grouped<2, true>), g, b, 0, st, (cScenDesc *)D, start, nscen, rev);
        else hipLaunchKernelGGL((k_tails_grouped<2, false>), g, b, 0, st, (cScenDesc *)D, start, nscen, rev);
    } else {
        if (write_derived) hipLaunchKernelGGL((k_tails_grouped<3, true>), g, b, 0, st, (cScenDesc *)D, start, nscen, rev);
        else hipLaunchKernelGGL((k_tails_grouped<3, false>), g, b, 0, st, (cScenDesc *)D, start, nscen, rev);
    }
    return hipGetLastError();
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
                          double f0, bool write_uc, int rev, bool stream)
{
    if (k.ny % 2 == 0 && k.ny >= 4) {
        const int rows = qx_tail_rows(k);
        hipLaunchKernelGGL(stream ? k_qx_tail2r<true> : k_qx_tail2r<false>, cell_grid(k.ny / 2, (k.nx + rows - 1) / rows),
                           dim3(256), 0, st, k, m, s, w, f0, (int)write_uc, rows, rev);
    } else {
        hipLaunchKernelGGL(k_qx_tail, cell_grid(k.ny, k.nx), dim3(256), 0, st, k, m, s, w, f0, (int)write_uc, rev);
    }
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
hipError_t launch_qy_tail(hipStream_t st, const Consts &k, const State &s, bool write_vc, int rev, bool stream)
{
    if (k.ny % 2 == 0 && k.ny >= 4)
        hipLaunchKernelGGL(stream ? k_qy_tail2<true> : k_qy_tail2<false>, cell_grid(k.ny / 2, k.nx), dim3(256), 0, st, k,
                           s, (int)write_vc, rev);
    else
        hipLaunchKernelGGL(k_qy_tail, cell_grid(k.ny, k.nx), dim3(256), 0, st, k, s, (int)write_vc, rev);
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

hipError_t launch_flag_clear(hipStream_t st, int *flag)
{
    hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, st, flag);
    return hipGetLastError();
}

hipError_t launch_init_zs(hipStream_t st, double *Zs, const double *D, const double *Z, const double *hb, long n)
{
    hipLaunchKernelGGL(k_init_zs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Zs, D, Z, hb, n);
    return hipGetLastError();
}

}  // namespace m70
