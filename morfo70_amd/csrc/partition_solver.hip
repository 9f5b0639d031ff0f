// partition_solver.hip -- workgroup-parallel tridiagonal line solver for gfx950.
//
// Solves many independent tridiagonal systems (periodic or not) of n unknowns each,
//     b_k x_{k-1} + a_k x_k + c_k x_{k+1} = f_k,   k = 0..n-1,
// with x_{-1} = x_{n-1} and x_n = x_0 when periodic (corners b_0 and c_{n-1}, the layout
// of Math::TridiagonalSolverPeriodic, src/Math.cpp:68-118) and b_0 = c_{n-1} = 0 otherwise
// (Math::TridiagonalSolver, src/Math.cpp:15-65).
//
// Each line is split into S segments (S a power of two), one per thread:
//   1. Two-sided elimination inside the segment, in registers: the downward pass removes
//      the sub-diagonal and leaves a fill-in column on the segment's first unknown, the
//      upward pass removes the super-diagonal and leaves one on its last unknown.
//   2. The first/last unknowns of all segments form a cyclic tridiagonal system of
//      N = 2S rows, solved by parallel cyclic reduction in LDS: log2(N)-1 reduction steps
//      leave rows i and i+N/2 coupled only to each other, then a 2x2 solve.
//   3. Interior unknowns follow from the two boundary values.
// The matrices are diagonally dominant (the 1/dt term), so no pivoting is needed. The
// rounding differs from the serial recurrences of the reference; results agree to a few
// ulps per solve.
#include <type_traits>

#include "adjoint_batch.h"
#include "assembly.h"
#include "hydro_device.h"
#include "hydro_kernels.h"

namespace m70 {


struct LineGeom {
    long ld;        // COLS: row stride of the arrays; ROWS: line stride
    long ldy;       // output stride (as ld, for Y)
    int n;          // unknowns per line
    int nL;         // number of lines
    int S;          // segments per line (power of two, n >= 2S, n <= SEG_M*S)
    int log2S;
    int lines_per_block;
    int periodic;
    int xcd_remap;  // remap block index so neighbouring tiles share an XCD's L2
};

template <bool COLS>
__device__ __forceinline__ long elem_index(const LineGeom &g, int line, int k, long ld)
{
    return COLS ? (long)k * ld + line : (long)line * ld + k;
}

// Coefficient providers: get(line, k, a, b, c, f) yields element k of line `line`.
template <bool COLS, int SITE = NT_NONE, bool STR = false>
struct ProvArrays {   // coefficients stored in four arrays
    static constexpr int nt_site = SITE;
    static constexpr bool nt_stream = STR;
    const double *A, *B, *C, *F;
    long ld;
    __device__ __forceinline__ void get(int line, int k, double &a, double &b, double &c, double &f) const
    {
        constexpr bool NTL = Nt<SITE, STR>::L;
        const long e = COLS ? (long)k * ld + line : (long)line * ld + k;
        a = ldn<NTL>(A + e); b = ldn<NTL>(B + e); c = ldn<NTL>(C + e); f = ldn<NTL>(F + e);
    }
};
// The cache-policy site of a coefficient provider (NtSite, NT_NONE if it names none) and
// whether it runs under the streaming policy.
template <class P, class = void> struct ProvSite { static constexpr int value = NT_NONE; };
template <class P> struct ProvSite<P, std::void_t<decltype(P::nt_site)>> { static constexpr int value = P::nt_site; };
template <class P, class = void> struct ProvStream { static constexpr bool value = false; };
template <class P> struct ProvStream<P, std::void_t<decltype(P::nt_stream)>> { static constexpr bool value = P::nt_stream; };
// The store policy of a provider's kernel.
template <class P> constexpr bool prov_nts() { return Nt<ProvSite<P>::value, ProvStream<P>::value>::S; }
// Row providers with get2() also yield elements kk and kk + 1 together (kk even, even
// line length) from loads of column pairs.
__device__ __forceinline__ void st_pair(double *p, long e, double a, double b)
{
    *reinterpret_cast<double2 *>(p + e) = make_double2(a, b);
}

// Qx projector rows with Qx* (the predictor's solution, row i+1 = unknown i of each column)
// formed from the predictor's three-pass column solve instead of read back after a
// back-substitution pass: the FUSED step skips that pass and never stores Qx*.
// V_c row i at column c is 0.5*(V[i+1][c] + V[i+1][c-1]) (periodic in c), the Qy tail's own
// expression for it. With VCF a sweep forms the V_c values it needs from V rows i+1, i+2 instead
// of reading V_c rows i, i+1: the same bits wherever V_c was stored by a step, which holds
// from a Run's second step on, so the steps between need not store V_c.
__device__ __forceinline__ void qxy_vc_from_v(const Consts &k, const State &s, int i, int j, QxyIn &q)
{
    const int ny = k.ny;
    const int jp = wrap_inc(j, ny), jm = wrap_dec(j, ny), jmm = wrap_dec(jm, ny);
    const double *v1 = s.V + (long)(i + 1) * ny, *v2 = v1 + ny;
    q.Vcp0 = 0.5 * (v1[jp] + v1[j]); q.Vcp1 = 0.5 * (v2[jp] + v2[j]);
    q.Vcm0 = 0.5 * (v1[jm] + v1[jmm]); q.Vcm1 = 0.5 * (v2[jm] + v2[jmm]);
}
// Elements e, e+1 (e even, ny even): V columns e-2..e+2 from three pair loads per row.
template <bool STR>
__device__ __forceinline__ void qxy_vc_from_v2(const Consts &k, const State &s, int i, int e, QxyIn &a, QxyIn &b)
{
    constexpr bool NTL = Nt<NT_QXY_ROWS, STR>::L;
    const int ny = k.ny;
    const int em = (e == 0) ? ny - 2 : e - 2, ep = (e + 2 == ny) ? 0 : e + 2;
    const long r1 = (long)(i + 1) * ny, r2 = r1 + ny;
    double x, p2, p1, p0, q1, q2;   // V row i+1 at columns e-2, e-1, e, e+1, e+2
    ld_pair<NTL>(s.V, r1 + em, p2, p1);
    ld_pair<NTL>(s.V, r1 + e, p0, q1);
    ld_pair<NTL>(s.V, r1 + ep, q2, x);
    double t2, t1, t0, u1, u2;      // V row i+2
    ld_pair<NTL>(s.V, r2 + em, t2, t1);
    ld_pair<NTL>(s.V, r2 + e, t0, u1);
    ld_pair<NTL>(s.V, r2 + ep, u2, x);
    // V_c at columns e-1 (m), e (0), e+1 (n), e+2 (p) of rows i and i+1
    const double c0m = 0.5 * (p1 + p2), c00 = 0.5 * (p0 + p1), c0n = 0.5 * (q1 + p0), c0p = 0.5 * (q2 + q1);
    const double c1m = 0.5 * (t1 + t2), c10 = 0.5 * (t0 + t1), c1n = 0.5 * (u1 + t0), c1p = 0.5 * (u2 + u1);
    a.Vcp0 = c0n; a.Vcp1 = c1n; a.Vcm0 = c0m; a.Vcm1 = c1m;
    b.Vcp0 = c0p; b.Vcp1 = c1p; b.Vcm0 = c00; b.Vcm1 = c10;
}

template <bool VCF, bool STR>
struct ProvQxyBs {
    static constexpr int nt_site = NT_QXY_ROWS;
    static constexpr bool nt_stream = STR;
    Consts k; State s; RunFields rf; Work w; ColsWork cw;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    {
        QxyIn q = qxy_load(k, s, rf, w, line, kk);
        if (VCF) qxy_vc_from_v(k, s, line, kk, q);
        q.qx = cols_value(cw, k.nx - 1, k.ny, line, kk);
        asm_qxy_vals(k, q, a, b, c, f);
    }
    __device__ __forceinline__ void get2(int line, int kk, double &a0, double &b0, double &c0, double &f0, double &a1,
                                         double &b1, double &c1, double &f1) const
    {
        QxyIn q0, q1;
        qxy_load2<STR>(k, s, rf, w, line, kk, q0, q1);
        if (VCF) qxy_vc_from_v2<STR>(k, s, line, kk, q0, q1);
        cols_value2<STR>(cw, k.nx - 1, k.ny, line, kk, q0.qx, q1.qx);
        asm_qxy_vals(k, q0, a0, b0, c0, f0);
        asm_qxy_vals(k, q1, a1, b1, c1, f1);
    }
};
constexpr int CM = 8;   // rows per column segment in the three-pass column solver

// Column segment i = k0..k0+mm-1 of the Qx predictor at column j (diagonal di, lower lo,
// upper up, right-hand side rh), with the rows the elements share held in a register
// window: each element loads only the rows no earlier element of the segment loaded
// (15 values instead of about 35). Rows of the nx-row fields past nx-1 are read only
// where asm_qxx_vals does not use them, so they are clamped to row nx-1.
// nuD_x, dnuD_dx and dnuDx_dy are formed from nuD rows i, i+1 at columns j, j+-1 by
// run_fields_body's expressions (one array in the window instead of three), and with the
// constant drag model (CDC) cD_x is cD0 itself (0.5*(cD0+cD0) is exact), not read.
template <bool CDC, bool VCF, bool STR>
__device__ __forceinline__ void qxx_segment(const Consts &k, const Metrics &mt, const State &s, const RunFields &rf,
                                            const Work &w, double f0, int j, int k0, int mm, double (&lo)[CM],
                                            double (&di)[CM], double (&up)[CM], double (&rh)[CM],
                                            bool cdc_rt = false)
{
    const int ny = k.ny, nx = k.nx;
    const int jm = wrap_dec(j, ny), jp = wrap_inc(j, ny);
    auto at = [&](int i) { return (long)i * ny + j; };
    auto atm = [&](int i) { return (long)i * ny + jm; };
    auto atp = [&](int i) { return (long)i * ny + jp; };
    auto cl = [&](int i) { return i < nx ? i : nx - 1; };
    constexpr bool NTL = Nt<NT_QXX, STR>::L, NTS = Nt<NT_QXX, STR>::S;
    constexpr bool NTF = NTL || Nt<NT_QXX_STREAM, STR>::L;   // single-use fields
    auto ld = [&](const double *f, long e) { return ldn<NTL>(f + e); };    // stencil-shared fields
    auto ldf = [&](const double *f, long e) { return ldn<NTF>(f + e); };   // single-use fields
    const double *nuD = rf.nuD;
    // cD0 in a register: selected against a load, a field of k (which may live in a constant
    // table) would make the select one of pointers in two address spaces, a flat load
    const double cD0 = k.cD0;
    double n0 = ld(nuD, at(k0)), n0p = ld(nuD, atp(k0)), n0m = ld(nuD, atm(k0));   // nuD row i
    QxxIn q;
    int wis[CM];
    double drags[CM];
    {
        const int i = k0, im = i > 0 ? i - 1 : 0, i2 = cl(i + 2);
        q.Dm = ld(s.D, at(im)); q.D0 = ld(s.D, at(i)); q.D1 = ld(s.D, at(i + 1)); q.D2 = ld(s.D, at(i2));
        q.Zm = ld(s.Zs, at(im)); q.Z0 = ld(s.Zs, at(i)); q.Z1 = ld(s.Zs, at(i + 1)); q.Z2 = ld(s.Zs, at(i2));
        q.Qm = ld(s.dQx, at(im)); q.Q0 = ld(s.dQx, at(i)); q.Q1 = ld(s.dQx, at(i + 1)); q.Q2 = ld(s.dQx, at(i2));
        if (!VCF) { q.Vc0 = ld(s.Vc, at(i)); q.Vc1 = ld(s.Vc, at(i + 1)); }
        q.U0 = ld(s.U, at(i)); q.U1 = ld(s.U, at(i + 1)); q.U2 = ld(s.U, at(i + 2));
        q.dqy0 = ldf(s.dQy, at(i)); q.dqy1 = ldf(s.dQy, at(i + 1));
        q.v1j = ld(s.V, at(i + 1)); q.v1m = ld(s.V, atm(i + 1)); q.v2j = ld(s.V, at(i + 2)); q.v2m = ld(s.V, atm(i + 2));
    }
#pragma unroll
    for (int t = 0; t < CM; ++t) {
        if (t < mm) {
            const int i = k0 + t;
            if (t > 0) {
                const int i2 = cl(i + 2);
                q.Dm = q.D0; q.D0 = q.D1; q.D1 = q.D2; q.D2 = ld(s.D, at(i2));
                q.Zm = q.Z0; q.Z0 = q.Z1; q.Z1 = q.Z2; q.Z2 = ld(s.Zs, at(i2));
                q.Qm = q.Q0; q.Q0 = q.Q1; q.Q1 = q.Q2; q.Q2 = ld(s.dQx, at(i2));
                if (!VCF) { q.Vc0 = q.Vc1; q.Vc1 = ld(s.Vc, at(i + 1)); }
                q.U0 = q.U1; q.U1 = q.U2; q.U2 = ld(s.U, at(i + 2));
                q.dqy0 = q.dqy1; q.dqy1 = ldf(s.dQy, at(i + 1));
                q.v1j = q.v2j; q.v1m = q.v2m; q.v2j = ld(s.V, at(i + 2)); q.v2m = ld(s.V, atm(i + 2));
            }
            if (VCF) {   // V_c rows i, i+1 from the V window (rows i+1, i+2 at j, j-1)
                q.Vc0 = 0.5 * (q.v1j + q.v1m);
                q.Vc1 = 0.5 * (q.v2j + q.v2m);
            }
            q.Qx1 = ldf(s.Qx, at(i + 1));
            q.cD = (CDC || cdc_rt) ? cD0 : ldf(rf.cD_x, at(i));
            q.Urms = ldf(rf.Urms_x, at(i)); q.Fx = ldf(rf.Fx, at(i));
            const double n1 = ld(nuD, at(i + 1)), n1p = ld(nuD, atp(i + 1)), n1m = ld(nuD, atm(i + 1));
            q.nuD_x = 0.5 * (n0 + n1);
            q.dnuD_dx = mt.inv_dx[i] * (n1 - n0);
            const double nux_p = 0.5 * (n0p + n1p), nux_m = 0.5 * (n0m + n1m);
            q.dnuDx_dy = k.inv_2dy_c2c * (nux_p - nux_m);
            n0 = n1; n0p = n1p; n0m = n1m;
            asm_qxx_vals(k, mt, q, f0, i, di[t], lo[t], up[t], rh[t], wis[t], drags[t]);
        }
    }
    // the side outputs go out after every load of the segment: stores in the loop would
    // keep later rows' loads (and the per-row metrics' scalar loads) behind them
#pragma unroll
    for (int t = 0; t < CM; ++t) {
        if (t < mm) {
            stn<NTS>(w.w + at(k0 + t), (uint8_t)wis[t]);
            stn<NTS>(w.drag + at(k0 + t), drags[t]);
        }
    }
}

template <bool CDC, bool VCF, bool STR>
struct ProvQxx {      // Qx predictor, lines along x (columns): line = j, k = i
    static constexpr int nt_site = NT_QXX;
    static constexpr bool nt_stream = STR;
    Consts k; Metrics m; State s; RunFields rf; Work w; double f0;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    { asm_qxx(k, m, s, rf, w, f0, kk, line, a, b, c, f); }
    __device__ __forceinline__ void segment(int line, int k0, int mm, double (&lo)[CM], double (&di)[CM],
                                            double (&up)[CM], double (&rh)[CM]) const
    { qxx_segment<CDC, VCF, STR>(k, m, s, rf, w, f0, line, k0, mm, lo, di, up, rh); }
};
// Column segment i = k0..k0+mm-1 of the Qy projector at column j (asm_qyx's arithmetic,
// src/HydroQy.cpp:235-362), with the rows the elements share held in a register window:
// element i reads guess_D rows i-1..i+1 at columns j and j+1, U rows i and i+1 and IsWater
// rows i-1..i+1, so each element loads only row i+1's (9 loads instead of 17), and the
// reciprocal of guess_D's face value of a row serves the three elements that use it. The
// same expressions on the same values: the same bits as asm_qyx.
template <bool STR>
__device__ __forceinline__ void qyx_segment(const Consts &k, const Metrics &mt, const State &s, const RunFields &rf,
                                            const Work &w, int j, int k0, int mm, double (&lo)[CM], double (&di)[CM],
                                            double (&up)[CM], double (&rh)[CM])
{
    constexpr bool NTL = Nt<NT_QYX, STR>::L;
    constexpr bool NTF = NTL || Nt<NT_QYX_STREAM, STR>::L;   // single-use fields
    const int ny = k.ny, nx = k.nx;
    const int jp = wrap_inc(j, ny);
    const double dt = k.dt;
    const double *gD = w.gD, *U = s.U;
    auto ld = [&](const double *f, long e) { return ldn<NTL>(f + e); };
    auto ldf = [&](const double *f, long e) { return ldn<NTF>(f + e); };
    auto inv_face = [&](int i) {   // inv_Di row i+1 (getD_y(i)) at column j
        const long r = (long)i * ny;
        return 1 / (0.5 * (ld(gD, r + jp) + ld(gD, r + j)));
    };
    // window: rows i-1 (m), i (0), i+1 (1)
    double invm = 0.0, inv0 = inv_face(k0), inv1 = inv_face(k0 + 1);
    int wm = k0 > 0 ? (int)ldn<NTL>(w.w + (long)(k0 - 1) * ny + j) : 0;
    int w0 = ldn<NTL>(w.w + (long)k0 * ny + j), w1 = ldn<NTL>(w.w + (long)(k0 + 1) * ny + j);
    if (k0 > 0) invm = inv_face(k0 - 1);
    const long r00 = (long)k0 * ny;
    double u0j = ld(U, r00 + j), u0p = ld(U, r00 + jp), u1j = ld(U, r00 + ny + j), u1p = ld(U, r00 + ny + jp);
#pragma unroll
    for (int t = 0; t < CM; ++t) {
        if (t < mm) {
            const int i = k0 + t;
            const long r = (long)i * ny;
            if (t > 0) {   // rows i-1..i+1 from the window, row i+1 loaded
                invm = inv0; inv0 = inv1; inv1 = inv_face(i + 1);
                wm = w0; w0 = w1; w1 = ldn<NTL>(w.w + r + ny + j);
                u0j = u1j; u0p = u1p; u1j = ld(U, r + ny + j); u1p = ld(U, r + ny + jp);
            }
            const int wi = w0;
            const double drag = ldf(w.drag, r + j);
            const double nud = ldf(rf.nuD_y, r + j), dnud_dx = ldf(rf.dnuDy_dx, r + j);
            // setSolverQyx_f (:235-248)
            const double fv = wi * ldf(s.Qy, r + ny + j) * (1 + dt * drag);
            // setSolverQyx_d (:250-285)
            const double x2c_d = mt.x2c_d[i], x2c_up = mt.x2c_up[i];
            double dv;
            {
                const double dx_d = mt.difx_c_d[i], d2x_d = mt.dif2x_c_d[i];
                const double inv_dx = mt.inv_dxx[i];
                const double inv_d = inv0;
                const int isl = !wi;
                const double du_dx = 0.5 * inv_dx * ((u1p + u1j) - (u0p + u0j));
                const double u_y = 0.5 * (x2c_up * (u1p + u1j) + x2c_d * (u0p + u0j));
                dv = wi * ((1 + dt * drag) + dt * (u_y * dx_d + du_dx) -
                           dt * (inv_d * nud * d2x_d + inv_d * dnud_dx * dx_d)) + isl;
            }
            // Ui_j row i (C2Y of U_c), U_c formed from U rows i, i+1 as the Qx tail forms it
            const double uc_j = x2c_d * u0j + x2c_up * u1j;
            const double uc_p = x2c_d * u0p + x2c_up * u1p;
            const double uy = 0.5 * (uc_p + uc_j);
            // setSolverQyx_up (:287-311); inv_Di row i+2 = getD_y(i+1); IsWater row nx-1 is false
            double uv;
            {
                const double inv_up = inv1;
                const int wu = (i + 1 <= nx - 2) ? w1 : 0;
                const double ww = wi && wu;
                const double dx_up = mt.difx_c_up[i], d2x_up = mt.dif2x_c_up[i];
                uv = ww * dt * (uy * dx_up - nud * d2x_up * inv_up - dnud_dx * dx_up * inv_up);
            }
            // setSolverQyx0_low (:340-362) / setSolverQyx_low (:313-338)
            double lv;
            {
                const double dx_low = mt.difx_c_low[i], d2x_low = mt.dif2x_c_low[i];
                if (i == 0) {
                    // getD_y0 (:426-440); the wrap element reads d[0] where the interior
                    // reads row 1 (:438)
                    double Dy0;
                    if (j < ny - 1)
                        Dy0 = 1 * (gD[j + 1] + gD[j]) - 0.5 * (gD[ny + j + 1] + gD[ny + j]);
                    else
                        Dy0 = 1 * (gD[0] + gD[ny - 1]) - 0.5 * (gD[0] + gD[2 * ny - 1]);
                    const double inv_low = 1 / Dy0;
                    lv = wi * dt * (uy * dx_low - nud * d2x_low * inv_low - dnud_dx * dx_low * inv_low);
                } else {
                    const double inv_low = invm;
                    const double ww = wi && wm;
                    lv = ww * dt * (uy * dx_low - nud * d2x_low * inv_low - dnud_dx * dx_low * inv_low);
                }
            }
            if (i == 0) dv += k.bcQy * lv;
            if (i == nx - 2) dv += uv * k.inv_3;
            di[t] = dv;
            lo[t] = lv;
            up[t] = uv;
            rh[t] = fv;
        }
    }
}

template <bool STR>
struct ProvQyx {      // Qy projector, columns
    static constexpr int nt_site = NT_QYX;
    static constexpr bool nt_stream = STR;
    Consts k; Metrics m; State s; RunFields rf; Work w;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    { asm_qyx<STR>(k, m, s, rf, w, kk, line, a, b, c, f); }
    __device__ __forceinline__ void segment(int line, int k0, int mm, double (&lo)[CM], double (&di)[CM],
                                            double (&up)[CM], double (&rh)[CM]) const
    { qyx_segment<STR>(k, m, s, rf, w, line, k0, mm, lo, di, up, rh); }
};
// nuD_y and dnuD_dy of elements e, e+1 of row i formed from nuD row i by run_fields_body's
// expressions, in place of qyy_load2's loads of the two arrays (which become dead).
template <bool STR>
__device__ __forceinline__ void qyy_nud_pair(const Consts &k, const RunFields &rf, int i, int e, QyyIn &a, QyyIn &b)
{
    constexpr bool NTL = Nt<NT_QYY_ROWS, STR>::L;
    const int ny = k.ny;
    const long r = (long)i * ny;
    const int ep = (e + 2 == ny) ? 0 : e + 2;
    double n0, n1, n2, x;
    ld_pair<NTL>(rf.nuD, r + e, n0, n1);
    ld_pair<NTL>(rf.nuD, r + ep, n2, x);
    a.nuD_y = 0.5 * (n1 + n0);
    a.dnuD_dy = (n1 - n0) * k.inv_dy;
    b.nuD_y = 0.5 * (n2 + n1);
    b.dnuD_dy = (n2 - n1) * k.inv_dy;
}

// CDC: constant drag model, cD_y is cD0 itself (0.5*(cD0+cD0) is exact) and is not read.
template <bool CDC, bool STR>
struct ProvQyy {      // Qy predictor, rows
    static constexpr int nt_site = NT_QYY_ROWS;
    static constexpr bool nt_stream = STR;
    Consts k; Metrics m; State s; RunFields rf; Work w;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    { asm_qyy(k, m, s, rf, w, line, kk, a, b, c, f); }
    __device__ __forceinline__ void get2(int line, int kk, double &a0, double &b0, double &c0, double &f0, double &a1,
                                         double &b1, double &c1, double &f1) const
    {
        QyyIn q0, q1;
        qyy_load2<STR>(k, s, rf, w, line, kk, q0, q1, !CDC);
        qyy_nud_pair<STR>(k, rf, line, kk, q0, q1);
        double d0, d1;
        asm_qyy_vals(k, m, q0, line, a0, b0, c0, f0, d0);
        asm_qyy_vals(k, m, q1, line, a1, b1, c1, f1, d1);
        stn2<Nt<NT_QYY_ROWS, STR>::S>(w.drag + (long)line * k.ny + kk, d0, d1);
    }
};

// ---------------------------------------------------------------------------------
// Building blocks shared by the partition kernels.
// ---------------------------------------------------------------------------------
struct RedRows {   // the two reduced rows of one segment
    double A_F, C_F, R_F, A_L, B_L, R_L, lo0, upL;
};

// Two-sided elimination of one segment of m (2..SEG_M) unknowns held in registers:
// lo/di/up/rh are the lower, diagonal, upper and right-hand side. On return lo[q] is the
// fill-in on the segment's first unknown, up[q] the fill-in on its last unknown, rh[q]
// the modified right-hand side and di[q] the inverse pivot of interior row q.
template <int SEG_M>
__device__ __forceinline__ RedRows seg_eliminate(int m, double (&lo)[SEG_M], double (&di)[SEG_M],
                                                 double (&up)[SEG_M], double (&rh)[SEG_M])
{
    RedRows r;
    // downward pass: row q -> s_q x_0 + d_q x_q + up_q x_{q+1} = g_q (q >= 1)
    r.lo0 = lo[0];
    double dLast = di[1];
    di[1] = 1.0 / di[1];
#pragma unroll
    for (int q = 2; q < SEG_M; ++q) {
        if (q < m) {
            const double tq = lo[q] * di[q - 1];
            lo[q] = -tq * lo[q - 1];
            const double dq = di[q] - tq * up[q - 1];
            rh[q] = rh[q] - tq * rh[q - 1];
            di[q] = 1.0 / dq;
            if (q == m - 1) dLast = dq;
        }
    }
    {   // row L (q = m-1): A_L x_0 + B_L x_{m-1} + upL x_m = R_L
        double sL = lo[1], gL = rh[1], uL = up[1];
#pragma unroll
        for (int q = 2; q < SEG_M; ++q)
            if (q == m - 1) { sL = lo[q]; gL = rh[q]; uL = up[q]; }
        r.A_L = sL; r.B_L = dLast; r.R_L = gL; r.upL = uL;
    }
    // upward pass: row q -> s'_q x_0 + d_q x_q + q_q x_{m-1} = g'_q, q = m-2..1
#pragma unroll
    for (int q = SEG_M - 2; q >= 1; --q) {
        if (q < m - 2) {
            const double tq = up[q] * di[q + 1];
            lo[q] = lo[q] - tq * lo[q + 1];
            up[q] = -tq * up[q + 1];
            rh[q] = rh[q] - tq * rh[q + 1];
        }
    }
    // row F (q = 0): lo0 x_{-1} + A_F x_0 + C_F x_{m-1} = R_F
    if (m >= 3) {
        const double tq = up[0] * di[1];
        r.A_F = di[0] - tq * lo[1];
        r.C_F = -tq * up[1];
        r.R_F = rh[0] - tq * rh[1];
    } else {
        r.A_F = di[0];
        r.C_F = up[0];
        r.R_F = rh[0];
    }
    return r;
}

// One cyclic-reduction step of a row (lower lo, diagonal dg, upper up, rhs rh) against
// its neighbours at -stride (m0..m3) and +stride (p0..p3), as (lower, diagonal, upper,
// rhs): eliminates the couplings to them, coupling the row to rows at twice the stride.
__device__ __forceinline__ void pcr_combine(double &lo, double &dg, double &up, double &rh, double m0, double m1,
                                            double m2, double m3, double p0, double p1, double p2, double p3)
{
    const double k1m = lo / m1;
    const double k2p = up / p1;
    const double na = -m0 * k1m;
    const double nc = -p2 * k2p;
    const double nb = dg - m2 * k1m - p0 * k2p;
    const double nd = rh - m3 * k1m - p3 * k2p;
    lo = na; dg = nb; up = nc; rh = nd;
}

// Last step of cyclic PCR on 2S rows: rows i and i+S are coupled only to each other,
//   bi x_i + ei x_j = di,  ej x_i + bj x_j = dj  (e = lower + upper);  returns x_i.
__device__ __forceinline__ double pcr_pair(double bi, double ei, double di, double bj, double ej, double dj)
{
    return (di * bj - ei * dj) / (bi * bj - ei * ej);
}

// Cyclic PCR on N = 2S reduced rows of one line (lower, diagonal, upper, rhs), row i of
// the line at base + slot(i)*rs of the four arrays P[0..3] or Q[0..3], where slot(2s+w) =
// w*S + s: the first rows of all segments, then their last rows, so that a wavefront's
// accesses to one of a thread's two rows are consecutive words (no LDS bank conflicts;
// rows 2s and 2s+1 side by side would put lanes 16 bytes apart). The steps alternate
// between the two buffers, so one barrier per step orders every read before the next
// overwrite, and each thread keeps its own two rows in registers, reading only the
// neighbours' from LDS. Workgroups holding several lines interleave them (rs = lines per
// workgroup), so that lanes of neighbouring lines touch neighbouring LDS words. Every
// thread of the block must call it (it holds barriers); `active` threads own rows 2*seg
// and 2*seg+1. Returns the solution of those two rows.
__device__ __forceinline__ void pcr_cyclic(double *const (&P)[4], double *const (&Q)[4], int base, int rs, int seg,
                                           int S, bool active, const RedRows &r, double &xF, double &xL)
{
    const int N = 2 * S;
    auto at = [&](int i) { return base + ((i & 1) * S + (i >> 1)) * rs; };
    double lo[2] = {r.lo0, r.A_L}, dg[2] = {r.A_F, r.B_L}, up[2] = {r.C_F, r.upL}, rh[2] = {r.R_F, r.R_L};
    const double *c0 = P[0], *c1 = P[1], *c2 = P[2], *c3 = P[3];
    double *n0 = Q[0], *n1 = Q[1], *n2 = Q[2], *n3 = Q[3];
    if (active) {
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const int ii = base + (w * S + seg) * rs;
            P[0][ii] = lo[w]; P[1][ii] = dg[w]; P[2][ii] = up[w]; P[3][ii] = rh[w];
        }
    }
    __syncthreads();
    for (int st = 1; st < S; st <<= 1) {   // strides 1, 2, ..., N/4
        if (active) {
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int i = 2 * seg + w;
                const int im = at((i - st) & (N - 1)), ip = at((i + st) & (N - 1));
                pcr_combine(lo[w], dg[w], up[w], rh[w], c0[im], c1[im], c2[im], c3[im], c0[ip], c1[ip], c2[ip], c3[ip]);
            }
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int ii = base + (w * S + seg) * rs;
                n0[ii] = lo[w]; n1[ii] = dg[w]; n2[ii] = up[w]; n3[ii] = rh[w];
            }
        }
        __syncthreads();
        const double *t0 = n0, *t1 = n1, *t2 = n2, *t3 = n3;
        n0 = const_cast<double *>(c0); n1 = const_cast<double *>(c1);
        n2 = const_cast<double *>(c2); n3 = const_cast<double *>(c3);
        c0 = t0; c1 = t1; c2 = t2; c3 = t3;
    }
    xF = 0.0;
    xL = 0.0;
    if (active) {
        // rows i and j = i + N/2 are coupled only to each other:
        //   b_i x_i + e_i x_j = d_i,  e_j x_i + b_j x_j = d_j,  e = lower + upper
        double xs[2];
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const int i = 2 * seg + w;
            const int jj = at((i + S) & (N - 1));
            xs[w] = pcr_pair(dg[w], lo[w] + up[w], rh[w], c1[jj], c0[jj] + c2[jj], c3[jj]);
        }
        xF = xs[0];
        xL = xs[1];
    }
}

// Reads a double from lane `src` of the wavefront (the source lane must be active).
__device__ __forceinline__ double lane_read(double v, int src)
{
    const int a = src << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(a, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(a, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// The steps of cyclic PCR after the first, inside one wavefront, with no LDS and no
// barrier. Lane l (0..G-1, G a power of two, 2 <= G <= 64) holds the two rows (w = 0, 1)
// of segment l, cyclic over the G lanes. The steps at strides 2, 4, ..., G/2 read the
// rows of the same w from the lanes 1, 2, ..., G/4 away by cross-lane permutes and
// combine them (pcr_combine); then rows G/2 lanes apart are coupled only to each other
// (pcr_pair). Each row's arithmetic is pcr_cyclic's, step for step. Lanes G..63 take part
// in the permutes with whatever they hold and return garbage. Every lane of the
// wavefront must call it.
__device__ __forceinline__ void pcr_lane_steps(int G, int l, double (&lo)[2], double (&dg)[2], double (&up)[2],
                                               double (&rh)[2], double &x0, double &x1)
{
    const int msk = G - 1;
    for (int d = 1; d < (G >> 1); d <<= 1) {
        const int lm = (l - d) & msk, lp = (l + d) & msk;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const double m0 = lane_read(lo[w], lm), m1 = lane_read(dg[w], lm), m2 = lane_read(up[w], lm),
                         m3 = lane_read(rh[w], lm);
            const double p0 = lane_read(lo[w], lp), p1 = lane_read(dg[w], lp), p2 = lane_read(up[w], lp),
                         p3 = lane_read(rh[w], lp);
            pcr_combine(lo[w], dg[w], up[w], rh[w], m0, m1, m2, m3, p0, p1, p2, p3);
        }
    }
    const int pl = (l + (G >> 1)) & msk;
    double xs[2];
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const double e = lo[w] + up[w];
        const double bj = lane_read(dg[w], pl), ej = lane_read(e, pl), dj = lane_read(rh[w], pl);
        xs[w] = pcr_pair(dg[w], e, rh[w], bj, ej, dj);
    }
    x0 = xs[0];
    x1 = xs[1];
}

// pcr_cyclic for S <= 64 segments held by lanes 0..S-1 of one wavefront (lane l, segment
// l): the stride-1 step mixes a lane's two rows with its neighbours' (row 2l with rows
// 2l-1 and 2l+1, row 2l+1 with rows 2l and 2l+2), the other steps are pcr_lane_steps.
// Same steps and arithmetic as pcr_cyclic, so the same bits, without LDS or barriers.
// Every lane of the wavefront calls it; lanes l >= S return garbage.
__device__ __forceinline__ void pcr_wave(int S, int l, const RedRows &r, double &xF, double &xL)
{
    double lo[2] = {r.lo0, r.A_L}, dg[2] = {r.A_F, r.B_L}, up[2] = {r.C_F, r.upL}, rh[2] = {r.R_F, r.R_L};
    if (S == 1) {   // rows 0 and 1 are the coupled pair
        xF = pcr_pair(dg[0], lo[0] + up[0], rh[0], dg[1], lo[1] + up[1], rh[1]);
        xL = pcr_pair(dg[1], lo[1] + up[1], rh[1], dg[0], lo[0] + up[0], rh[0]);
        return;
    }
    const int msk = S - 1, sl = l & msk;
    const int lm = (sl - 1) & msk, lp = (sl + 1) & msk;
    const double m0 = lane_read(lo[1], lm), m1 = lane_read(dg[1], lm), m2 = lane_read(up[1], lm),
                 m3 = lane_read(rh[1], lm);
    const double p0 = lane_read(lo[0], lp), p1 = lane_read(dg[0], lp), p2 = lane_read(up[0], lp),
                 p3 = lane_read(rh[0], lp);
    const double a0 = lo[0], b0 = dg[0], c0 = up[0], d0 = rh[0];
    pcr_combine(lo[0], dg[0], up[0], rh[0], m0, m1, m2, m3, lo[1], dg[1], up[1], rh[1]);
    pcr_combine(lo[1], dg[1], up[1], rh[1], a0, b0, c0, d0, p0, p1, p2, p3);
    pcr_lane_steps(S, sl, lo, dg, up, rh, xF, xL);
}

// Interior unknown q of an eliminated segment.
template <int SEG_M>
__device__ __forceinline__ double seg_value(int q, int m, const double (&lo)[SEG_M], const double (&di)[SEG_M],
                                            const double (&up)[SEG_M], const double (&rh)[SEG_M], double xF,
                                            double xL)
{
    if (q == 0) return xF;
    if (q == m - 1) return xL;
    return (rh[q] - lo[q] * xF - up[q] * xL) * di[q];
}

// One block: lines_per_block lines x S segments. COLS: lines are columns of row-major
// arrays (element (k, line) at k*ld + line) and thread t maps to (line t % L, segment
// t / L) so that a wavefront reads runs of consecutive columns; ROWS: lines are rows
// (element (line, k) at line*ld + k) and t maps to (line t / S, segment t % S).
template <bool COLS, int BLOCK, int SEG_M, class Prov>
__device__ __forceinline__ void partition_body(const Prov &prov, double *__restrict__ Y, const LineGeom &g, int blk,
                                               double (*red)[4][BLOCK * 2])
{
    const int t = threadIdx.x;
    const int S = g.S, L = g.lines_per_block;
    const int lloc = COLS ? (t % L) : (t / S);
    const int seg = COLS ? (t / L) : (t % S);
    const int line = blk * L + lloc;
    const bool active = (t < L * S) && line < g.nL;
    const int n = g.n;
    const int k0 = (int)(((long)seg * n) / S);
    const int k1 = (int)(((long)(seg + 1) * n) / S);
    const int m = k1 - k0;   // 2..SEG_M

    double lo[SEG_M], di[SEG_M], up[SEG_M], rh[SEG_M];
    RedRows rr{};
    if (active) {
#pragma unroll
        for (int q = 0; q < SEG_M; ++q)
            if (q < m) prov.get(line, k0 + q, di[q], lo[q], up[q], rh[q]);
        if (!g.periodic) {
            if (k0 == 0) lo[0] = 0.0;
            if (k1 == n) {   // static indices keep the segment arrays in registers
#pragma unroll
                for (int q = 0; q < SEG_M; ++q)
                    if (q == m - 1) up[q] = 0.0;
            }
        }
        rr = seg_eliminate<SEG_M>(m, lo, di, up, rh);
    }
    double xF, xL;
    double *const P[4] = {red[0][0], red[0][1], red[0][2], red[0][3]};
    double *const Q[4] = {red[1][0], red[1][1], red[1][2], red[1][3]};
    if (COLS) pcr_cyclic(P, Q, lloc, L, seg, S, active, rr, xF, xL);
    else pcr_cyclic(P, Q, lloc * 2 * S, 1, seg, S, active, rr, xF, xL);
    if (active) {
#pragma unroll
        for (int q = 0; q < SEG_M; ++q)
            if (q < m)
                stn<prov_nts<Prov>()>(Y + elem_index<COLS>(g, line, k0 + q, g.ldy),
                                                   seg_value<SEG_M>(q, m, lo, di, up, rh, xF, xL));
    }
}

template <bool COLS, int BLOCK, int SEG_M, class Prov>
__global__ void __launch_bounds__(BLOCK) k_partition_solve(Prov prov, double *__restrict__ Y, LineGeom g)
{
    __shared__ double red[2][4][BLOCK * 2];   // reduced rows (lower, diag, upper, rhs), two buffers
    const int blk = g.xcd_remap ? xcd_chunk((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    partition_body<COLS, BLOCK, SEG_M>(prov, Y, g, blk, red);
}

// Row lines staged through LDS: one line (n <= RB*RM) per block of RB threads. The block
// assembles the line's coefficients in coalesced order and hands each thread its segment
// through a two-array LDS tile (element e at e + e/8, so that the per-thread segment reads
// are free of bank conflicts), in two passes: the diagonal and lower coefficients first,
// while each thread keeps the upper and right-hand-side values it assembled in registers,
// then those. The tile (two arrays) then holds the reduced system's two PCR buffers and
// finally the solution on its way to a coalesced store. At 2*RB*RM*9/8 doubles the tile
// leaves room for two 512-thread or four 256-thread blocks per CU, so one block's
// elimination and PCR overlap another block's assembly.
constexpr int RM = 8;                  // unknowns per thread
template <int RB> struct RowsLds {     // RB threads per block, lines up to RB*RM
    static constexpr int NMAX = RB * RM;
    static constexpr int LDS = NMAX + NMAX / 8;
};
__device__ __forceinline__ int lpad(int e) { return e + (e >> 3); }

template <class P, class = void> struct HasGet2 : std::false_type {};
template <class P> struct HasGet2<P, std::void_t<decltype(&P::get2)>> : std::true_type {};

template <int RB, class Prov>
__device__ __forceinline__ void rows_lds_body(const Prov &prov, double *__restrict__ Y, long ldy, int n, int S,
                                              int periodic, int line, double (*sh)[RowsLds<RB>::LDS])
{
    const int t = threadIdx.x;
    double cu[RM], cf[RM];   // upper and rhs of the elements this thread assembles
    // Elements of pass it: t + it*RB, or with get2() and an even line the pairs
    // 2*(t + it*RB) + {0, 1} (passes it < RM/2).
    bool pairs = false;
    if constexpr (HasGet2<Prov>::value) {
        pairs = (n & 1) == 0;
        if (pairs) {
#pragma unroll
            for (int it = 0; it < RM / 2; ++it) {
                const int e = 2 * (t + it * RB);
                if (e < n) {
                    double a0, b0, a1, b1;
                    prov.get2(line, e, a0, b0, cu[2 * it], cf[2 * it], a1, b1, cu[2 * it + 1], cf[2 * it + 1]);
                    const int p0 = lpad(e), p1 = lpad(e + 1);
                    sh[0][p0] = a0; sh[1][p0] = b0;
                    sh[0][p1] = a1; sh[1][p1] = b1;
                }
            }
        }
    }
    if (!pairs) {
#pragma unroll
        for (int it = 0; it < RM; ++it) {
            const int e = t + it * RB;
            if (e < n) {
                double a, b;
                prov.get(line, e, a, b, cu[it], cf[it]);
                const int p = lpad(e);
                sh[0][p] = a; sh[1][p] = b;
            }
        }
    }
    __syncthreads();
    const bool active = t < S;
    const int seg = t;
    const int k0 = (int)(((long)seg * n) / S);
    const int k1 = (int)(((long)(seg + 1) * n) / S);
    const int m = k1 - k0;
    double lo[RM], di[RM], up[RM], rh[RM];
    if (active) {
#pragma unroll
        for (int q = 0; q < RM; ++q) {
            if (q < m) {
                const int p = lpad(k0 + q);
                di[q] = sh[0][p]; lo[q] = sh[1][p];
            }
        }
    }
    __syncthreads();
    if (pairs) {
#pragma unroll
        for (int it = 0; it < RM / 2; ++it) {
            const int e = 2 * (t + it * RB);
            if (e < n) {
                const int p0 = lpad(e), p1 = lpad(e + 1);
                sh[0][p0] = cu[2 * it]; sh[1][p0] = cf[2 * it];
                sh[0][p1] = cu[2 * it + 1]; sh[1][p1] = cf[2 * it + 1];
            }
        }
    } else {
#pragma unroll
        for (int it = 0; it < RM; ++it) {
            const int e = t + it * RB;
            if (e < n) {
                const int p = lpad(e);
                sh[0][p] = cu[it]; sh[1][p] = cf[it];
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int q = 0; q < RM; ++q) {
            if (q < m) {
                const int p = lpad(k0 + q);
                up[q] = sh[0][p]; rh[q] = sh[1][p];
            }
        }
        if (!periodic) {
            if (k0 == 0) lo[0] = 0.0;
            if (k1 == n) {   // static indices keep the segment arrays in registers
#pragma unroll
                for (int q = 0; q < RM; ++q)
                    if (q == m - 1) up[q] = 0.0;
            }
        }
    }
    __syncthreads();   // the coefficient tile is free from here on
    RedRows rr{};
    if (active) rr = seg_eliminate<RM>(m, lo, di, up, rh);
    double xF = 0.0, xL = 0.0;
    if (S <= 64) {
        // every segment is in wavefront 0: PCR by cross-lane permutes, no LDS
        if (t < 64) pcr_wave(S, t, rr, xF, xL);
    } else {
        const int N = 2 * S;   // 4N <= 4*RB*2 < LDS doubles per array
        double *const P[4] = {sh[0], sh[0] + N, sh[0] + 2 * N, sh[0] + 3 * N};
        double *const Q[4] = {sh[1], sh[1] + N, sh[1] + 2 * N, sh[1] + 3 * N};
        pcr_cyclic(P, Q, 0, 1, seg, S, active, rr, xF, xL);
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int q = 0; q < RM; ++q)
            if (q < m) sh[0][lpad(k0 + q)] = seg_value<RM>(q, m, lo, di, up, rh, xF, xL);
    }
    __syncthreads();
    double *yl = Y + (long)line * ldy;
    for (int e = t; e < n; e += RB) stn<prov_nts<Prov>()>(yl + e, sh[0][lpad(e)]);
}

template <int RB, class Prov>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_rows_lds(Prov prov, double *__restrict__ Y, long ldy, int n, int S, int periodic, int rev)
{
    __shared__ double sh[2][RowsLds<RB>::LDS];
    // rev: lines last to first
    rows_lds_body<RB>(prov, Y, ldy, n, S, periodic, rev ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x, sh);
}

// Picks S (power of two, n >= 2S, ceil(n/S) <= M) for a line length n; 0 if none.
static int pick_segments(int n, int M, int smax)
{
    int S = 1;
    while (S < smax && (n + S - 1) / S > M) S <<= 1;
    while (S > 1 && n < 2 * S) S >>= 1;
    if ((n + S - 1) / S > M || n < 2) return 0;
    return S;
}

// Lines up to 8192 unknowns: rows through the LDS row kernel (up to 1024 segments of 8),
// columns through the partition kernel (up to 1024 segments of 8 in 1024-thread blocks).
bool partition_supported(int n) { return pick_segments(n, 8, 1024) > 0; }

// smax bounds the segment count (0: the default, BLOCK for rows and BLOCK/4 for columns).
template <bool COLS, int BLOCK, int M, class Prov>
static hipError_t launch_partition(hipStream_t st, const Prov &prov, double *Y, long ldy, int n, int nL, bool periodic,
                                   int smax = 0)
{
    LineGeom g{};
    g.ld = 0; g.ldy = ldy; g.n = n; g.nL = nL; g.periodic = periodic ? 1 : 0;
    // columns: by default at least 4 lines per block, so a row access covers 32
    // contiguous bytes
    g.S = pick_segments(n, M, smax > 0 ? smax : (COLS ? BLOCK / 4 : BLOCK));
    if (g.S == 0) return hipErrorInvalidValue;
    int lg = 0;
    while ((1 << lg) < g.S) ++lg;
    g.log2S = lg;
    g.lines_per_block = BLOCK / g.S;
    const int blocks = (nL + g.lines_per_block - 1) / g.lines_per_block;
    g.xcd_remap = COLS ? 1 : 0;
    hipLaunchKernelGGL((k_partition_solve<COLS, BLOCK, M, Prov>), dim3((unsigned)blocks), dim3(BLOCK), 0, st, prov, Y,
                       g);
    return hipGetLastError();
}

// Workgroup size (1024, 256 or 64 threads) of the column partition solve of nL lines of n
// unknowns with S segments each.
static int col_workgroup(int n, int nL, int S)
{
    const int sizes[3] = {1024, 256, 64};
    int best = 1024;
    for (int B : sizes) {
        if (pick_segments(n, 8, B / 4) != S || B / S < 1) continue;
        best = B;
        if ((nL + B / S - 1) / (B / S) >= 256) break;
    }
    return best;
}

// Column lines (x-sweeps): tile = unknowns per thread (8 or 16; 0 = default 8).
template <class Prov>
static hipError_t dispatch_cols(hipStream_t st, const Prov &prov, double *Y, long ldy, int n, int nL, int tile)
{
    if (tile == 16 && pick_segments(n, 16, 64) > 0) return launch_partition<true, 256, 16>(st, prov, Y, ldy, n, nL, false);
    const int S = pick_segments(n, 8, 256);
    if (S > 0) {
        // A workgroup of B threads holds B/S lines. With few lines (columns) 1024-thread
        // workgroups leave CUs idle, so take the largest B that still gives one
        // workgroup per CU. The segment count, and so every result, is the same for all B.
        const int wg = col_workgroup(n, nL, S);
        if (wg == 64) return launch_partition<true, 64, 8>(st, prov, Y, ldy, n, nL, false);
        if (wg == 256) return launch_partition<true, 256, 8>(st, prov, Y, ldy, n, nL, false);
        return launch_partition<true, 1024, 8>(st, prov, Y, ldy, n, nL, false);
    }
    // columns longer than 2048: up to 1024 segments, so fewer than 4 columns per
    // 1024-thread workgroup (2 up to 4096, 1 up to 8192)
    return launch_partition<true, 1024, 8>(st, prov, Y, ldy, n, nL, false, 1024);
}

template <class Prov>
static hipError_t launch_rows_lds(hipStream_t st, const Prov &prov, double *Y, long ldy, int n, int nL, bool periodic,
                                  int rb, int rev = 0);

// Row lines (y-sweeps): tile = unknowns per thread (8 or 16; 0 = default 8).
template <class Prov>
static hipError_t dispatch_rows(hipStream_t st, const Prov &prov, double *Y, long ldy, int n, int nL, bool periodic,
                                int tile)
{
    if (tile == 16 && pick_segments(n, 16, 256) > 0)
        return launch_partition<false, 256, 16>(st, prov, Y, ldy, n, nL, periodic);
    // rows up to 8192: staged through LDS (coalesced loads); the same segments and
    // arithmetic as the partition kernel below
    if (pick_segments(n, RM, 1024) > 0) return launch_rows_lds(st, prov, Y, ldy, n, nL, periodic, 0);
    if (pick_segments(n, 8, 512) > 0) return launch_partition<false, 512, 8>(st, prov, Y, ldy, n, nL, periodic);
    return launch_partition<false, 256, 16>(st, prov, Y, ldy, n, nL, periodic);
}

hipError_t launch_partition_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                 double *Y, int n, int m, long ld, int tile)
{
    return dispatch_cols(st, ProvArrays<true>{A, B, C, F, ld}, Y, ld, n, m, tile);
}

hipError_t launch_partition_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                 long ld, double *Y, long ldy, int nL, int n, bool periodic, int tile)
{
    return dispatch_rows(st, ProvArrays<false>{A, B, C, F, ld}, Y, ldy, n, nL, periodic, tile);
}

// Transposed systems for the adjoint: A^T has diagonal a_k, lower c_{k-1} and upper
// b_{k+1} (periodic: indices wrap, so the corners swap as well).
struct ProvTransCols {   // columns, non-periodic; element (k, line) at k*ld + line
    const double *A, *B, *C, *F;
    long ld;
    int n;
    __device__ __forceinline__ void get(int line, int k, double &a, double &b, double &c, double &f) const
    {
        const long e = (long)k * ld + line;
        a = A[e];
        b = (k >= 1) ? C[e - ld] : 0.0;
        c = (k <= n - 2) ? B[e + ld] : 0.0;
        f = F[e];
    }
};
struct ProvTransRows {   // rows, periodic; element (line, k) at line*ld + k
    const double *A, *B, *C, *F;
    long ld;
    int n;
    __device__ __forceinline__ void get(int line, int k, double &a, double &b, double &c, double &f) const
    {
        const long r = (long)line * ld;
        a = A[r + k];
        b = C[r + (k == 0 ? n - 1 : k - 1)];
        c = B[r + (k == n - 1 ? 0 : k + 1)];
        f = F[r + k];
    }
};

hipError_t launch_transposed_cols(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                  double *Y, int n, int m, long ld, int tile)
{
    return dispatch_cols(st, ProvTransCols{A, B, C, F, ld, n}, Y, ld, n, m, tile);
}

hipError_t launch_transposed_rows(hipStream_t st, const double *A, const double *B, const double *C, const double *F,
                                  long ld, double *Y, long ldy, int nL, int n, int tile)
{
    return dispatch_rows(st, ProvTransRows{A, B, C, F, ld, n}, Y, ldy, n, nL, true, tile);
}

// ---------------------------------------------------------------------------------
// Column lines (x-sweeps) in three coalesced passes. Thread (column j, segment s) with
// 64 consecutive columns per wavefront, so every access is a 512-byte run:
//   A. assemble the segment's CM rows, eliminate it, store the eliminated segment and
//      its two reduced rows;
//   B. solve the 2*nseg-row reduced systems of all columns (partition solver);
//   C. back-substitute the segment from its two boundary values.
// ---------------------------------------------------------------------------------
template <class P, class = void> struct HasSegment : std::false_type {};
template <class P> struct HasSegment<P, std::void_t<decltype(&P::segment)>> : std::true_type {};

int cols_segments(int n) { return (n + CM - 1) / CM; }

template <class Prov>
__device__ __forceinline__ void cols_elim_body(const Prov &prov, const ColsWork &cw, int n, int ny, int bx, int by)
{
    const int j = bx * 64 + threadIdx.x;
    // a wavefront is one threadIdx.y of the 64x4 workgroup: its segment, and so its rows,
    // are uniform, which keeps the per-row metrics in scalar loads
    const int sg = __builtin_amdgcn_readfirstlane(by * 4 + (int)threadIdx.y);
    const int nseg = cw.nseg;
    if (j >= ny || sg >= nseg) return;
    const int k0 = (int)(((long)sg * n) / nseg);
    const int k1 = (int)(((long)(sg + 1) * n) / nseg);
    const int m = k1 - k0;   // 2..CM
    double lo[CM], di[CM], up[CM], rh[CM];
    if constexpr (HasSegment<Prov>::value) {
        prov.segment(j, k0, m, lo, di, up, rh);
    } else {
#pragma unroll
        for (int q = 0; q < CM; ++q)
            if (q < m) prov.get(j, k0 + q, di[q], lo[q], up[q], rh[q]);
    }
    if (k0 == 0) lo[0] = 0.0;
    if (k1 == n) {   // static indices keep the segment arrays in registers
#pragma unroll
        for (int q = 0; q < CM; ++q)
            if (q == m - 1) up[q] = 0.0;
    }
    const RedRows r = seg_eliminate<CM>(m, lo, di, up, rh);
    constexpr bool NTS = prov_nts<Prov>();
#pragma unroll
    for (int q = 1; q < CM - 1; ++q) {
        if (q < m - 1) {
            const long o = (long)(k0 + q) * ny + j;
            stn<NTS>(cw.E0 + o, lo[q] * di[q]);
            stn<NTS>(cw.E1 + o, up[q] * di[q]);
            stn<NTS>(cw.E2 + o, rh[q] * di[q]);
        }
    }
    const long oF = (long)(2 * sg) * ny + j, oL = oF + ny;
    stn<NTS>(cw.R0 + oF, r.lo0); stn<NTS>(cw.R1 + oF, r.A_F); stn<NTS>(cw.R2 + oF, r.C_F); stn<NTS>(cw.R3 + oF, r.R_F);
    stn<NTS>(cw.R0 + oL, r.A_L); stn<NTS>(cw.R1 + oL, r.B_L); stn<NTS>(cw.R2 + oL, r.upL); stn<NTS>(cw.R3 + oL, r.R_L);
}

template <bool STR>
__device__ __forceinline__ void cols_backsub_body(const ColsWork &cw, double *__restrict__ Y, int n, int ny, int bx,
                                                  int by)
{
    const int j = bx * 64 + threadIdx.x;
    // a wavefront is one threadIdx.y of the 64x4 workgroup: its segment bounds are scalars
    const int sg = __builtin_amdgcn_readfirstlane(by * 4 + (int)threadIdx.y);
    const int nseg = cw.nseg;
    if (j >= ny || sg >= nseg) return;
    const int k0 = (int)(((long)sg * n) / nseg);
    const int k1 = (int)(((long)(sg + 1) * n) / nseg);
    const int m = k1 - k0;
    constexpr bool NTL = Nt<NT_BACKSUB, STR>::L, NTS = Nt<NT_BACKSUB, STR>::S;
    const double xF = ldn<NTL>(cw.X + (long)(2 * sg) * ny + j);
    const double xL = ldn<NTL>(cw.X + (long)(2 * sg + 1) * ny + j);
#pragma unroll
    for (int q = 0; q < CM; ++q) {
        if (q < m) {
            const long o = (long)(k0 + q) * ny + j;
            double x;
            if (q == 0) x = xF;
            else if (q == m - 1) x = xL;
            else x = ldn<NTL>(cw.E2 + o) - ldn<NTL>(cw.E0 + o) * xF - ldn<NTL>(cw.E1 + o) * xL;
            stn<NTS>(Y + o, x);
        }
    }
}

// rev: take the segment groups last to first (see sweep order in m70h_api.hip).
template <class Prov>
__global__ void __launch_bounds__(256) k_cols_elim(Prov prov, ColsWork cw, int n, int ny, int rev)
{
    cols_elim_body(prov, cw, n, ny, (int)blockIdx.x, rev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y);
}

template <bool STR>
__global__ void __launch_bounds__(256) k_cols_backsub(ColsWork cw, double *__restrict__ Y, int n, int ny, int rev)
{
    cols_backsub_body<STR>(cw, Y, n, ny, (int)blockIdx.x, rev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y);
}

// One pass of the three-pass column solve: COLS_ELIM (assemble + eliminate the segments),
// COLS_REDUCED (the reduced systems, 2*nseg rows per column) or COLS_BACKSUB. The passes
// run in this order on one stream; each is its own launch so it can be timed alone. rev:
// the elimination and back-substitution passes take the rows bottom to top.
template <class Prov>
static hipError_t launch_cols3(hipStream_t st, const Prov &prov, const ColsWork &cw, double *Y, int n, int ny, int tile,
                               int pass, int rev)
{
    if (n < 3 || cols_segments(n) != cw.nseg) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((ny + 63) / 64), (unsigned)((cw.nseg + 3) / 4)), block(64, 4);
    switch (pass) {
    case COLS_ELIM:
        hipLaunchKernelGGL((k_cols_elim<Prov>), grid, block, 0, st, prov, cw, n, ny, rev);
        return hipGetLastError();
    case COLS_REDUCED:
        // diagonal R1, lower R0, upper R2, rhs R3; 2*nseg rows per column
        return dispatch_cols(st, ProvArrays<true, NT_REDUCED, ProvStream<Prov>::value>{cw.R1, cw.R0, cw.R2, cw.R3, (long)ny},
                             cw.X, ny, 2 * cw.nseg, ny, tile);
    case COLS_BACKSUB:
        hipLaunchKernelGGL(k_cols_backsub<ProvStream<Prov>::value>, grid, block, 0, st, cw, Y, n, ny, rev);
        return hipGetLastError();
    default:
        return hipErrorInvalidValue;
    }
}

template <bool CDC, bool VCF>
static hipError_t launch_cols3_qxx_t(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                                     const RunFields &rf, const Work &w, double f0, const ColsWork &cw, int tile,
                                     int pass, int rev, bool stream)
{
    double *Y = s.Qx + k.ny;
    const int n = k.nx - 1, ny = k.ny;
    return stream ? launch_cols3(st, ProvQxx<CDC, VCF, true>{k, m, s, rf, w, f0}, cw, Y, n, ny, tile, pass, rev)
                  : launch_cols3(st, ProvQxx<CDC, VCF, false>{k, m, s, rf, w, f0}, cw, Y, n, ny, tile, pass, rev);
}
hipError_t launch_cols3_qxx(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                            const Work &w, double f0, const ColsWork &cw, int tile, int pass, bool vc_from_v, int rev,
                            bool stream)
{
    if (k.drag_model == 0)
        return vc_from_v ? launch_cols3_qxx_t<true, true>(st, k, m, s, rf, w, f0, cw, tile, pass, rev, stream)
                         : launch_cols3_qxx_t<true, false>(st, k, m, s, rf, w, f0, cw, tile, pass, rev, stream);
    return vc_from_v ? launch_cols3_qxx_t<false, true>(st, k, m, s, rf, w, f0, cw, tile, pass, rev, stream)
                     : launch_cols3_qxx_t<false, false>(st, k, m, s, rf, w, f0, cw, tile, pass, rev, stream);
}
hipError_t launch_cols3_qyx(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                            const Work &w, const ColsWork &cw, int tile, int pass, int rev, bool stream)
{
    return stream ? launch_cols3(st, ProvQyx<true>{k, m, s, rf, w}, cw, s.Qy + k.ny, k.nx - 1, k.ny, tile, pass, rev)
                  : launch_cols3(st, ProvQyx<false>{k, m, s, rf, w}, cw, s.Qy + k.ny, k.nx - 1, k.ny, tile, pass, rev);
}

bool rows_lds_supported(int n) { return pick_segments(n, RM, 1024) > 0; }

// Row block: 256 threads (lines up to 2048, 36 KB of LDS; four blocks per CU), 512 threads
// (lines up to 4096, 72 KB; two blocks per CU) or 1024 threads (lines up to 8192, 144 KB;
// one block per CU). rb = 0 picks the smallest fit; a block too small for the line is
// replaced by the smallest that fits.
int rows_block_for(int n, int rb)
{
    if (n > RowsLds<512>::NMAX) return 1024;
    if (rb == 256 && pick_segments(n, RM, 256) > 0 && n <= RowsLds<256>::NMAX) return 256;
    if (rb == 512) return 512;
    return n <= RowsLds<256>::NMAX ? 256 : 512;
}

// ---------------------------------------------------------------------------------
// Grouped (multi-scenario) versions of the FUSED stages. Providers read the scenario's
// descriptor in place.
// ---------------------------------------------------------------------------------
// VCF: V_c formed from V (see ProvQxx); CDC: every scenario of the batch has the constant
// drag model (cD_x is cD0), otherwise each scenario's own model decides
template <bool VCF, bool CDC>
struct GProvQxx {
    static constexpr int nt_site = NT_QXX;
    static constexpr bool nt_stream = GROUPED_STREAM;
    const ScenDesc *d;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    { asm_qxx(d->k, d->m, d->s, d->rf, d->w, scen_f0(*d), kk, line, a, b, c, f); }
    __device__ __forceinline__ void segment(int line, int k0, int mm, double (&lo)[CM], double (&di)[CM],
                                            double (&up)[CM], double (&rh)[CM]) const
    {
        qxx_segment<CDC, VCF, GROUPED_STREAM>(d->k, d->m, d->s, d->rf, d->w, scen_f0(*d), line, k0, mm, lo, di, up, rh,
                                              !CDC && d->k.drag_model == 0);
    }
};
struct GProvQyx {
    static constexpr int nt_site = NT_QYX;
    static constexpr bool nt_stream = GROUPED_STREAM;
    const ScenDesc *d;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    { asm_qyx<GROUPED_STREAM>(d->k, d->m, d->s, d->rf, d->w, kk, line, a, b, c, f); }
    __device__ __forceinline__ void segment(int line, int k0, int mm, double (&lo)[CM], double (&di)[CM],
                                            double (&up)[CM], double (&rh)[CM]) const
    { qyx_segment<GROUPED_STREAM>(d->k, d->m, d->s, d->rf, d->w, line, k0, mm, lo, di, up, rh); }
};
struct GProvQyy {
    static constexpr int nt_site = NT_QYY_ROWS;
    static constexpr bool nt_stream = GROUPED_STREAM;
    const ScenDesc *d;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    { asm_qyy(d->k, d->m, d->s, d->rf, d->w, line, kk, a, b, c, f); }
    __device__ __forceinline__ void get2(int line, int kk, double &a0, double &b0, double &c0, double &f0, double &a1,
                                         double &b1, double &c1, double &f1) const
    {
        QyyIn q0, q1;
        qyy_load2<GROUPED_STREAM>(d->k, d->s, d->rf, d->w, line, kk, q0, q1, d->k.drag_model != 0);
        qyy_nud_pair<GROUPED_STREAM>(d->k, d->rf, line, kk, q0, q1);
        double d0, d1;
        asm_qyy_vals(d->k, d->m, q0, line, a0, b0, c0, f0, d0);
        asm_qyy_vals(d->k, d->m, q1, line, a1, b1, c1, f1, d1);
        stn2<Nt<NT_QYY_ROWS, GROUPED_STREAM>::S>(d->w.drag + (long)line * d->k.ny + kk, d0, d1);
    }
};

// Qx projector rows of a grouped step, with Qx* formed from the scenario's column passes
// (as ProvQxyBs).
template <bool VCF>
struct GProvQxyBs {
    static constexpr int nt_site = NT_QXY_ROWS;
    static constexpr bool nt_stream = GROUPED_STREAM;
    const ScenDesc *d;
    __device__ __forceinline__ void get(int line, int kk, double &a, double &b, double &c, double &f) const
    {
        QxyIn q = qxy_load(d->k, d->s, d->rf, d->w, line, kk);
        if (VCF) qxy_vc_from_v(d->k, d->s, line, kk, q);
        q.qx = cols_value(d->cw, d->k.nx - 1, d->k.ny, line, kk);
        asm_qxy_vals(d->k, q, a, b, c, f);
    }
    __device__ __forceinline__ void get2(int line, int kk, double &a0, double &b0, double &c0, double &f0, double &a1,
                                         double &b1, double &c1, double &f1) const
    {
        QxyIn q0, q1;
        qxy_load2<GROUPED_STREAM>(d->k, d->s, d->rf, d->w, line, kk, q0, q1);
        if (VCF) qxy_vc_from_v2<GROUPED_STREAM>(d->k, d->s, line, kk, q0, q1);
        cols_value2<GROUPED_STREAM>(d->cw, d->k.nx - 1, d->k.ny, line, kk, q0.qx, q1.qx);
        asm_qxy_vals(d->k, q0, a0, b0, c0, f0);
        asm_qxy_vals(d->k, q1, a1, b1, c1, f1);
    }
};

// WHICH 0: Qx projector rows, 1: Qy predictor rows; VCF: the Qx projector forms V_c from V
template <int WHICH, int RB, bool VCF>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_rows_lds_grouped(cScenDesc *D, const int *__restrict__ start, int nscen, int rev)
{
    __shared__ double sh[2][RowsLds<RB>::LDS];
    const int b = rev ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;   // rev: last scenario's last row first
    const int sid = scen_of(start, nscen, b);
    const ScenDesc *d = (const ScenDesc *)(D + sid);
    const int ny = d->k.ny;
    const int line = b - start[sid];
    if (WHICH == 0) rows_lds_body<RB>(GProvQxyBs<VCF>{d}, d->s.Qx + ny, ny, ny, d->S_rows, 1, line, sh);
    else rows_lds_body<RB>(GProvQyy{d}, d->s.Qy + ny, ny, ny, d->S_rows, 1, line, sh);
}

// WHICH 0: Qx predictor columns, 1: Qy projector columns; VCF: the Qx predictor forms V_c
// from V; CDC: constant drag model in every scenario. The Qx predictor takes 181-185 VGPRs
// here (two waves per SIMD); held to three waves per SIMD it spills about 60 bytes per lane
// and the 64-scenario batch still steps 0.8% faster (7.34e9 -> 7.40e9 cell-steps/s).
template <int WHICH, bool VCF, bool CDC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_cols_elim_grouped(cScenDesc *D,
                                                           const int *__restrict__ start, int nscen, int rev)
{
    const int b = rev ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const ScenDesc *d = (const ScenDesc *)(D + sid);
    const int ny = d->k.ny;
    const int lb = b - start[sid];
    const int gx = (ny + 63) / 64;
    const int by = lb / gx, bx = lb - by * gx;
    if (WHICH == 0) cols_elim_body(GProvQxx<VCF, CDC>{d}, d->cw, d->k.nx - 1, ny, bx, by);
    else cols_elim_body(GProvQyx{d}, d->cw, d->k.nx - 1, ny, bx, by);
}

template <int WHICH>
__global__ void __launch_bounds__(256) k_cols_backsub_grouped(cScenDesc *D,
                                                              const int *__restrict__ start, int nscen, int rev)
{
    const int b = rev ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const ScenDesc *d = (const ScenDesc *)(D + sid);
    const int ny = d->k.ny;
    const int lb = b - start[sid];
    const int gx = (ny + 63) / 64;
    const int by = lb / gx, bx = lb - by * gx;
    cols_backsub_body<GROUPED_STREAM>(d->cw, (WHICH == 0 ? d->s.Qx : d->s.Qy) + ny, d->k.nx - 1, ny, bx, by);
}

// Reduced column systems of every scenario (partition solver, 8 unknowns per thread).
// Reduced systems of the FUSED x-sweeps for a batch: the partition solver with the
// scenario's segment count, RED_BLOCK-thread workgroups of RED_BLOCK/S lines.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_red_pcr_grouped(cScenDesc *D, const int *__restrict__ start,
                                                           int nscen)
{
    __shared__ double red[2][4][BLOCK * 2];
    const int b = blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const ScenDesc *d = (const ScenDesc *)(D + sid);
    const int ny = d->k.ny;
    LineGeom g{};
    g.ld = ny; g.ldy = ny; g.n = 2 * d->cw.nseg; g.nL = ny; g.S = d->S_red; g.lines_per_block = d->L_red;
    g.periodic = 0;
    partition_body<true, BLOCK, 8>(ProvArrays<true, NT_REDUCED, GROUPED_STREAM>{d->cw.R1, d->cw.R0, d->cw.R2, d->cw.R3, (long)ny}, d->cw.X, g,
                                   b - start[sid], red);
}

// ---------------------------------------------------------------------------------
// Line solves of the grouped adjoint (adjoint_batch.h): the taped step's systems and their
// transposes, every scenario with its own segments, as its single-scenario launch.
// ---------------------------------------------------------------------------------
template <int SYS, int TRANS>
__global__ void __launch_bounds__(1024) k_adj_cols_grouped(cAdjScen *D, const int *__restrict__ start,
                                                           int nscen)
{
    __shared__ double red[2][4][1024 * 2];
    const int b = blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const AdjScen *d = (const AdjScen *)(D + sid);
    const int ny = d->k.ny;
    LineGeom g{};
    g.ld = ny; g.ldy = ny; g.n = d->k.nx - 1; g.nL = ny; g.S = d->S_cols; g.lines_per_block = d->L_cols;
    g.periodic = 0;
    const double *A = SYS == 1 ? d->T.a1 : d->T.a4, *B = SYS == 1 ? d->T.b1 : d->T.b4;
    const double *C = SYS == 1 ? d->T.c1 : d->T.c4;
    if (TRANS) {
        partition_body<true, 1024, 8>(ProvTransCols{A, B, C, d->W.rhs, (long)ny, g.n}, d->W.lam, g, b - start[sid], red);
    } else {
        double *Y = SYS == 1 ? d->T.qxs : d->s.Qy + ny;
        partition_body<true, 1024, 8>(ProvArrays<true>{A, B, C, d->T.f, (long)ny}, Y, g, b - start[sid], red);
    }
}

template <int RB, int SYS, int TRANS>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_adj_rows_grouped(cAdjScen *D, const int *__restrict__ start, int nscen)
{
    __shared__ double sh[2][RowsLds<RB>::LDS];
    const int b = blockIdx.x;
    const int sid = scen_of(start, nscen, b);
    const AdjScen *d = (const AdjScen *)(D + sid);
    const int ny = d->k.ny;
    const int line = b - start[sid];
    const double *A = SYS == 2 ? d->T.a2 : d->T.a3, *B = SYS == 2 ? d->T.b2 : d->T.b3;
    const double *C = SYS == 2 ? d->T.c2 : d->T.c3;
    if (TRANS) {
        rows_lds_body<RB>(ProvTransRows{A, B, C, d->W.rhs, (long)ny, ny}, d->W.lam, ny, ny, d->S_rows, 1, line, sh);
    } else {
        double *Y = SYS == 2 ? d->s.Qx + ny : d->T.qys;
        rows_lds_body<RB>(ProvArrays<false>{A, B, C, d->T.f, (long)ny}, Y, ny, ny, d->S_rows, 1, line, sh);
    }
}

hipError_t launch_adj_cols_grouped(hipStream_t st, int sys, int trans, const AdjScen *D, const int *start, int nscen,
                                   int blocks)
{
    if (blocks <= 0) return hipSuccess;
    const dim3 gr((unsigned)blocks), bl(1024);
    if (sys == 1 && !trans) hipLaunchKernelGGL((k_adj_cols_grouped<1, 0>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
    else if (sys == 1) hipLaunchKernelGGL((k_adj_cols_grouped<1, 1>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
    else if (sys == 4 && !trans) hipLaunchKernelGGL((k_adj_cols_grouped<4, 0>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
    else if (sys == 4) hipLaunchKernelGGL((k_adj_cols_grouped<4, 1>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_adj_rows_grouped(hipStream_t st, int rb, int sys, int trans, const AdjScen *D, const int *start,
                                   int nscen, int blocks)
{
    if (blocks <= 0) return hipSuccess;
    const dim3 gr((unsigned)blocks);
    if (sys != 2 && sys != 3) return hipErrorInvalidValue;
    if (rb == 256) {
        const dim3 bl(256);
        if (sys == 2 && !trans) hipLaunchKernelGGL((k_adj_rows_grouped<256, 2, 0>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
        else if (sys == 2) hipLaunchKernelGGL((k_adj_rows_grouped<256, 2, 1>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
        else if (!trans) hipLaunchKernelGGL((k_adj_rows_grouped<256, 3, 0>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
        else hipLaunchKernelGGL((k_adj_rows_grouped<256, 3, 1>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
    } else if (rb == 512) {
        const dim3 bl(512);
        if (sys == 2 && !trans) hipLaunchKernelGGL((k_adj_rows_grouped<512, 2, 0>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
        else if (sys == 2) hipLaunchKernelGGL((k_adj_rows_grouped<512, 2, 1>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
        else if (!trans) hipLaunchKernelGGL((k_adj_rows_grouped<512, 3, 0>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
        else hipLaunchKernelGGL((k_adj_rows_grouped<512, 3, 1>), gr, bl, 0, st, (cAdjScen *)D, start, nscen);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Segments of a grouped adjoint scenario: column lines of n = nx-1 unknowns as the single
// scenario's column solves (8 unknowns per thread), row lines of ny as its LDS row solver
// with rb threads; 0 when the scenario does not fit.
int adj_cols_segments(int n) { return pick_segments(n, 8, 256); }
int adj_rows_segments(int n, int rb) { return pick_segments(n, RM, rb); }

// Block counts of the grouped stages for one scenario.
void grouped_stage_blocks(int nx, int ny, int rb, int *rows, int *cols, int *red, int *redp, int *S_rows, int *S_red,
                          int *L_red)
{
    *rows = nx - 1;
    *S_rows = pick_segments(ny, RM, rb);
    const int nseg = cols_segments(nx - 1);
    *cols = ((ny + 63) / 64) * ((nseg + 3) / 4);
    *S_red = pick_segments(2 * nseg, 8, 256);
    // 1024-thread workgroups of 1024/S columns (rows read in runs of 8*1024/S bytes), or
    // 256 threads when the scenario has fewer columns than such a workgroup holds (and at
    // most 64 segments). For a 2048 x 4096 beach 1024 threads take 25 us where 256 (four
    // columns per workgroup) took 46 us.
    const int B = (*S_red > 0 && *S_red <= 64 && ny < 1024 / *S_red) ? 256 : 1024;
    *L_red = *S_red > 0 ? B / *S_red : 0;
    if (B == 256) {
        *red = (ny + *L_red - 1) / *L_red;
        *redp = 0;
    } else {
        *red = 0;
        *redp = *L_red > 0 ? (ny + *L_red - 1) / *L_red : 0;
    }
}

hipError_t launch_rows_grouped(hipStream_t st, int which, int rb, const ScenDesc *D, const int *start, int nscen,
                               int blocks, int rev, bool vc_from_v)
{
    const dim3 g((unsigned)blocks);
    cScenDesc *Dc = (cScenDesc *)D;
    if (rb == 256) {
        if (which == 1) hipLaunchKernelGGL((k_rows_lds_grouped<1, 256, false>), g, dim3(256), 0, st, Dc, start, nscen, rev);
        else if (vc_from_v)
            hipLaunchKernelGGL((k_rows_lds_grouped<0, 256, true>), g, dim3(256), 0, st, Dc, start, nscen, rev);
        else hipLaunchKernelGGL((k_rows_lds_grouped<0, 256, false>), g, dim3(256), 0, st, Dc, start, nscen, rev);
    } else {
        if (which == 1) hipLaunchKernelGGL((k_rows_lds_grouped<1, 512, false>), g, dim3(512), 0, st, Dc, start, nscen, rev);
        else if (vc_from_v)
            hipLaunchKernelGGL((k_rows_lds_grouped<0, 512, true>), g, dim3(512), 0, st, Dc, start, nscen, rev);
        else hipLaunchKernelGGL((k_rows_lds_grouped<0, 512, false>), g, dim3(512), 0, st, Dc, start, nscen, rev);
    }
    return hipGetLastError();
}

hipError_t launch_cols3_grouped(hipStream_t st, int which, const ScenDesc *D, const int *start_cols,
                                const int *start_red, const int *start_redp, int nscen, int blocks_cols, int blocks_red,
                                int blocks_redp, bool backsub, int rev_elim, int rev_backsub, bool vc_from_v,
                                bool const_drag)
{
    const dim3 g((unsigned)blocks_cols), blk(64, 4);
    cScenDesc *Dc = (cScenDesc *)D;
    auto qxx = [&](auto vcf, auto cdc) {
        hipLaunchKernelGGL((k_cols_elim_grouped<0, decltype(vcf)::value, decltype(cdc)::value>), g, blk, 0, st, Dc,
                           start_cols, nscen, rev_elim);
    };
    using T = std::true_type;
    using F = std::false_type;
    if (which == 1) hipLaunchKernelGGL((k_cols_elim_grouped<1, false, false>), g, blk, 0, st, Dc, start_cols, nscen, rev_elim);
    else if (vc_from_v) const_drag ? qxx(T{}, T{}) : qxx(T{}, F{});
    else const_drag ? qxx(F{}, T{}) : qxx(F{}, F{});
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (blocks_red > 0) {   // scenarios whose reduced solve fits 256-thread workgroups
        hipLaunchKernelGGL(k_red_pcr_grouped<256>, dim3(blocks_red), dim3(256), 0, st, Dc, start_red, nscen);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (blocks_redp > 0) {  // the others, 1024-thread workgroups
        hipLaunchKernelGGL(k_red_pcr_grouped<1024>, dim3(blocks_redp), dim3(1024), 0, st, Dc, start_redp, nscen);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (!backsub) return hipSuccess;
    if (which == 0)
        hipLaunchKernelGGL(k_cols_backsub_grouped<0>, dim3(blocks_cols), dim3(64, 4), 0, st, Dc, start_cols, nscen,
                           rev_backsub);
    else hipLaunchKernelGGL(k_cols_backsub_grouped<1>, dim3(blocks_cols), dim3(64, 4), 0, st, Dc, start_cols, nscen,
                            rev_backsub);
    return hipGetLastError();
}

template <class Prov>
static hipError_t launch_rows_lds(hipStream_t st, const Prov &prov, double *Y, long ldy, int n, int nL, bool periodic,
                                  int rb, int rev)
{
    rb = rows_block_for(n, rb);
    const int S = pick_segments(n, RM, rb);
    if (S == 0) return hipErrorInvalidValue;
    if (rb == 256)
        hipLaunchKernelGGL((k_rows_lds<256, Prov>), dim3((unsigned)nL), dim3(256), 0, st, prov, Y, ldy, n, S,
                           periodic ? 1 : 0, rev);
    else if (rb == 512)
        hipLaunchKernelGGL((k_rows_lds<512, Prov>), dim3((unsigned)nL), dim3(512), 0, st, prov, Y, ldy, n, S,
                           periodic ? 1 : 0, rev);
    else
        hipLaunchKernelGGL((k_rows_lds<1024, Prov>), dim3((unsigned)nL), dim3(1024), 0, st, prov, Y, ldy, n, S,
                           periodic ? 1 : 0, rev);
    return hipGetLastError();
}

// Row sweeps with the assembly staged through LDS (one row per block).
template <bool STR>
static hipError_t launch_lds_qxy_bs_t(hipStream_t st, const Consts &k, const State &s, const RunFields &rf,
                                      const Work &w, const ColsWork &cw, int rb, bool vc_from_v, int rev)
{
    if (vc_from_v)
        return launch_rows_lds(st, ProvQxyBs<true, STR>{k, s, rf, w, cw}, s.Qx + k.ny, k.ny, k.ny, k.nx - 1, true, rb,
                               rev);
    return launch_rows_lds(st, ProvQxyBs<false, STR>{k, s, rf, w, cw}, s.Qx + k.ny, k.ny, k.ny, k.nx - 1, true, rb,
                           rev);
}
hipError_t launch_lds_qxy_bs(hipStream_t st, const Consts &k, const State &s, const RunFields &rf, const Work &w,
                             const ColsWork &cw, int rb, bool vc_from_v, int rev, bool stream)
{
    return stream ? launch_lds_qxy_bs_t<true>(st, k, s, rf, w, cw, rb, vc_from_v, rev)
                  : launch_lds_qxy_bs_t<false>(st, k, s, rf, w, cw, rb, vc_from_v, rev);
}
template <bool STR>
static hipError_t launch_lds_qyy_t(hipStream_t st, const Consts &k, const Metrics &m, const State &s,
                                   const RunFields &rf, const Work &w, int rb, int rev)
{
    if (k.drag_model == 0)
        return launch_rows_lds(st, ProvQyy<true, STR>{k, m, s, rf, w}, s.Qy + k.ny, k.ny, k.ny, k.nx - 1, true, rb, rev);
    return launch_rows_lds(st, ProvQyy<false, STR>{k, m, s, rf, w}, s.Qy + k.ny, k.ny, k.ny, k.nx - 1, true, rb, rev);
}
hipError_t launch_lds_qyy(hipStream_t st, const Consts &k, const Metrics &m, const State &s, const RunFields &rf,
                          const Work &w, int rb, int rev, bool stream)
{
    return stream ? launch_lds_qyy_t<true>(st, k, m, s, rf, w, rb, rev) : launch_lds_qyy_t<false>(st, k, m, s, rf, w, rb, rev);
}

}  // namespace m70
