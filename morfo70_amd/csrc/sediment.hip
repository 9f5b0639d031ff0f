// sediment.hip -- sediment transport and bed update (class Sediment, src/Sediment.cpp)
// for gfx950.
//
// Three kernels per flux evaluation and one per bed update, all on the hydro handle's
// stream and reading its device state:
//   k_sed_col_*    the running maximum of the breaking dissipation over x
//                  (getMaxWaveBreakingDissipation, :252-262) and the shoreline surface
//                  level scan (getShorelineZs, :264-283): per chunk of rows, then per
//                  column;
//   k_sed_coeffs_fluxes  per cell, by runs of rows: stirring, downslope and swash
//                  coefficients (:194-250), then the cross-shore (rows 0..nx-2) and
//                  alongshore fluxes (:156-192);
//   k_sed_update   per cell, rows 1..nx-1: Adams-Bashforth-2 bed update, moving the
//                  hydro's Zs with the bed (:105-151).
// Arithmetic keeps the reference's operation order (built with -ffp-contract=off), except
// that 1 + tanh(z) is evaluated through exp (one_plus_tanh below).
#include "hydro_device.h"
#include "sediment.h"

namespace m70 {

constexpr double PI_C = 3.141592653589793;   // include/OptionsParametersConstants.h:15
constexpr double S_REL = 2.65;               // :17
constexpr double TANPHI_I = 0.6249;          // :22

// The column pass is two reductions over x, so it runs as chunks of rows in parallel
// followed by a combine per column:
//  * the running maximum of Dw: max is exact and order-free;
//  * the shoreline level: the serial recurrence zs = (1-isSh) zs + Zs[i] isSh over
//    i = 1..nx-1 leaves Zs of the LAST row with isSh = 1 (isSh = not both rows i-1 and i
//    dry), or the previous value if no row has it; so each chunk finds its last such row
//    and the combine takes the largest. The rows after it add Zs[i]*0 = +-0, which changes
//    nothing but the sign of a zero result.
__global__ void __launch_bounds__(256) k_sed_col_partial(SedArgs a, int rows_per_chunk)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.ny) return;
    const int ny = a.ny;
    const int i0 = blockIdx.y * rows_per_chunk;
    const int i1 = min(i0 + rows_per_chunk, a.nx);
    const double Dm = a.Dmin + 1E-4;
    double mx = a.Dw[(long)i0 * ny + j];
    int last = -1;
    int dryLow = (i0 > 0) ? (a.D[(long)(i0 - 1) * ny + j] <= Dm) : 1;
    for (int i = i0; i < i1; ++i) {
        const long q = (long)i * ny + j;
        const double b = a.Dw[q];
        mx = (mx < b) ? b : mx;
        const int isDry = (a.D[q] <= Dm);
        if (i > 0 && !(isDry * dryLow)) last = i;
        dryLow = isDry;
    }
    a.part_max[(long)blockIdx.y * ny + j] = mx;
    a.part_last[(long)blockIdx.y * ny + j] = last;
}

// 64 columns per workgroup, the chunks split over 4 wavefronts, combined through LDS.
__global__ void __launch_bounds__(256) k_sed_col_final(SedArgs a, int chunks)
{
    __shared__ double smax[4][64];
    __shared__ int slast[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    const int ny = a.ny;
    double mx = -INFINITY;
    int last = -1;
    if (j < ny) {
        for (int c = wv; c < chunks; c += 4) {
            const double b = a.part_max[(long)c * ny + j];
            mx = (mx < b) ? b : mx;
            const int l = a.part_last[(long)c * ny + j];
            last = (l > last) ? l : last;
        }
    }
    smax[wv][lane] = mx;
    slast[wv][lane] = last;
    __syncthreads();
    if (wv == 0 && j < ny) {
        double m = a.maxDw[j];   // std::max(running, every row's Dw)
        for (int w = 0; w < 4; ++w) {
            const double b = smax[w][lane];
            m = (m < b) ? b : m;
            last = (slast[w][lane] > last) ? slast[w][lane] : last;
        }
        a.maxDw[j] = m;
        if (last >= 1) a.Zs_sh[j] = 0 * a.Zs_sh[j] + a.Zs[(long)last * ny + j] * 1;
    }
}

// 1 + tanh(z), the only form in which tanh enters the coefficients (src/Sediment.cpp:
// 228,252,279), evaluated as 2 / (1 + exp(-2z)): one exp and one division instead of a
// tanh (a third of the FP64 instructions of the kernel). Within a few ulps of 1 + tanh(z)
// in double; 0 and 2 at the saturated ends, as 1 + tanh gives there.
__device__ __forceinline__ double one_plus_tanh(double z) { return 2.0 / (1.0 + exp(-2.0 * z)); }

// The three coefficients of cell (i, j) (src/Sediment.cpp:194-250): stirring Alpha,
// downslope gamma, swash Gamma, from its depth and bed level and the column's shoreline
// level and running maximum of Dw.
struct SedCoef { double alpha, gamma, Gamma; };
__device__ __forceinline__ SedCoef sed_coef(const SedArgs &a, double d, double zb, double zs_sh, double max_dw)
{
    SedCoef c;
    {   // getWaveStirringCWS
        const double alpha0_min = a.p.alpha0_min;
        const double incrAlpha = a.p.alpha0_max - alpha0_min;
        const double Dmax = a.p.D_alpha0_max;
        const double Dmax_min = a.p.D_alpha0_min - Dmax;
        double x = 1 + 2 * (Dmax - d) / Dmax_min;
        c.alpha = alpha0_min + incrAlpha * 0.5 * one_plus_tanh(PI_C * x);
    }
    {   // getGammaDownslopeCWS (the EGU form)
        const double inv_tanPhi_i = 1 / TANPHI_I;
        const double gamma0_tanPhi = a.p.gamma0 * inv_tanPhi_i;
        const double Dmin = a.Dmin;
        const double A_swash = a.p.A_swash;
        const double u_bsl = one_plus_tanh(PI_C * (2 * (d - Dmin) / (A_swash - Dmin) - 1)) / 2;
        c.gamma = u_bsl * gamma0_tanPhi;
    }
    {   // getGammaSwash
        const double G0 = a.p.Gamma0;
        const double Dmin = a.Dmin + 1E-4;
        const double A_swash = a.p.A_swash;
        const double inv_rho_g_sm1 = 1 / (G * RHO * (S_REL - 1));
        const double _2_Doff_on = 2 / (A_swash - Dmin);
        const int isDry = (d <= Dmin);
        const double d_dry = (1 - isDry) * d + isDry * (fabs(zb + zs_sh));
        const double sf_swash = 1 - one_plus_tanh(PI_C * (_2_Doff_on * (d_dry - Dmin) - 1)) * 0.5;
        c.Gamma = G0 * inv_rho_g_sm1 * sf_swash * max_dw;
    }
    return c;
}

// Coefficients and fluxes in one pass (src/Sediment.cpp:56-65, 156-250): thread j takes a
// run of SED_RUN rows of column j. Each cell's coefficients are formed once and stored
// (they are public outputs); the qx flux of row i reads row i+1's from the run's registers,
// the qy flux reads column j+1's from lane l+1 (cross-lane, no memory access), and only
// the last lane of a wavefront and column ny-1 (whose right neighbour is column 0) form
// their right neighbour's themselves. The values are the ones the coefficient arrays would
// hold, so the fluxes are those of the two-pass form, bit for bit. Every lane of a
// wavefront holds the same rows (the runs are uniform per workgroup).
// Rows per run, measured at 2048 x 4096 (ms per sediment step, fluxes + bed update):
// 2 rows 0.341, 4 rows 0.327, 8 rows 0.331, 16 rows 0.351; the two-pass form took 0.370.
constexpr int SED_RUN = 4;
__global__ void __launch_bounds__(256) k_sed_coeffs_fluxes(SedArgs a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int ny = a.ny, nx = a.nx;
    const int i0 = blockIdx.y * SED_RUN;
    if (j >= ny || i0 >= nx) return;
    const int i1 = min(i0 + SED_RUN, nx);
    const int jp = wrap_inc(j, ny);
    const bool own_p = (threadIdx.x & 63) < 63 && j + 1 < ny;   // column j+1 is lane l+1's
    const double zs_sh = a.Zs_sh[j], max_dw = a.maxDw[j];
    const double zs_sh_p = own_p ? 0.0 : a.Zs_sh[jp], max_dw_p = own_p ? 0.0 : a.maxDw[jp];
    long q = (long)i0 * ny + j;
    SedCoef c0 = sed_coef(a, a.D[q], a.Zb[q], zs_sh, max_dw);
    double h0 = a.h[q];
    for (int i = i0; i < i1; ++i, q += ny) {
        const bool has_qx = i < nx - 1;
        const long qp = (long)i * ny + jp;
        // loads of this row first: a store ahead of them would hold them behind it
        double d1 = 0.0, zb1 = 0.0, h1 = 0.0, u1 = 0.0;
        if (has_qx) { d1 = a.D[q + ny]; zb1 = a.Zb[q + ny]; h1 = a.h[q + ny]; u1 = a.U[q + ny]; }
        const double v1 = a.V[q + ny];
        double dp = 0.0, zbp = 0.0, hp_own = 0.0;
        if (!own_p) { dp = a.D[qp]; zbp = a.Zb[qp]; hp_own = a.h[qp]; }
        // column j+1 of row i from lane l+1 (every lane takes part)
        const double al_l = __shfl_down(c0.alpha, 1), ga_l = __shfl_down(c0.gamma, 1);
        const double Ga_l = __shfl_down(c0.Gamma, 1), h_l = __shfl_down(h0, 1);
        SedCoef cp;
        double hp;
        if (own_p) { cp.alpha = al_l; cp.gamma = ga_l; cp.Gamma = Ga_l; hp = h_l; }
        else { cp = sed_coef(a, dp, zbp, zs_sh_p, max_dw_p); hp = hp_own; }
        SedCoef c1 = c0;
        double qx = 0.0;
        if (has_qx) {   // getqx: rows 0..nx-2, U row i+1
            c1 = sed_coef(a, d1, zb1, zs_sh, max_dw);
            const double inv_dx = a.inv_dx[i];
            double A_x = 0.5 * (c0.alpha + c1.alpha);
            double g_x = 0.5 * (c0.gamma + c1.gamma);
            double G_x = 0.5 * (c0.Gamma + c1.Gamma);
            double dh_dx = (h1 - h0) * inv_dx;
            qx = A_x * (u1 - g_x * dh_dx) - G_x * dh_dx;
        }
        double qy;
        {   // getqy: rows 0..nx-1, V row i+1, periodic in j
            double A_y = 0.5 * (c0.alpha + cp.alpha);
            double g_y = 0.5 * (c0.gamma + cp.gamma);
            double G_y = 0.5 * (c0.Gamma + cp.Gamma);
            double dh_dy = (hp - h0) * a.inv_dy;
            qy = A_y * (v1 - g_y * dh_dy) - G_y * dh_dy;
        }
        a.Alpha[q] = c0.alpha;
        a.gamma[q] = c0.gamma;
        a.Gamma[q] = c0.Gamma;
        if (has_qx) a.qx[q] = qx;
        a.qy[q] = qy;
        c0 = c1;
        h0 = h1;
    }
}

__global__ void __launch_bounds__(256) k_sed_update(SedArgs a, double dt_2, double _1_m1p)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y + 1;
    const int ny = a.ny;
    if (j >= ny || i >= a.nx) return;
    const long q = (long)i * ny + j;
    const double inv_dx = a.inv_dxx[i];
    double dqy_dy = a.inv_dy * (a.qy[q] - a.qy[(long)i * ny + wrap_dec(j, ny)]);
    double dqx_dx = inv_dx * (a.qx[q] - a.qx[q - ny]);   // qx row nx-1 stays zero
    double dh_dt = -_1_m1p * (dqx_dx + dqy_dy);
    double incr_h = dt_2 * (3 * dh_dt - a.dh_dt[q]);
    a.h[q] += incr_h;
    a.Zb[q] -= incr_h;
    a.Zs[q] += incr_h;
    a.dh_dt[q] = dh_dt;
}

__global__ void __launch_bounds__(256) k_sed_dhdt(SedArgs a, double _1_m1p)
{   // dh_dt from given fluxes (setSedimentFluxesFromResumeFile's arithmetic, :77-103)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y + 1;
    const int ny = a.ny;
    if (j >= ny || i >= a.nx) return;
    const long q = (long)i * ny + j;
    double dqy_dy = a.inv_dy * (a.qy[q] - a.qy[(long)i * ny + wrap_dec(j, ny)]);
    double dqx_dx = a.inv_dxx[i] * (a.qx[q] - a.qx[q - ny]);
    a.dh_dt[q] = -_1_m1p * (dqx_dx + dqy_dy);
}

static dim3 sed_grid(const SedArgs &a, int rows) { return dim3((unsigned)((a.ny + 255) / 256), (unsigned)rows); }

int sed_row_chunks(int nx, int *rows_per_chunk)
{
    int r = (nx + SED_MAX_CHUNKS - 1) / SED_MAX_CHUNKS;
    if (r < 8) r = 8;
    if (rows_per_chunk) *rows_per_chunk = r;
    return (nx + r - 1) / r;
}

hipError_t launch_sed_columns(hipStream_t st, const SedArgs &a)
{
    int r = 0;
    const int chunks = sed_row_chunks(a.nx, &r);
    const unsigned bx = (unsigned)((a.ny + 255) / 256);
    hipLaunchKernelGGL(k_sed_col_partial, dim3(bx, (unsigned)chunks), dim3(256), 0, st, a, r);
    hipLaunchKernelGGL(k_sed_col_final, dim3((unsigned)((a.ny + 63) / 64)), dim3(256), 0, st, a, chunks);
    return hipGetLastError();
}

hipError_t launch_sed_coeffs_fluxes(hipStream_t st, const SedArgs &a)
{
    hipLaunchKernelGGL(k_sed_coeffs_fluxes, sed_grid(a, (a.nx + SED_RUN - 1) / SED_RUN), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_sed_update(hipStream_t st, const SedArgs &a, double dt)
{
    const double _1_m1p = 1 / (1 - a.p.p);
    hipLaunchKernelGGL(k_sed_update, sed_grid(a, a.nx - 1), dim3(256), 0, st, a, 0.5 * dt, _1_m1p);
    return hipGetLastError();
}

hipError_t launch_sed_dhdt(hipStream_t st, const SedArgs &a)
{
    const double _1_m1p = 1 / (1 - a.p.p);
    hipLaunchKernelGGL(k_sed_dhdt, sed_grid(a, a.nx - 1), dim3(256), 0, st, a, _1_m1p);
    return hipGetLastError();
}

}  // namespace m70
