// tails.h -- per-cell bodies of the two elementwise tails of the Hydro ADI step (one
// column per thread), shared by the step's kernels (hydro_kernels.hip) and the grouped
// adjoint's taped step (adjoint_batch.hip).
#pragma once
#include "hydro_device.h"

namespace m70 {

// After the Qx projector: setQx_BC (src/HydroBoundaryConditions.cpp:13-23), difC2X of Qx
// (src/HydroQx.cpp:80), solveZs_guess (src/Hydro.cpp:230-250), getU0/getU
// (src/HydroQx.cpp:420-451), X2C (include/MeshOperations.h:36-47), and the Qy
// predictor's IsWater (setWaterNodes_y, src/HydroQy.cpp:92-122) into w.w. Rows 0..nx-1.
// guess_Zs enters only that mask and is not stored.
__host__ __device__ __forceinline__ double qx_at(const State &s, const Consts &k, long rowoff, int row, int j,
                                        double f0)
{
    return row == 0 ? k.f1 * s.Qx[k.ny + j] + f0 : s.Qx[rowoff + j];
}

__host__ __device__ __forceinline__ void qx_tail_body(const Consts &k, const Metrics &m, const State &s, const Work &w, double f0,
                                             int i, int j, bool write_uc = true)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    (void)r;
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
    const double gzs_i = s.Zs[r + j] + inc_i;
    {   // IsWater of face j: guess_D and guess_Zs of column j+1 by the same expressions
        const int jp = wrap_inc(j, ny);
        const double qa = qx_at(s, k, r, i, jp, f0), qb = qx_at(s, k, r + ny, i + 1, jp, f0);
        const double dq = m.inv_dxx[i] * (qb - qa);
        const double incrZs = -dt * (dq + s.dQy[r + jp]);
        const double gd_p = s.D[r + jp] + incrZs;
        const double gzs_p = s.Zs[r + jp] + incrZs;
        w.w[r + j] = (uint8_t)wet_rule(gd[1], gd_p, gzs_i, gzs_p, s.dQy[r + j], s.dQy[r + jp], k.Dmask);
    }
    s.dQx[r + j] = dq_i;
    w.gD[r + j] = gd[1];
    s.U[r + j] = ui;
    if (write_uc) s.Uc[r + j] = m.x2c_d[i] * ui + m.x2c_up[i] * ui1;
    if (i == 0) s.Qx[j] = q[1];
}

// After the Qy projector: setQy_BC (src/HydroBoundaryConditions.cpp:62-79), difY2C
// (src/HydroMeshOperations.cpp:148-159), solveZs (src/Hydro.cpp:253-271), getV/getV0
// (src/HydroQy.cpp:458-487), Y2C (src/HydroMeshOperations.cpp:88-97). Rows 0..nx-1.
// Reads D, writes the new depth to Dn.
__host__ __device__ __forceinline__ double qy_row(const State &s, const Consts &k, int row, int j)
{   // Qy row `row` in 1..nx after the BC: row nx is Qy[nx-1]*inv_3
    return row < k.nx ? s.Qy[(long)row * k.ny + j] : s.Qy[(long)(k.nx - 1) * k.ny + j] * k.inv_3;
}

__host__ __device__ __forceinline__ double dqy_at(const State &s, const Consts &k, int i, int j)
{
    const int jm = wrap_dec(j, k.ny);
    return (qy_row(s, k, i + 1, j) - qy_row(s, k, i + 1, jm)) * k.inv_dy;
}

__host__ __device__ __forceinline__ double new_depth(const State &s, const Consts &k, int i, int j)
{
    const long q = (long)i * k.ny + j;
    const double incrZs = -k.dt * (s.dQx[q] + dqy_at(s, k, i, j));
    return s.D[q] + incrZs;
}

__host__ __device__ __forceinline__ void qy_tail_body(const Consts &k, const State &s, int i, int j, bool write_vc = true)
{
    const int ny = k.ny;
    const long r = (long)i * ny;
    (void)r;
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
    if (write_vc) s.Vc[r + j] = 0.5 * (v_j + v_m);
    if (i == 0) {
        const double qy0 = k.bcQy * s.Qy[ny + j];
        const double d_y0 = 2 * dn - new_depth(s, k, 1, j);
        s.Qy[j] = qy0;
        s.V[j] = qy0 / d_y0;
    }
    if (i == nx - 1) s.Qy[(long)nx * ny + j] = s.Qy[(long)(nx - 1) * ny + j] * k.inv_3;
}

}  // namespace m70
