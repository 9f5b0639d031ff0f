// hydro_device.h -- device-side data structures and helpers of the Hydro ADI step.
//
// Every arithmetic expression in the kernels keeps the operand order of the reference's
// C++ (src/HydroQx.cpp, src/HydroQy.cpp, src/Hydro.cpp, src/HydroParametrizations.cpp,
// src/HydroMeshOperations.cpp, src/HydroBoundaryConditions.cpp) and the library is built
// with -ffp-contract=off, so each value is the IEEE double the reference computes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace m70 {

// include/OptionsParametersConstants.h:14-19
constexpr double G = 9.8100000000000000;
constexpr double RHO = 1025.00000000000;
constexpr double SQRT_2 = 1.414213562373095;

// Per-row grid metrics (src/Grid.cpp:32-92) plus the per-row reciprocals the reference
// forms inline (1/dx[i] at src/HydroQx.cpp:105, 1/dx_x[i] at src/HydroQy.cpp:129,266 and
// src/HydroMeshOperations.cpp:106, inv_rho/dx[i] at src/HydroQx.cpp:372).
struct Metrics {
    const double *inv_dx;     // [nx-1]
    const double *inv_dxx;    // [nx]
    const double *inv_rhodx;  // [nx-1]
    const double *difx_x_up, *difx_x_d, *difx_x_low;     // [nx-1]
    const double *dif2x_x_up, *dif2x_x_d, *dif2x_x_low;  // [nx-1]
    const double *difx_c_up, *difx_c_d, *difx_c_low;     // [nx]
    const double *dif2x_c_up, *dif2x_c_d, *dif2x_c_low;  // [nx]
    const double *x2c_up, *x2c_d;                        // [nx]
};

// Scalars. Each derived constant is formed on the host with the expression the reference
// uses at the cited line, so it holds the same bits.
struct Consts {
    int nx, ny;
    int drag_model, visc_model;
    double dt;
    double inv_dt;        // 1/dt                       src/HydroQx.cpp:103
    double inv_dy;        // 1/dy                       src/HydroQx.cpp:104
    double inv_2dy;       // 0.5/dy                     src/HydroQx.cpp:301
    double inv_dy2_xy;    // 4*inv_2dy*inv_2dy          src/HydroQx.cpp:302
    double inv_dy2_yy;    // inv_dy*inv_dy              src/HydroQy.cpp:187
    double two_inv_dy2;   // 2/(dy*dy)                  src/HydroQx.cpp:282
    double alpha_f2;      // alpha_f*alpha_f            src/HydroParametrizations.cpp:20
    double inv_sqrt2;     // 1/sqrt_2                   src/HydroParametrizations.cpp:22
    double Dmask;         // Dmin+1E-4                  src/HydroQx.cpp:236
    double f1;            // bcOffshoreQx_f1            src/Hydro.cpp:51,55
    double bcQy;          // bcOffshoreQy               src/Hydro.cpp:60
    double inv_3;         // 1.0/3.0                    src/HydroBoundaryConditions.cpp:54
    double inv_rho;       // 1/rho                      src/HydroQx.cpp:360
    double inv_4dy_rho;   // (0.25/dy)*inv_rho          src/HydroQx.cpp:361-362
    double inv_dy_rho;    // (1/dy)*inv_rho             src/HydroQy.cpp:372-373
    double inv_2dy_c2c;   // 0.5/dy                     src/HydroMeshOperations.cpp:204
    double cD0, nu0, M, ka, z0, alpha_r;
    float inv_rho_f;      // float 1/rho                src/HydroParametrizations.cpp:107
};

struct Forcing {
    const double *Sxx, *Sxy, *Syy, *Urms, *Dw, *D_r, *H;
};

// Fields computed once per Run (src/Hydro.cpp:175-192).
struct RunFields {
    double *cD, *nuD;                  // F
    double *Fx, *Urms_x, *cD_x, *nuD_x, *dnuD_dx, *dnuDx_dy;   // rows 0..nx-2
    double *Fy, *Urms_y, *cD_y, *nuD_y, *dnuD_dy, *dnuDy_dx;   // rows 0..nx-1
};

// State (include/Hydro.h:26-34 plus Bathymetry::D). D is double-buffered: the Qy tail
// reads D and writes Dn, then the host swaps them.
struct State {
    double *Zs, *D, *Dn, *Qx, *Qy, *U, *V, *Uc, *Vc, *dQx, *dQy;
};

// Per-step work arrays.
struct Work {
    double *a, *b, *c, *f;     // tridiagonal coefficients (diag, low, up, rhs), row-major F
    double *gD;                // guess_D (src/Hydro.cpp:230-250); guess_Zs is used only for the
                               // Qy predictor's IsWater, which the Qx tail forms, so it is not kept
    double *drag;              // DragTerm of the predictor, reused by the projector
    uint8_t *w;                // IsWater of the predictor, reused by the projector
};

// Work arrays of the three-pass column solver: eliminated segments E0..E2 (F each: the
// fill-ins on the segment's first and last unknowns and the right-hand side, each times
// the inverse pivot), reduced rows R0..R3 and their solution X (2*nseg rows of ny each).
struct ColsWork {
    double *E0, *E1, *E2;
    double *R0, *R1, *R2, *R3;
    double *X;
    int nseg;
};

// One scenario of a grouped (multi-scenario) launch: everything a stage needs, read by
// the blocks of that scenario from a device-side table.
struct ScenDesc {
    Consts k;
    Metrics m;
    State s;
    RunFields rf;
    Forcing fw;
    Work w;
    ColsWork cw;
    double f0;            // offshore offset when the scenario has no series
    const double *f0s;    // the scenario's f0 series on the device, or NULL
    long f0_n;            // series length
    long f0_base;         // series index of the batch run's first step
    const int *step;      // steps done in the current batch run (one counter per batch)
    int S_rows;   // segments per row line in the LDS row solver
    int S_red;    // segments per reduced column system
    int L_red;    // reduced column systems per workgroup
    int pad;
};

// Grouped kernels take their ScenDesc table through the constant address space: pointers
// loaded from it then count as global ones, so loads through them are global loads and the
// per-row metrics, uniform in a wavefront, scalar loads. Through a generic table pointer
// they were flat vector loads (194 of them per wavefront in the Qx predictor, against 121
// global loads in the single-scenario kernel).
typedef __attribute__((address_space(4))) const ScenDesc cScenDesc;

// Offshore offset of the scenario's current step: series[min(base + step, n - 1)], the
// value m70h_run would use at the same step.
__device__ __forceinline__ double scen_f0(const ScenDesc &d)
{
    if (!d.f0s) return d.f0;
    long q = d.f0_base + *d.step;
    if (q > d.f0_n - 1) q = d.f0_n - 1;
    return d.f0s[q];
}

// Stages of a grouped step; start[stage][s] is the first block of scenario s.
// GS_RED / GS_REDP: reduced column systems in 256- / 1024-thread workgroups; GS_TX / GS_TY:
// the Qx and Qy tails.
enum GroupStage { GS_CELLS = 0, GS_ROWS = 1, GS_COLS = 2, GS_RED = 3, GS_REDP = 4, GS_TX = 5, GS_TY = 6, GS_COUNT = 7 };

// Scenario owning block b: start[] has nscen+1 ascending entries, start[0] = 0.
__device__ __forceinline__ int scen_of(const int *__restrict__ start, int nscen, int b)
{
    int lo = 0, hi = nscen - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8),
// and each XCD has its own L2. This bijection on [0, N) gives XCD x one contiguous range of
// logical tiles, so tiles that share lines run on the same XCD and hit its L2. Used by the
// reduced-system column solver; on the streaming row and column kernels it cut fetched
// bytes by 5-20% but made them slower (they are latency-bound, and eight far-apart
// streams of rows spread the DRAM pages), so they keep the plain order.
__host__ __device__ __forceinline__ int xcd_chunk(int b, int N)
{
    const int q = N >> 3, r = N & 7, x = b & 7, l = b >> 3;
    return x * q + (x < r ? x : r) + l;
}

// Cache policy of the FUSED step's streamed accesses, per kernel ("site"). A non-temporal
// access sets the nt bit of gfx950's global load or store instruction. On this GPU a
// streaming kernel that reads ten arrays and writes four moves 5.2 TB/s with plain
// accesses, 6.2 TB/s with non-temporal loads, and one that reads four and writes four goes
// from 5.3 to 7.3 TB/s with non-temporal stores. In the step, a field that another lane,
// row or workgroup reads again (stencil neighbours, the rows a column segment shares with
// the next one) must stay a plain load: made non-temporal, its second reader misses and
// the step got 11% slower. So: every store of the step is non-temporal; loads are where
// nothing reads the line again (the single-use fields of each sweep, the reduced systems,
// the back-substitution's eliminated segments). Measured at 2048 x 4096: 0.956 -> 0.909 ms
// per step, results bitwise unchanged. Bit s of M70_NTL / M70_NTS selects site s; a build
// may override the measured defaults.
enum NtSite {
    NT_NONE = -1,
    NT_QXX = 0,        // Qx predictor column elimination (every load)
    NT_QYX = 1,        // Qy projector column elimination (every load)
    NT_BACKSUB = 2,    // column back-substitution
    NT_QXY_ROWS = 3,   // Qx projector rows (every load)
    NT_QYY_ROWS = 4,   // Qy predictor rows (every load)
    NT_QX_TAIL = 5,
    NT_QY_TAIL = 6,
    NT_REDUCED = 7,    // reduced column systems
    // loads of the fields a kernel reads at one point only (no neighbouring lane, row or
    // workgroup reads them again)
    NT_QXX_STREAM = 8,
    NT_QYY_STREAM = 9,
    NT_QXY_STREAM = 10,
    NT_QYX_STREAM = 11,
};
#ifndef M70_NTL
#define M70_NTL ((1 << NT_BACKSUB) | (1 << NT_REDUCED) | (1 << NT_QXX_STREAM) | (1 << NT_QYY_STREAM) | \
                 (1 << NT_QXY_STREAM) | (1 << NT_QYX_STREAM))
#endif
#ifndef M70_NTS
#define M70_NTS 0xff
#endif
// STR: the kernel runs under the streaming policy (stream_policy), otherwise every access
// is plain.
template <int SITE, bool STR> struct Nt {
    static constexpr bool L = STR && SITE >= 0 && ((M70_NTL >> (SITE < 0 ? 0 : SITE)) & 1);
    static constexpr bool S = STR && SITE >= 0 && ((M70_NTS >> (SITE < 0 ? 0 : SITE)) & 1);
};
// A grid streams when one kernel's traffic (about 120 B per cell) outgrows the 256 MB MALL
// behind the L2s: up to then a producer's plain stores leave its output in the MALL for
// the kernel after it, which non-temporal stores would forgo. Measured step times, plain
// against streaming policy (ms): 768 x 1024 0.139 / 0.152, 1024 x 1024 0.160 / 0.171,
// 1024 x 2048 0.276 / 0.275, 2048 x 2048 0.517 / 0.502, 2048 x 4096 0.955 / 0.911.
// Grouped launches over many scenarios take the streaming policy, together with 1024-thread
// workgroups for their reduced column systems (grouped_stage_blocks): 64 uneven beaches
// 7.39e9 -> 7.42e9 cell-steps/s, one 2048 x 4096 beach 1.145 -> 1.036 ms per step. With
// 256-thread reduced-system workgroups the streaming policy made the 64 beaches 1.5% slower.
constexpr long STREAM_MIN_CELLS = 2L << 20;
constexpr bool GROUPED_STREAM = true;
__host__ __device__ __forceinline__ bool stream_policy(int nx, int ny) { return (long)nx * ny >= STREAM_MIN_CELLS; }

typedef double m70_v2d __attribute__((ext_vector_type(2)));

template <bool NT> __host__ __device__ __forceinline__ double ldn(const double *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (NT) return __builtin_nontemporal_load(p);
#endif
    return *p;
}
template <bool NT> __host__ __device__ __forceinline__ uint8_t ldn(const uint8_t *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (NT) return __builtin_nontemporal_load(p);
#endif
    return *p;
}
template <bool NT> __host__ __device__ __forceinline__ void stn(double *p, double v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (NT) { __builtin_nontemporal_store(v, p); return; }
#endif
    *p = v;
}
template <bool NT> __host__ __device__ __forceinline__ void stn(uint8_t *p, uint8_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (NT) { __builtin_nontemporal_store(v, p); return; }
#endif
    *p = v;
}
// Two adjacent doubles at a 16-byte aligned address.
template <bool NT> __host__ __device__ __forceinline__ void ldn2(const double *p, double &a, double &b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const m70_v2d v = NT ? __builtin_nontemporal_load(reinterpret_cast<const m70_v2d *>(p))
                         : *reinterpret_cast<const m70_v2d *>(p);
    a = v.x;
    b = v.y;
#else
    a = p[0];
    b = p[1];
#endif
}
template <bool NT> __host__ __device__ __forceinline__ void stn2(double *p, double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    m70_v2d v;
    v.x = a;
    v.y = b;
    if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<m70_v2d *>(p));
    else *reinterpret_cast<m70_v2d *>(p) = v;
#else
    p[0] = a;
    p[1] = b;
#endif
}

__host__ __device__ __forceinline__ int wrap_inc(int j, int ny) { return j + 1 == ny ? 0 : j + 1; }
__host__ __device__ __forceinline__ int wrap_dec(int j, int ny) { return j == 0 ? ny - 1 : j - 1; }

// Body of Hydro::setWaterNodes_x/_y (src/HydroQx.cpp:245-253, src/HydroQy.cpp:103-109):
// integer arithmetic on the six comparisons, then conversion to bool.
__host__ __device__ __forceinline__ int wet_rule(double d, double dup, double zs, double zsup, double s,
                                        double sup, double Dmin)
{
    int dup_gt_min = dup > Dmin;
    int d_gt_min = d > Dmin;
    int zsup_gt_zs = zsup > zs;
    int zs_gt_zsup = zs > zsup;
    int sink = s > 0;
    int sinkup = sup > 0;
    return (d_gt_min * dup_gt_min +
            (dup_gt_min * (!d_gt_min) * zsup_gt_zs + d_gt_min * (!dup_gt_min) * zs_gt_zsup) * !(sink + sinkup)) != 0;
}

// Hydro::calculateDragTerm, src/HydroParametrizations.cpp:19-28 (one element).
__host__ __device__ __forceinline__ double drag_term(double cd, double urms, double u, double v, double inv_d,
                                            const Consts &k)
{
    return cd * urms * k.inv_sqrt2 * sqrt(k.alpha_f2 + 2 * (v * v + u * u) / (urms * urms)) * inv_d;
}

}  // namespace m70
