// m70h_api.hip -- C ABI of the Hydro ADI step (include/m70h.h): handle, device buffers,
// the per-Run pre-loop and the per-step launch sequence of Hydro::Run
// (src/Hydro.cpp:170-202).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/m70h.h"
#include "adjoint_batch.h"
#include "adjoint_kernels.h"
#include "hydro_device.h"
#include "hydro_kernels.h"
#include "sediment.h"
#include "netcdf3.h"
#include "outer.h"
#include "wave.h"

using namespace m70;

struct AdjCtx;

namespace {

thread_local std::string g_create_error;

enum KernelId {
    K_DRAG_VISC, K_RUN_FIELDS, K_QXX_ASM, K_THOMAS_QX, K_QXY_ASM, K_PERIODIC_QX, K_QX_TAIL,
    K_QYY_ASM, K_PERIODIC_QY, K_QYX_ASM, K_THOMAS_QY, K_QY_TAIL, K_FUSED_QXY, K_FUSED_QYY,
    K_COLS_QX_ELIM, K_COLS_QX_RED, K_COLS_QX_BACK, K_COLS_QY_ELIM, K_COLS_QY_RED, K_COLS_QY_BACK,
    K_SED_COLUMNS, K_SED_FLUXES, K_SED_UPDATE, K_COUNT
};
const char *const kKernelNames[K_COUNT] = {
    "drag_viscosity", "run_fields", "qxx_assemble", "thomas_qx", "qxy_assemble", "periodic_qx", "qx_tail",
    "qyy_assemble", "periodic_qy", "qyx_assemble", "thomas_qy", "qy_tail", "fused_qxy", "fused_qyy",
    "cols_qx_elim", "cols_qx_reduced", "cols_qx_backsub", "cols_qy_elim", "cols_qy_reduced", "cols_qy_backsub",
    "sed_columns", "sed_fluxes", "sed_update"};

}  // namespace

struct m70h {
    int nx = 0, ny = 0;
    double dy = 0.0;
    long F = 0, Fp = 0;
    m70h_params p{};
    Consts k{};
    double f0 = 0.0;
    std::vector<double> f0_series;
    long f0_pos = 0;
    int sweep_rev = 0;   // row order of the next FUSED step's first kernel (run_step)
    int device = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    int solver = M70H_SOLVER_EXACT;
    int tile_cols = 0, tile_rows = 0;   // partition-solver unknowns per thread (0 = default)
    int rows_block = 0;                 // FUSED row-solver block size (0 = smallest that fits)
    std::string err;
    std::vector<void *> allocs;
    double *metric_buf = nullptr;
    Metrics m{};
    State s{};
    Forcing fw{};
    double *fw_buf[7] = {};
    bool forcing_set = false;
    RunFields rf{};
    Work w{};
    double *sr = nullptr, *sv = nullptr, *sk = nullptr;
    ColsWork cw{};
    int *flag = nullptr;
    int *run_flag = nullptr;      // device: a Run's non-finite flag
    int *run_flag_host = nullptr; // pinned copy of it, read by m70h_synchronize
    bool run_check_pending = false;
    double *f0_dev = nullptr;     // device copy of f0_series (read by batched steps)
    long f0_dev_n = 0;
    AdjCtx *adj = nullptr;   // checkpoints and work of the adjoint (m70h_adjoint_*)
    long run_gen = 0;        // bumped by every Run and forcing change
    // profiling
    bool prof = false;
    int prof_only = -1;   // kernel id to time alone, -1 for every launch
    struct Rec { int kid; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double prof_ms[K_COUNT] = {};
    long prof_n[K_COUNT] = {};
};

namespace {

int fail(m70h *h, int code, const std::string &msg)
{
    if (h) h->err = msg; else g_create_error = msg;
    return code;
}

int hip_fail(m70h *h, hipError_t e, const char *what)
{
    return fail(h, M70H_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(h, expr)                                   \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

template <typename T>
hipError_t dalloc(m70h *h, T **ptr, size_t count)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 64);
    if (e != hipSuccess) return e;
    h->allocs.push_back(p);
    *ptr = (T *)p;
    return hipMemset(p, 0, count * sizeof(T) + 64);
}

hipEvent_t next_event(m70h *h)
{
    if (h->ev_used == h->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->ev_pool.push_back(e);
    }
    return h->ev_pool[h->ev_used++];
}

// Runs one launch, bracketed by events when profiling is on.
template <typename Fn>
int timed(m70h *h, int kid, Fn &&fn)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (h->prof && (h->prof_only < 0 || h->prof_only == kid)) {
        a = next_event(h);
        b = next_event(h);
        if (a && b) (void)hipEventRecord(a, h->stream);
    }
    hipError_t e = fn();
    if (e != hipSuccess) return hip_fail(h, e, kKernelNames[kid]);
    if (a && b) {
        (void)hipEventRecord(b, h->stream);
        h->recs.push_back({kid, a, b});
    }
    return M70H_OK;
}

int collect_profile(m70h *h)
{
    if (h->recs.empty()) return M70H_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (auto &r : h->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            h->prof_ms[r.kid] += ms;
            h->prof_n[r.kid] += 1;
        }
    }
    h->recs.clear();
    h->ev_used = 0;
    return M70H_OK;
}

double f0_for_step(m70h *h)
{
    if (!h->f0_series.empty()) {
        long q = h->f0_pos < (long)h->f0_series.size() ? h->f0_pos : (long)h->f0_series.size() - 1;
        return h->f0_series[q];
    }
    return h->f0;
}

// Per-Run pre-loop, src/Hydro.cpp:175-192.
int run_preloop(m70h *h)
{
    int rc = timed(h, K_DRAG_VISC, [&] { return launch_drag_viscosity(h->stream, h->k, h->s.D, h->fw, h->rf); });
    if (rc) return rc;
    return timed(h, K_RUN_FIELDS, [&] { return launch_run_fields(h->stream, h->k, h->m, h->fw, h->rf); });
}

// One ADI step: solveQx (src/HydroQx.cpp:18-93) then solveQy (src/HydroQy.cpp:17-85).
// last: the final step of a Run, which leaves every public state field in memory.
// first: the first step of a Run, which reads V_c as the caller left it; the FUSED steps
// after it form V_c from V, so only a Run's last step stores V_c and U_c.
int run_step(m70h *h, bool last = true, bool first = true)
{
    const double f0 = f0_for_step(h);
    h->f0_pos++;
    const int nx = h->nx, ny = h->ny;
    const Consts &k = h->k;
    State &s = h->s;
    Work &w = h->w;
    int rc;
    hipStream_t st = h->stream;
    if (h->solver == M70H_SOLVER_FUSED) {
        // x-sweeps: three coalesced passes (assemble+eliminate, reduced solve, back-sub);
        // y-sweeps: one launch each, the row's coefficients staged through LDS. The Qx
        // projector forms the Qx predictor's solution itself from the first two passes.
        // Sweep order: each kernel takes the grid rows in the order opposite to the
        // kernel before it, so it starts on the rows that kernel wrote last, which are
        // still in the 256 MB MALL. Seven kernels of a step have a row order, so the
        // order flips from step to step (sweep_rev). Results do not depend on it.
        const int a = h->sweep_rev, b = 1 - a;
        h->sweep_rev = b;
        // cache policy by grid size: large grids stream (non-temporal accesses, see
        // stream_policy), grids whose kernels fit the MALL keep plain ones
        const bool str = stream_policy(nx, ny);
        for (int pass = COLS_ELIM; pass <= COLS_REDUCED; ++pass)
            if ((rc = timed(h, K_COLS_QX_ELIM + pass, [&] {
                     return launch_cols3_qxx(st, k, h->m, s, h->rf, w, f0, h->cw, h->tile_cols, pass, !first, a, str);
                 })))
                return rc;
        if ((rc = timed(h, K_FUSED_QXY,
                        [&] { return launch_lds_qxy_bs(st, k, s, h->rf, w, h->cw, h->rows_block, !first, b, str); })))
            return rc;
        if ((rc = timed(h, K_QX_TAIL, [&] { return launch_qx_tail(st, k, h->m, s, w, f0, last, a, str); }))) return rc;
        if ((rc = timed(h, K_FUSED_QYY,
                        [&] { return launch_lds_qyy(st, k, h->m, s, h->rf, w, h->rows_block, b, str); })))
            return rc;
        for (int pass = COLS_ELIM; pass <= COLS_BACKSUB; ++pass)
            if ((rc = timed(h, K_COLS_QY_ELIM + pass, [&] {
                     return launch_cols3_qyx(st, k, h->m, s, h->rf, w, h->cw, h->tile_cols, pass,
                                             pass == COLS_BACKSUB ? b : a, str);
                 })))
                return rc;
        if ((rc = timed(h, K_QY_TAIL, [&] { return launch_qy_tail(st, k, s, last, a, str); }))) return rc;
        std::swap(s.D, s.Dn);
        return M70H_OK;
    }
    // Qx predictor: assembly + Thomas along x (src/HydroQx.cpp:29-51)
    if ((rc = timed(h, K_QXX_ASM, [&] { return launch_qxx_assemble(st, k, h->m, s, h->rf, w, f0); }))) return rc;
    const bool fast = h->solver == M70H_SOLVER_PARTITION;
    if ((rc = timed(h, K_THOMAS_QX, [&] {
             return fast ? launch_partition_cols(st, w.a, w.b, w.c, w.f, s.Qx + ny, nx - 1, ny, ny, h->tile_cols)
                         : launch_thomas_cols(st, w.a, w.b, w.c, w.f, s.Qx + ny, h->sv, nx - 1, ny, ny);
         })))
        return rc;
    // Qx projector: assembly + periodic solve along y per row (src/HydroQx.cpp:59-70)
    if ((rc = timed(h, K_QXY_ASM, [&] { return launch_qxy_assemble(st, k, s, h->rf, w); }))) return rc;
    if ((rc = timed(h, K_PERIODIC_QX, [&] {
             return fast ? launch_partition_rows(st, w.a, w.b, w.c, w.f, ny, s.Qx + ny, ny, nx - 1, ny, true, h->tile_rows)
                         : launch_periodic_rows(st, w.a, w.b, w.c, w.f, ny, s.Qx + ny, ny, h->sr, h->sv, h->sk,
                                                nx - 1, ny);
         })))
        return rc;
    // BC, dQx_dx, guess, U, U_c (src/HydroQx.cpp:76-92)
    if ((rc = timed(h, K_QX_TAIL, [&] { return launch_qx_tail(st, k, h->m, s, w, f0); }))) return rc;
    // Qy predictor (src/HydroQy.cpp:26-42)
    if ((rc = timed(h, K_QYY_ASM, [&] { return launch_qyy_assemble(st, k, h->m, s, h->rf, w); }))) return rc;
    if ((rc = timed(h, K_PERIODIC_QY, [&] {
             return fast ? launch_partition_rows(st, w.a, w.b, w.c, w.f, ny, s.Qy + ny, ny, nx - 1, ny, true, h->tile_rows)
                         : launch_periodic_rows(st, w.a, w.b, w.c, w.f, ny, s.Qy + ny, ny, h->sr, h->sv, h->sk,
                                                nx - 1, ny);
         })))
        return rc;
    // Qy projector (src/HydroQy.cpp:52-67)
    if ((rc = timed(h, K_QYX_ASM, [&] { return launch_qyx_assemble(st, k, h->m, s, h->rf, w); }))) return rc;
    if ((rc = timed(h, K_THOMAS_QY, [&] {
             return fast ? launch_partition_cols(st, w.a, w.b, w.c, w.f, s.Qy + ny, nx - 1, ny, ny, h->tile_cols)
                         : launch_thomas_cols(st, w.a, w.b, w.c, w.f, s.Qy + ny, h->sv, nx - 1, ny, ny);
         })))
        return rc;
    // BC, dQy_dy, Zs, D, V, V_c (src/HydroQy.cpp:73-84)
    if ((rc = timed(h, K_QY_TAIL, [&] { return launch_qy_tail(st, k, s); }))) return rc;
    std::swap(s.D, s.Dn);
    return M70H_OK;
}

// Scalars of the kernels, each formed with the reference's expression (see Consts).
// The boundary constants depend on the offshore option at construction only
// (src/Hydro.cpp:44-60), as in the reference.
void set_consts(m70h *h, const m70h_params &pp, double dx0, double dy, bool with_bc)
{
    const m70h_params *p = &pp;
    const int nx = h->nx, ny = h->ny;
    const double inv_rho = 1 / RHO;
    Consts &k = h->k;
    h->p = pp;
    k.nx = nx; k.ny = ny;
    k.drag_model = p->drag_model; k.visc_model = p->visc_model;
    k.dt = p->dt;
    k.inv_dt = 1 / p->dt;
    k.inv_dy = 1 / dy;
    k.inv_2dy = 0.5 / dy;
    k.inv_dy2_xy = 4 * k.inv_2dy * k.inv_2dy;
    k.inv_dy2_yy = k.inv_dy * k.inv_dy;
    k.two_inv_dy2 = 2 / (dy * dy);
    k.alpha_f2 = (p->alpha_f * p->alpha_f);
    k.inv_sqrt2 = 1 / SQRT_2;
    k.Dmask = p->Dmin + 1E-4;
    if (with_bc) {
        const double lSponge = dx0 * 5;
        const double exp_dx_2ls = std::exp(-0.5 * dx0 / lSponge);
        const double exp_dx_ls = std::exp(-dx0 / lSponge);
        k.f1 = (p->offshore_bc == M70H_BC_SPONGE) ? 0.5 * exp_dx_2ls / (1 - 0.5 * exp_dx_2ls) : 1;
        k.bcQy = exp_dx_ls;
    }
    k.inv_3 = 1.0 / 3.0;
    k.inv_rho = inv_rho;
    {
        double inv_4dy = 0.25 / dy;
        inv_4dy *= inv_rho;
        k.inv_4dy_rho = inv_4dy;
        double inv_dyr = 1 / dy;
        inv_dyr *= inv_rho;
        k.inv_dy_rho = inv_dyr;
    }
    k.inv_2dy_c2c = 0.5 / dy;
    k.cD0 = p->cD0; k.nu0 = p->nu0; k.M = p->M; k.ka = p->ka; k.z0 = p->z0; k.alpha_r = p->alpha_r;
    k.inv_rho_f = (float)(1 / RHO);
}

// Grid metrics as Grid::Grid (src/Grid.cpp:32-92, closed forms :127-145).
void grid_metrics(int nx, const double *x, std::vector<double> &dx, std::vector<double> &dx_x,
                  std::vector<double> (&xx)[6], std::vector<double> (&cc)[6], std::vector<double> &x2c_up,
                  std::vector<double> &x2c_d)
{
    auto dif = [](double dx0, double dx1, double &up, double &d, double &low) {
        double dx02 = dx0 * dx0;
        double dx12 = dx1 * dx1;
        double dx0dx1 = dx0 * dx1;
        double dx0_p_dx1 = dx0 + dx1;
        d = (dx12 - dx02) / (dx0_p_dx1 * dx0dx1);
        up = dx02 / (dx0_p_dx1 * dx0dx1);
        low = -dx12 / (dx0_p_dx1 * dx0dx1);
    };
    auto dif2 = [](double dx0, double dx1, double &up, double &d, double &low) {
        double dx0dx1 = dx0 * dx1;
        double dx0_p_dx1 = dx0 + dx1;
        d = -2 / (dx0dx1);
        up = 2 / (dx0_p_dx1 * dx1);
        low = 2 / (dx0_p_dx1 * dx0);
    };
    dx.assign(nx - 1, 0.0);
    dx_x.assign(nx, 0.0);
    for (auto &v : xx) v.assign(nx - 1, 0.0);
    for (auto &v : cc) v.assign(nx, 0.0);
    x2c_up.assign(nx, 0.0);
    x2c_d.assign(nx, 0.0);
    for (int i = 0; i < nx - 1; i++) dx[i] = x[i + 1] - x[i];
    dx_x[0] = dx[0];
    for (int i = 1; i < nx - 1; i++) dx_x[i] = (dx[i] + dx[i - 1]) / 2;
    dx_x[nx - 1] = dx[nx - 2];
    // xx: up, d, low, up2, d2, low2 at x-faces; cc: same at centres
    for (int i = 0; i < nx - 1; i++) {
        dif(dx_x[i], dx_x[i + 1], xx[0][i], xx[1][i], xx[2][i]);
        dif2(dx_x[i], dx_x[i + 1], xx[3][i], xx[4][i], xx[5][i]);
    }
    dif(dx[0], dx[1], cc[0][0], cc[1][0], cc[2][0]);
    dif2(dx[0], dx[1], cc[3][0], cc[4][0], cc[5][0]);
    for (int i = 1; i < nx - 1; i++) {
        dif(dx[i - 1], dx[i], cc[0][i], cc[1][i], cc[2][i]);
        dif2(dx[i - 1], dx[i], cc[3][i], cc[4][i], cc[5][i]);
    }
    dif(dx[nx - 2], dx[nx - 2] / 2, cc[0][nx - 1], cc[1][nx - 1], cc[2][nx - 1]);
    dif2(dx[nx - 2], dx[nx - 2] / 2, cc[3][nx - 1], cc[4][nx - 1], cc[5][nx - 1]);
    double dx0 = dx[0];
    double dx1 = dx0;
    double inv_dx1pdx0 = 1 / (dx1 + dx0);
    x2c_up[0] = dx0 * inv_dx1pdx0;
    x2c_d[0] = dx1 * inv_dx1pdx0;
    for (int i = 1; i < nx - 1; i++) {
        dx0 = dx1;
        dx1 = dx[i];
        inv_dx1pdx0 = 1 / (dx1 + dx0);
        x2c_up[i] = dx0 * inv_dx1pdx0;
        x2c_d[i] = dx1 * inv_dx1pdx0;
    }
    x2c_up[nx - 1] = 0.5;
    x2c_d[nx - 1] = 0.5;
}

void adj_release(m70h *h);

int copy_fields(m70h *h, double *const *dst, const double *const *src, const long *n, int count, hipMemcpyKind kind)
{
    for (int q = 0; q < count; ++q) {
        if (!src[q] || !dst[q]) continue;
        HIPCHK(h, hipMemcpyAsync(dst[q], src[q], (size_t)n[q] * sizeof(double), kind, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return M70H_OK;
}

}  // namespace

extern "C" {

const char *m70h_last_error(const m70h *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int m70h_create(int nx, int ny, const double *x, double dy, const m70h_params *p, m70h **out)
{
    if (!out) return fail(nullptr, M70H_EINVAL, "out is NULL");
    *out = nullptr;
    if (nx < 3 || ny < 3) return fail(nullptr, M70H_EINVAL, "nx and ny must be >= 3");
    if (!x || !p) return fail(nullptr, M70H_EINVAL, "x and p must be non-NULL");
    if (!(dy > 0.0)) return fail(nullptr, M70H_EINVAL, "dy must be > 0");
    if (p->drag_model < 0 || p->drag_model > 2 || p->visc_model < 0 || p->visc_model > 3 || p->offshore_bc < 0 ||
        p->offshore_bc > 1)
        return fail(nullptr, M70H_EINVAL, "bad model enum");
    for (int i = 0; i < nx - 1; ++i)
        if (!(x[i + 1] > x[i])) return fail(nullptr, M70H_EINVAL, "x must be strictly increasing");
    m70h *h = new m70h();
    h->nx = nx;
    h->ny = ny;
    h->dy = dy;
    h->F = (long)nx * ny;
    h->Fp = h->F + ny;
    h->p = *p;
    auto bail = [&](int rc) {
        g_create_error = h->err;
        m70h_destroy(h);
        return rc;
    };
    hipError_t e;
    if ((e = hipGetDevice(&h->device)) != hipSuccess) return bail(hip_fail(h, e, "hipGetDevice"));
    if ((e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(h, e, "hipStreamCreate"));
    h->stream = h->own_stream;

    // metrics
    std::vector<double> dx, dx_x, xx[6], cc[6], x2c_up, x2c_d;
    grid_metrics(nx, x, dx, dx_x, xx, cc, x2c_up, x2c_d);
    const double inv_rho = 1 / RHO;
    std::vector<double> host((size_t)22 * nx, 0.0);
    double *hp = host.data();
    auto put = [&](const std::vector<double> &v) { double *at = hp; std::copy(v.begin(), v.end(), hp); hp += nx; return at - host.data(); };
    std::vector<double> inv_dx(nx - 1), inv_dxx(nx), inv_rhodx(nx - 1);
    for (int i = 0; i < nx - 1; ++i) { inv_dx[i] = 1 / dx[i]; inv_rhodx[i] = inv_rho / dx[i]; }
    for (int i = 0; i < nx; ++i) inv_dxx[i] = 1 / dx_x[i];
    long o_inv_dx = put(inv_dx), o_inv_dxx = put(inv_dxx), o_inv_rhodx = put(inv_rhodx);
    long o_xx[6], o_cc[6];
    for (int q = 0; q < 6; ++q) o_xx[q] = put(xx[q]);
    for (int q = 0; q < 6; ++q) o_cc[q] = put(cc[q]);
    long o_x2c_up = put(x2c_up), o_x2c_d = put(x2c_d);
    if ((e = dalloc(h, &h->metric_buf, host.size())) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc"));
    if ((e = hipMemcpy(h->metric_buf, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
        return bail(hip_fail(h, e, "hipMemcpy"));
    const double *mb = h->metric_buf;
    h->m.inv_dx = mb + o_inv_dx;
    h->m.inv_dxx = mb + o_inv_dxx;
    h->m.inv_rhodx = mb + o_inv_rhodx;
    h->m.difx_x_up = mb + o_xx[0]; h->m.difx_x_d = mb + o_xx[1]; h->m.difx_x_low = mb + o_xx[2];
    h->m.dif2x_x_up = mb + o_xx[3]; h->m.dif2x_x_d = mb + o_xx[4]; h->m.dif2x_x_low = mb + o_xx[5];
    h->m.difx_c_up = mb + o_cc[0]; h->m.difx_c_d = mb + o_cc[1]; h->m.difx_c_low = mb + o_cc[2];
    h->m.dif2x_c_up = mb + o_cc[3]; h->m.dif2x_c_d = mb + o_cc[4]; h->m.dif2x_c_low = mb + o_cc[5];
    h->m.x2c_up = mb + o_x2c_up;
    h->m.x2c_d = mb + o_x2c_d;

    // constants (src/Hydro.cpp:44-60 for the boundary constants)
    set_consts(h, *p, dx_x[0], dy, true);
    // buffers
    const long F = h->F, Fp = h->Fp;
    double **stF[] = {&h->s.Zs, &h->s.D, &h->s.Dn, &h->s.Uc, &h->s.Vc, &h->s.dQx, &h->s.dQy};
    for (auto q : stF)
        if ((e = dalloc(h, q, F)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc state"));
    double **stFp[] = {&h->s.Qx, &h->s.Qy, &h->s.U, &h->s.V};
    for (auto q : stFp)
        if ((e = dalloc(h, q, Fp)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc state"));
    for (int q = 0; q < 7; ++q)
        if ((e = dalloc(h, &h->fw_buf[q], F)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc forcing"));
    h->fw = Forcing{h->fw_buf[0], h->fw_buf[1], h->fw_buf[2], h->fw_buf[3], h->fw_buf[4], h->fw_buf[5], h->fw_buf[6]};
    double **rfp[] = {&h->rf.cD, &h->rf.nuD, &h->rf.Fx, &h->rf.Urms_x, &h->rf.cD_x, &h->rf.nuD_x, &h->rf.dnuD_dx,
                      &h->rf.dnuDx_dy, &h->rf.Fy, &h->rf.Urms_y, &h->rf.cD_y, &h->rf.nuD_y, &h->rf.dnuD_dy,
                      &h->rf.dnuDy_dx};
    for (auto q : rfp)
        if ((e = dalloc(h, q, F)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc run fields"));
    double **wk[] = {&h->w.a, &h->w.b, &h->w.c, &h->w.f, &h->w.gD, &h->w.drag};
    for (auto q : wk)
        if ((e = dalloc(h, q, F)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc work"));
    if ((e = dalloc(h, &h->w.w, F)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc mask"));
    const long nLp = periodic_rows_padded_lines(nx - 1);
    const long scr = (long)ny * nLp > F ? (long)ny * nLp : F;
    if ((e = dalloc(h, &h->sr, scr)) != hipSuccess || (e = dalloc(h, &h->sv, scr)) != hipSuccess ||
        (e = dalloc(h, &h->sk, scr)) != hipSuccess)
        return bail(hip_fail(h, e, "hipMalloc scratch"));
    if ((e = dalloc(h, &h->flag, 1)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc flag"));
    if ((e = dalloc(h, &h->run_flag, 1)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc flag"));
    if ((e = hipHostMalloc((void **)&h->run_flag_host, sizeof(int), hipHostMallocDefault)) != hipSuccess)
        return bail(hip_fail(h, e, "hipHostMalloc flag"));
    *h->run_flag_host = 0;
    {   // three-pass column solver: E0..E2 reuse the coefficient arrays a, b, c
        ColsWork &cw = h->cw;
        cw.nseg = cols_segments(nx - 1);
        cw.E0 = h->w.a; cw.E1 = h->w.b; cw.E2 = h->w.c;
        const long nr = 2L * cw.nseg * ny;
        double **rp[] = {&cw.R0, &cw.R1, &cw.R2, &cw.R3, &cw.X};
        for (auto q : rp)
            if ((e = dalloc(h, q, nr)) != hipSuccess) return bail(hip_fail(h, e, "hipMalloc column work"));
    }
    *out = h;
    return M70H_OK;
}

void m70h_destroy(m70h *h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    adj_release(h);
    for (void *q : h->allocs) (void)hipFree(q);
    if (h->f0_dev) (void)hipFree(h->f0_dev);
    if (h->run_flag_host) (void)hipHostFree(h->run_flag_host);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
}

int m70h_set_params(m70h *h, const m70h_params *p)
{
    if (!h || !p) return M70H_EINVAL;
    if (p->drag_model < 0 || p->drag_model > 2 || p->visc_model < 0 || p->visc_model > 3)
        return fail(h, M70H_EINVAL, "bad model enum");
    const double f1 = h->k.f1, bcQy = h->k.bcQy;
    const int bc = h->p.offshore_bc;
    set_consts(h, *p, 0.0, h->dy, false);
    h->k.f1 = f1;
    h->k.bcQy = bcQy;
    h->p.offshore_bc = bc;
    return M70H_OK;
}

int m70h_set_stream(m70h *h, void *hip_stream)
{
    if (!h) return M70H_EINVAL;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->stream = hip_stream ? (hipStream_t)hip_stream : h->own_stream;
    return M70H_OK;
}

int m70h_init_from_bathymetry(m70h *h, const double *D, const double *Z, const double *hb)
{
    if (!h || !D || !Z || !hb) return h ? fail(h, M70H_EINVAL, "NULL array") : M70H_EINVAL;
    State &s = h->s;
    double *zero[] = {s.Qx, s.Qy, s.U, s.V};
    for (double *q : zero) HIPCHK(h, hipMemsetAsync(q, 0, (size_t)h->Fp * sizeof(double), h->stream));
    double *zeroF[] = {s.Uc, s.Vc, s.dQx, s.dQy};
    for (double *q : zeroF) HIPCHK(h, hipMemsetAsync(q, 0, (size_t)h->F * sizeof(double), h->stream));
    const size_t nb = (size_t)h->F * sizeof(double);
    // Z and h pass through the work arrays a, b
    HIPCHK(h, hipMemcpyAsync(s.D, D, nb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->w.a, Z, nb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->w.b, hb, nb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_init_zs(h->stream, s.Zs, s.D, h->w.a, h->w.b, h->F));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return M70H_OK;
}

static int state_xfer(m70h *h, double *const *hostside, bool to_device, hipMemcpyKind kind)
{
    State &s = h->s;
    double *dev[10] = {s.Zs, s.D, s.Qx, s.Qy, s.U, s.V, s.Uc, s.Vc, s.dQx, s.dQy};
    const long n[10] = {h->F, h->F, h->Fp, h->Fp, h->Fp, h->Fp, h->F, h->F, h->F, h->F};
    if (to_device) return copy_fields(h, dev, hostside, n, 10, kind);
    return copy_fields(h, hostside, dev, n, 10, kind);
}

int m70h_set_state(m70h *h, const double *Zs, const double *D, const double *Qx, const double *Qy, const double *U,
                   const double *V, const double *U_c, const double *V_c, const double *dQx_dx, const double *dQy_dy)
{
    if (!h) return M70H_EINVAL;
    double *src[10] = {(double *)Zs, (double *)D, (double *)Qx, (double *)Qy, (double *)U, (double *)V,
                       (double *)U_c, (double *)V_c, (double *)dQx_dx, (double *)dQy_dy};
    return state_xfer(h, src, true, hipMemcpyHostToDevice);
}

int m70h_get_state(m70h *h, double *Zs, double *D, double *Qx, double *Qy, double *U, double *V, double *U_c,
                   double *V_c, double *dQx_dx, double *dQy_dy)
{
    if (!h) return M70H_EINVAL;
    double *dst[10] = {Zs, D, Qx, Qy, U, V, U_c, V_c, dQx_dx, dQy_dy};
    return state_xfer(h, dst, false, hipMemcpyDeviceToHost);
}

int m70h_set_state_dev(m70h *h, const double *Zs, const double *D, const double *Qx, const double *Qy,
                       const double *U, const double *V, const double *U_c, const double *V_c, const double *dQx_dx,
                       const double *dQy_dy)
{
    if (!h) return M70H_EINVAL;
    double *src[10] = {(double *)Zs, (double *)D, (double *)Qx, (double *)Qy, (double *)U, (double *)V,
                       (double *)U_c, (double *)V_c, (double *)dQx_dx, (double *)dQy_dy};
    return state_xfer(h, src, true, hipMemcpyDeviceToDevice);
}

int m70h_get_state_dev(m70h *h, double *Zs, double *D, double *Qx, double *Qy, double *U, double *V, double *U_c,
                       double *V_c, double *dQx_dx, double *dQy_dy)
{
    if (!h) return M70H_EINVAL;
    double *dst[10] = {Zs, D, Qx, Qy, U, V, U_c, V_c, dQx_dx, dQy_dy};
    return state_xfer(h, dst, false, hipMemcpyDeviceToDevice);
}

static int forcing_xfer(m70h *h, const double *const *src, hipMemcpyKind kind)
{
    for (int q = 0; q < 7; ++q)
        if (!src[q]) return fail(h, M70H_EINVAL, "all seven forcing arrays are required");
    long n[7];
    for (auto &v : n) v = h->F;
    int rc = copy_fields(h, h->fw_buf, src, n, 7, kind);
    if (rc == M70H_OK) h->forcing_set = true;
    h->run_gen++;
    return rc;
}

int m70h_set_forcing(m70h *h, const double *Sxx, const double *Sxy, const double *Syy, const double *Urms,
                     const double *Dw, const double *D_r, const double *H)
{
    if (!h) return M70H_EINVAL;
    const double *src[7] = {Sxx, Sxy, Syy, Urms, Dw, D_r, H};
    return forcing_xfer(h, src, hipMemcpyHostToDevice);
}

int m70h_set_forcing_dev(m70h *h, const double *Sxx, const double *Sxy, const double *Syy, const double *Urms,
                         const double *Dw, const double *D_r, const double *H)
{
    if (!h) return M70H_EINVAL;
    const double *src[7] = {Sxx, Sxy, Syy, Urms, Dw, D_r, H};
    return forcing_xfer(h, src, hipMemcpyDeviceToDevice);
}

int m70h_set_offshore_f0(m70h *h, double f0)
{
    if (!h) return M70H_EINVAL;
    h->f0 = f0;
    h->f0_series.clear();
    h->f0_pos = 0;
    return M70H_OK;
}

int m70h_set_offshore_f0_series(m70h *h, const double *f0, long n)
{
    if (!h) return M70H_EINVAL;
    if (n < 0 || (n > 0 && !f0)) return fail(h, M70H_EINVAL, "bad f0 series");
    h->f0_series.assign(f0, f0 + n);
    h->f0_pos = 0;
    // device copy for the batched steps, which read the step's value on the device
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->f0_dev && h->f0_dev_n < n) {
        HIPCHK(h, hipFree(h->f0_dev));
        h->f0_dev = nullptr;
        h->f0_dev_n = 0;
    }
    if (n > 0) {
        if (!h->f0_dev) {
            HIPCHK(h, hipMalloc(&h->f0_dev, n * sizeof(double)));
            h->f0_dev_n = n;
        }
        HIPCHK(h, hipMemcpy(h->f0_dev, f0, n * sizeof(double), hipMemcpyHostToDevice));
    }
    return M70H_OK;
}

int m70h_bc_constants(const m70h *h, double *f1, double *f0, double *bcQy)
{
    if (!h) return M70H_EINVAL;
    if (f1) *f1 = h->k.f1;
    if (f0) *f0 = h->f0;
    if (bcQy) *bcQy = h->k.bcQy;
    return M70H_OK;
}

}  // extern "C"

// Run(nIter) enqueued on the handle's stream. With `check`, a scan of the final state for
// NaN/Inf follows; its flag lands in pinned host memory and m70h_synchronize reports it.
// The device outer loop enqueues runs without the scan.
static int run_enqueue(m70h *h, int nIter, bool check)
{
    if (!h) return M70H_EINVAL;
    if (nIter < 0) return fail(h, M70H_EINVAL, "nIter < 0");
    if (!h->forcing_set) return fail(h, M70H_ESTATE, "forcing not set");
    h->run_gen++;
    int rc = run_preloop(h);
    if (rc) return rc;
    for (int it = 0; it < nIter; ++it)
        if ((rc = run_step(h, it == nIter - 1, it == 0))) return rc;
    if (check) {
        // a fill kernel rather than a memset, so that a caller may capture the Run in a graph
        HIPCHK(h, launch_flag_clear(h->stream, h->run_flag));
        State &s = h->s;
        const double *f[] = {s.Zs, s.D, s.Qx, s.Qy, s.U, s.V, s.Uc, s.Vc, s.dQx, s.dQy};
        const long n[] = {h->F, h->F, h->Fp, h->Fp, h->Fp, h->Fp, h->F, h->F, h->F, h->F};
        for (int q = 0; q < 10; ++q) HIPCHK(h, launch_nonfinite(h->stream, f[q], n[q], h->run_flag));
        HIPCHK(h, hipMemcpyAsync(h->run_flag_host, h->run_flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        h->run_check_pending = true;
    }
    return M70H_OK;
}

extern "C" {

int m70h_run_async(m70h *h, int nIter) { return run_enqueue(h, nIter, true); }

int m70h_synchronize(m70h *h)
{
    if (!h) return M70H_EINVAL;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->prof) {
        int rc = collect_profile(h);
        if (rc) return rc;
    }
    if (h->run_check_pending) {
        h->run_check_pending = false;
        if (*h->run_flag_host) return fail(h, M70H_ENONFINITE, "non-finite value in the state after Run");
    }
    return M70H_OK;
}

int m70h_run(m70h *h, int nIter)
{
    int rc = m70h_run_async(h, nIter);
    if (rc) return rc;
    return m70h_synchronize(h);
}

int m70h_set_solver(m70h *h, int solver)
{
    if (!h) return M70H_EINVAL;
    if (solver != M70H_SOLVER_EXACT && solver != M70H_SOLVER_PARTITION && solver != M70H_SOLVER_FUSED)
        return fail(h, M70H_EINVAL, "unknown solver");
    if (solver != M70H_SOLVER_EXACT && (!partition_supported(h->nx - 1) || !partition_supported(h->ny)))
        return fail(h, M70H_EINVAL, "partition solver supports line lengths up to 8192");
    if (solver == M70H_SOLVER_FUSED && !rows_lds_supported(h->ny))
        return fail(h, M70H_EINVAL, "fused solver supports ny up to 8192");
    h->solver = solver;
    return M70H_OK;
}

int m70h_set_solver_tiles(m70h *h, int tile_cols, int tile_rows)
{
    if (!h) return M70H_EINVAL;
    auto ok = [](int t) { return t == 0 || t == 8 || t == 16; };
    if (!ok(tile_cols) || !ok(tile_rows)) return fail(h, M70H_EINVAL, "tiles must be 0, 8 or 16");
    h->tile_cols = tile_cols;
    h->tile_rows = tile_rows;
    return M70H_OK;
}

int m70h_set_rows_block(m70h *h, int threads)
{
    if (!h) return M70H_EINVAL;
    if (threads != 0 && threads != 256 && threads != 512) return fail(h, M70H_EINVAL, "rows block must be 0, 256 or 512");
    if (threads == 256 && rows_block_for(h->ny, 256) != 256) return fail(h, M70H_EINVAL, "ny too large for 256");
    h->rows_block = threads;
    return M70H_OK;
}

int m70h_tridiag_solve(void *hip_stream, int solver, int lines_are_rows, int periodic, int n, int nlines, long ld,
                       const double *a_dev, const double *b_dev, const double *c_dev, const double *f_dev,
                       double *y_dev, long ldy, double *scratch_dev, long scratch_len, int tile)
{
    hipStream_t st = (hipStream_t)hip_stream;
    if (n < 3 || nlines < 1 || !a_dev || !b_dev || !c_dev || !f_dev || !y_dev) return M70H_EINVAL;
    hipError_t e;
    if (solver == M70H_SOLVER_PARTITION) {
        if (!partition_supported(n)) return M70H_EINVAL;
        e = lines_are_rows ? launch_partition_rows(st, a_dev, b_dev, c_dev, f_dev, ld, y_dev, ldy, nlines, n,
                                                   periodic != 0, tile)
                           : (periodic ? hipErrorInvalidValue
                                       : launch_partition_cols(st, a_dev, b_dev, c_dev, f_dev, y_dev, n, nlines, ld,
                                                               tile));
    } else if (solver == M70H_SOLVER_EXACT) {
        if (lines_are_rows && periodic) {
            const long nLp = periodic_rows_padded_lines(nlines);
            if (!scratch_dev || scratch_len < 3 * (long)n * nLp) return M70H_EINVAL;
            e = launch_periodic_rows(st, a_dev, b_dev, c_dev, f_dev, ld, y_dev, ldy, scratch_dev,
                                     scratch_dev + (long)n * nLp, scratch_dev + 2L * n * nLp, nlines, n);
        } else if (!lines_are_rows && !periodic) {
            if (ldy != ld || !scratch_dev || scratch_len < (long)n * ld) return M70H_EINVAL;
            e = launch_thomas_cols(st, a_dev, b_dev, c_dev, f_dev, y_dev, scratch_dev, n, nlines, ld);
        } else {
            return M70H_EINVAL;
        }
    } else {
        return M70H_EINVAL;
    }
    if (e != hipSuccess) {
        g_create_error = std::string("m70h_tridiag_solve: ") + hipGetErrorString(e);
        return e == hipErrorInvalidValue ? M70H_EINVAL : M70H_EHIP;
    }
    return M70H_OK;
}

}  // extern "C"

// Static scratch of m70h_tridiag_solve_noalloc: one copy per GPU, in the code object, so
// the entry point never allocates. Host code uses only the per-device address that
// hipGetSymbolAddress returns (the symbol's name is a host-side stand-in).
static __device__ __align__(256) double scratch[M70H_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on device `dev`, looked up once with that device current.
int noalloc_scratch(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDevices) {
        g_create_error = "m70h_tridiag_solve_noalloc: device ordinal out of range";
        return M70H_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            g_create_error = std::string("m70h_tridiag_solve_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return M70H_EHIP;
        }
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return M70H_OK;
}
}  // namespace

extern "C" {

int m70h_tridiag_solve_noalloc(void *hip_stream, int solver, int lines_are_rows, int periodic, int n, int nlines,
                               long ld, const double *a_dev, const double *b_dev, const double *c_dev,
                               const double *f_dev, double *y_dev, long ldy, int tile)
{
    if (n < 3 || nlines < 1) return M70H_EINVAL;
    long need = 0;
    if (solver == M70H_SOLVER_EXACT)
        need = lines_are_rows ? 3L * n * periodic_rows_padded_lines(nlines) : (long)n * ld;
    if (need > M70H_NOALLOC_SCRATCH_DOUBLES) {
        g_create_error = "m70h_tridiag_solve_noalloc: the solve needs " + std::to_string(need) +
                         " doubles of scratch, above the static maximum of " +
                         std::to_string(M70H_NOALLOC_SCRATCH_DOUBLES);
        return M70H_EINVAL;
    }
    int dev = 0;
    hipError_t e = hip_stream ? hipStreamGetDevice((hipStream_t)hip_stream, &dev) : hipGetDevice(&dev);
    if (e != hipSuccess) {
        g_create_error = std::string("m70h_tridiag_solve_noalloc: device of the stream: ") + hipGetErrorString(e);
        return M70H_EHIP;
    }
    double *sp = nullptr;
    int rc = noalloc_scratch(dev, &sp);
    if (rc) return rc;
    return m70h_tridiag_solve(hip_stream, solver, lines_are_rows, periodic, n, nlines, ld, a_dev, b_dev, c_dev, f_dev,
                              y_dev, ldy, sp, M70H_NOALLOC_SCRATCH_DOUBLES, tile);
}

int m70h_check_finite(m70h *h, int *finite)
{
    if (!h) return M70H_EINVAL;
    HIPCHK(h, hipMemsetAsync(h->flag, 0, sizeof(int), h->stream));
    State &s = h->s;
    const double *f[] = {s.Zs, s.D, s.Qx, s.Qy, s.U, s.V, s.Uc, s.Vc, s.dQx, s.dQy};
    const long n[] = {h->F, h->F, h->Fp, h->Fp, h->Fp, h->Fp, h->F, h->F, h->F, h->F};
    for (int q = 0; q < 10; ++q) HIPCHK(h, launch_nonfinite(h->stream, f[q], n[q], h->flag));
    int v = 0;
    HIPCHK(h, hipMemcpyAsync(&v, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (finite) *finite = v ? 0 : 1;
    return v ? fail(h, M70H_ENONFINITE, "non-finite value in the state") : M70H_OK;
}

int m70h_set_profiling(m70h *h, int enable)
{
    if (!h) return M70H_EINVAL;
    int rc = collect_profile(h);
    h->prof = enable != 0;
    h->prof_only = -1;
    for (int q = 0; q < K_COUNT; ++q) { h->prof_ms[q] = 0; h->prof_n[q] = 0; }
    return rc;
}

int m70h_set_profiling_only(m70h *h, const char *kernel)
{
    if (!h) return M70H_EINVAL;
    int kid = -1;
    if (kernel) {
        for (int q = 0; q < K_COUNT; ++q)
            if (std::strcmp(kKernelNames[q], kernel) == 0) kid = q;
        if (kid < 0) return fail(h, M70H_EINVAL, std::string("unknown kernel ") + kernel);
    }
    int rc = collect_profile(h);
    h->prof = true;
    h->prof_only = kid;
    for (int q = 0; q < K_COUNT; ++q) { h->prof_ms[q] = 0; h->prof_n[q] = 0; }
    return rc;
}

int m70h_profile(m70h *h, int kk, const char **name, double *ms, long *launches)
{
    if (!h) return M70H_EINVAL;
    if (kk < 0 || kk >= K_COUNT) return M70H_EINVAL;
    int rc = collect_profile(h);
    if (rc) return rc;
    if (name) *name = kKernelNames[kk];
    if (ms) *ms = h->prof_ms[kk];
    if (launches) *launches = h->prof_n[kk];
    return M70H_OK;
}

int m70h_state_ptr(m70h *h, const char *name, double **dev)
{
    if (!h || !name || !dev) return M70H_EINVAL;
    State &s = h->s;
    struct { const char *n; double *p; } tab[] = {{"Zs", s.Zs}, {"D", s.D}, {"Qx", s.Qx}, {"Qy", s.Qy}, {"U", s.U},
                                                  {"V", s.V}, {"U_c", s.Uc}, {"V_c", s.Vc}, {"dQx_dx", s.dQx},
                                                  {"dQy_dy", s.dQy}};
    for (auto &t : tab)
        if (std::strcmp(t.n, name) == 0) { *dev = t.p; return M70H_OK; }
    return fail(h, M70H_EINVAL, std::string("unknown field ") + name);
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Batches: many handles advanced by grouped launches (see include/m70h.h).
// ------------------------------------------------------------------------------------
struct m70h_batch {
    std::vector<m70h *> hs;
    hipStream_t stream = nullptr;
    ScenDesc *d_desc = nullptr;   // [2][nscen]: even steps (D, Dn) and odd steps (swapped)
    int *d_start = nullptr;       // [GS_COUNT][nscen+1]
    int *d_step = nullptr;        // steps done in the current run (offshore offset series)
    int blocks[GS_COUNT] = {};
    int cell_blocks = 0;
    bool use_graph = true;
    int rows_block = 0;   // row-solver block size for every scenario of the batch
    bool const_drag = true;   // every scenario has the constant drag model
    hipGraphExec_t gexec = nullptr;
    std::string err;
};

static int bfail(m70h_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg; else g_create_error = msg;
    return code;
}

#define BHIPCHK(b, expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return bfail(b, M70H_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static void fill_desc(const m70h *h, ScenDesc &d, bool odd, int rb, const int *step)
{
    int rows, cols, red, redp, S_rows, S_red, L_red;
    grouped_stage_blocks(h->nx, h->ny, rb, &rows, &cols, &red, &redp, &S_rows, &S_red, &L_red);
    d.k = h->k; d.m = h->m; d.s = h->s; d.rf = h->rf; d.fw = h->fw; d.w = h->w; d.cw = h->cw;
    if (odd) std::swap(d.s.D, d.s.Dn);
    d.f0 = h->f0;
    d.f0s = h->f0_series.empty() ? nullptr : h->f0_dev;
    d.f0_n = (long)h->f0_series.size();
    d.f0_base = h->f0_pos;
    d.step = step;
    d.S_rows = S_rows; d.S_red = S_red; d.L_red = L_red; d.pad = 0;
}

static int batch_upload(m70h_batch *b)
{
    const int n = (int)b->hs.size();
    std::vector<ScenDesc> host(2 * (size_t)n);
    for (int q = 0; q < n; ++q) {
        fill_desc(b->hs[q], host[q], false, b->rows_block, b->d_step);
        fill_desc(b->hs[q], host[n + q], true, b->rows_block, b->d_step);
    }
    BHIPCHK(b, hipMemcpyAsync(b->d_desc, host.data(), host.size() * sizeof(ScenDesc), hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, launch_flag_clear(b->stream, b->d_step));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return M70H_OK;
}

// One grouped step. first: the first step of a batched Run, which reads V_c as the handles
// hold it; later steps form it from V. last: the last step, which stores U_c and V_c (the
// steps before it leave them to be formed from U and V, as run_step does). The defaults
// give a step that reads and leaves every field in memory.
static hipError_t batch_step(m70h_batch *b, int odd, bool first = true, bool last = true)
{
    const int n = (int)b->hs.size();
    const ScenDesc *D = b->d_desc + (odd ? n : 0);
    const int *st_rows = b->d_start + GS_ROWS * (n + 1);
    const int *st_cols = b->d_start + GS_COLS * (n + 1), *st_red = b->d_start + GS_RED * (n + 1);
    const int *st_redp = b->d_start + GS_REDP * (n + 1);
    const int *st_tx = b->d_start + GS_TX * (n + 1), *st_ty = b->d_start + GS_TY * (n + 1);
    hipStream_t st = b->stream;
    hipError_t e;
    // sweep order as run_step's: each stage starts where the one before it ended (those
    // rows are still in the MALL), flipping from step to step with the D/Dn parity
    const int ra = odd, rb = 1 - odd;
    // the Qx projector rows form the Qx predictor's solution from its first two passes
    if ((e = launch_cols3_grouped(st, 0, D, st_cols, st_red, st_redp, n, b->blocks[GS_COLS], b->blocks[GS_RED],
                                  b->blocks[GS_REDP], false, ra, rb, !first, b->const_drag)) != hipSuccess)
        return e;
    if ((e = launch_rows_grouped(st, 0, b->rows_block, D, st_rows, n, b->blocks[GS_ROWS], rb, !first)) != hipSuccess)
        return e;
    if ((e = launch_tails_grouped(st, 2, D, st_tx, n, b->blocks[GS_TX], ra, last)) != hipSuccess) return e;
    if ((e = launch_rows_grouped(st, 1, b->rows_block, D, st_rows, n, b->blocks[GS_ROWS], rb)) != hipSuccess) return e;
    if ((e = launch_cols3_grouped(st, 1, D, st_cols, st_red, st_redp, n, b->blocks[GS_COLS], b->blocks[GS_RED],
                                  b->blocks[GS_REDP], true, ra, rb)) != hipSuccess)
        return e;
    return launch_tails_grouped(st, 3, D, st_ty, n, b->blocks[GS_TY], ra, last);
}

extern "C" {

const char *m70h_batch_last_error(const m70h_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int m70h_batch_create(int nscen, m70h *const *handles, m70h_batch **out)
{
    if (!out) return bfail(nullptr, M70H_EINVAL, "out is NULL");
    *out = nullptr;
    if (nscen < 1 || !handles) return bfail(nullptr, M70H_EINVAL, "need at least one handle");
    m70h_batch *b = new m70h_batch();
    auto bail = [&](int rc) { g_create_error = b->err; m70h_batch_destroy(b); return rc; };
    int dev0 = -1;
    int maxny = 0;
    for (int q = 0; q < nscen; ++q)
        if (handles[q] && handles[q]->ny > maxny) maxny = handles[q]->ny;
    b->rows_block = rows_block_for(maxny, 0);
    std::vector<int> starts((size_t)GS_COUNT * (nscen + 1), 0);
    for (int q = 0; q < nscen; ++q) {
        m70h *h = handles[q];
        if (!h) return bail(bfail(b, M70H_EINVAL, "NULL handle"));
        if (dev0 < 0) dev0 = h->device;
        if (h->device != dev0) return bail(bfail(b, M70H_EINVAL, "handles on different devices"));
        if (!rows_lds_supported(h->ny) || !partition_supported(h->nx - 1))
            return bail(bfail(b, M70H_EINVAL, "batched step supports ny and nx-1 up to 4096"));
        int rows, cols, red, redp, S_rows, S_red, L_red;
        grouped_stage_blocks(h->nx, h->ny, b->rows_block, &rows, &cols, &red, &redp, &S_rows, &S_red, &L_red);
        if (S_rows <= 0 || S_red <= 0) return bail(bfail(b, M70H_EINVAL, "grid too small for the batched step"));
        const int per[GS_COUNT] = {grouped_cell_blocks(h->nx, h->ny), rows, cols, red, redp,
                                   grouped_tail_blocks(h->nx, h->ny, 2), grouped_tail_blocks(h->nx, h->ny, 3)};
        for (int gs = 0; gs < GS_COUNT; ++gs)
            starts[(size_t)gs * (nscen + 1) + q + 1] = starts[(size_t)gs * (nscen + 1) + q] + per[gs];
        b->hs.push_back(h);
        if (h->p.drag_model != 0) b->const_drag = false;
    }
    for (int gs = 0; gs < GS_COUNT; ++gs) b->blocks[gs] = starts[(size_t)gs * (nscen + 1) + nscen];
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&b->d_desc, 2 * (size_t)nscen * sizeof(ScenDesc))) != hipSuccess ||
        (e = hipMalloc(&b->d_start, starts.size() * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&b->d_step, sizeof(int))) != hipSuccess ||
        (e = hipMemcpy(b->d_start, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
        return bail(bfail(b, M70H_EHIP, std::string("batch allocation: ") + hipGetErrorString(e)));
    *out = b;
    return M70H_OK;
}

void m70h_batch_destroy(m70h_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_start) (void)hipFree(b->d_start);
    if (b->d_step) (void)hipFree(b->d_step);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int m70h_batch_set_rows_block(m70h_batch *b, int threads)
{
    if (!b) return M70H_EINVAL;
    int maxny = 0;
    for (m70h *h : b->hs) maxny = h->ny > maxny ? h->ny : maxny;
    if (threads != 256 && threads != 512) return bfail(b, M70H_EINVAL, "rows block must be 256 or 512");
    if (threads == 256 && rows_block_for(maxny, 256) != 256) return bfail(b, M70H_EINVAL, "ny too large for 256");
    for (m70h *h : b->hs) {   // segment counts depend on the block size
        int rows, cols, red, redp, S_rows, S_red, L_red;
        grouped_stage_blocks(h->nx, h->ny, threads, &rows, &cols, &red, &redp, &S_rows, &S_red, &L_red);
        if (S_rows <= 0) return bfail(b, M70H_EINVAL, "grid too small for the row block");
    }
    b->rows_block = threads;
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    return M70H_OK;
}

int m70h_batch_set_graph(m70h_batch *b, int use_graph)
{
    if (!b) return M70H_EINVAL;
    b->use_graph = use_graph != 0;
    return M70H_OK;
}

int m70h_batch_stream(m70h_batch *b, void **hip_stream)
{
    if (!b || !hip_stream) return M70H_EINVAL;
    *hip_stream = (void *)b->stream;
    return M70H_OK;
}

int m70h_batch_run_async(m70h_batch *b, int nIter)
{
    if (!b) return M70H_EINVAL;
    if (nIter < 0) return bfail(b, M70H_EINVAL, "nIter < 0");
    for (m70h *h : b->hs) {
        if (!h->forcing_set) return bfail(b, M70H_ESTATE, "forcing not set on a handle");
        BHIPCHK(b, hipStreamSynchronize(h->stream));   // the handles' own queued work first
    }
    int rc = batch_upload(b);
    if (rc) return rc;
    const int n = (int)b->hs.size();
    hipStream_t st = b->stream;
    // per-Run pre-loop (src/Hydro.cpp:175-192) of every scenario
    BHIPCHK(b, launch_cells_grouped(st, 0, b->d_desc, b->d_start + GS_CELLS * (n + 1), n, b->blocks[GS_CELLS]));
    BHIPCHK(b, launch_cells_grouped(st, 1, b->d_desc, b->d_start + GS_CELLS * (n + 1), n, b->blocks[GS_CELLS]));
    // the first step reads V_c as the handles hold it and the last one stores U_c and V_c;
    // the steps between form them from U and V, two at a time from a graph: odd (Dn -> D)
    // then even (D -> Dn)
    int it = 0;
    if (nIter >= 1) {
        BHIPCHK(b, batch_step(b, 0, true, nIter == 1));
        it = 1;
    }
    if (b->use_graph && it + 2 <= nIter - 1) {
        if (!b->gexec) {
            hipGraph_t g = nullptr;
            BHIPCHK(b, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipError_t e1 = batch_step(b, 1, false, false);
            hipError_t e2 = e1 == hipSuccess ? batch_step(b, 0, false, false) : e1;
            hipError_t e3 = hipStreamEndCapture(st, &g);
            if (e2 != hipSuccess || e3 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(b, M70H_EHIP, std::string("graph capture: ") + hipGetErrorString(e2 != hipSuccess ? e2 : e3));
            }
            hipError_t e4 = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e4 != hipSuccess) {
                b->gexec = nullptr;
                return bfail(b, M70H_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e4));
            }
        }
        for (; it + 2 <= nIter - 1; it += 2) BHIPCHK(b, hipGraphLaunch(b->gexec, st));
    }
    for (; it < nIter; ++it) BHIPCHK(b, batch_step(b, it & 1, false, it == nIter - 1));
    if (nIter & 1)
        for (m70h *h : b->hs) std::swap(h->s.D, h->s.Dn);
    for (m70h *h : b->hs) h->f0_pos += nIter;
    return M70H_OK;
}

int m70h_batch_synchronize(m70h_batch *b)
{
    if (!b) return M70H_EINVAL;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return M70H_OK;
}

int m70h_batch_run(m70h_batch *b, int nIter)
{
    int rc = m70h_batch_run_async(b, nIter);
    if (rc) return rc;
    return m70h_batch_synchronize(b);
}

}  // extern "C"

// =====================================================================================
// Adjoint of Run (no counterpart in the reference): checkpointed forward, taped step,
// reverse sweep. The checkpoint schedule is the optimal binomial one (Griewank's
// Revolve), found by dynamic programming over (steps, free slots).
// =====================================================================================
struct AdjCtx {
    int N = 0;                   // steps of the recorded Run
    int S = 0;                   // free slots of the schedule (state buffers minus the start)
    long f0_pos0 = 0;            // offshore-hook position at the start of the Run
    long gen = 0;                // run generation the recording belongs to
    std::vector<void *> allocs;
    std::vector<double *> slot;  // state buffers, 6F + 4F+ doubles each
    std::vector<int> slot_step;  // step held by each buffer, -1 when free
    double *s0buf = nullptr;     // state before the step being reversed, when no slot holds it
    double *final_buf = nullptr; // state after the Run, restored at the end of the reverse sweep
    std::vector<int> split;      // first advance of the optimal schedule, (N+1) x (S+1)
    StepTape T{};
    AdjFields A[2]{};
    AdjWork W{};
    AdjRun R{};
    double *params = nullptr;    // device: adjoints of cD0, nu0, M
    long fwd_steps = 0, taped_steps = 0;
    int cur = 0;                 // A[cur] holds the adjoint of the current state
    bool tape_all = false;       // every step's state and intermediates kept by the forward sweep
    int slots_arg = -1;          // the slots argument the buffers were sized for
    std::vector<StepTape> tapes; // tape_all: one per step (slot[t] holds the state before step t)
};

namespace {

void adj_release(m70h *h)
{
    if (!h || !h->adj) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *q : h->adj->allocs) (void)hipFree(q);
    delete h->adj;
    h->adj = nullptr;
}

hipError_t adj_alloc(AdjCtx *a, double **p, long n)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)n * sizeof(double) + 64);
    if (e != hipSuccess) return e;
    a->allocs.push_back(q);
    *p = (double *)q;
    return hipMemset(q, 0, (size_t)n * sizeof(double) + 64);
}

State buf_state(const m70h *h, double *b)
{
    const long F = h->F, Fp = h->Fp;
    State s{};
    s.Zs = b;
    s.D = b + F;
    s.Qx = b + 2 * F;
    s.Qy = s.Qx + Fp;
    s.U = s.Qy + Fp;
    s.V = s.U + Fp;
    s.Uc = s.V + Fp;
    s.Vc = s.Uc + F;
    s.dQx = s.Vc + F;
    s.dQy = s.dQx + F;
    return s;
}

hipError_t copy_state(m70h *h, const State &dst, const State &src)
{
    return launch_copy_state(h->stream, dst, src, h->F, h->Fp);
}

hipError_t zero_fields(m70h *h, const AdjFields &A)
{
    const long F = h->F, Fp = h->Fp;
    double *d[10] = {A.Zs, A.D, A.Qx, A.Qy, A.U, A.V, A.Uc, A.Vc, A.dQx, A.dQy};
    const long n[10] = {F, F, Fp, Fp, Fp, Fp, F, F, F, F};
    for (int q = 0; q < 10; ++q) {
        hipError_t e = hipMemsetAsync(d[q], 0, (size_t)n[q] * sizeof(double), h->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

double f0_at(m70h *h, long pos)
{
    if (!h->f0_series.empty()) {
        const long q = pos < (long)h->f0_series.size() ? pos : (long)h->f0_series.size() - 1;
        return h->f0_series[q];
    }
    return h->f0;
}

// Tape buffers of one step (the right-hand side scratch f is shared).
hipError_t adj_alloc_tape(AdjCtx *a, StepTape &T, long F, double *shared_f)
{
    hipError_t e;
    double **tp[] = {&T.a1, &T.b1, &T.c1, &T.qxs, &T.a2, &T.b2, &T.c2, &T.a3, &T.b3, &T.c3, &T.qys,
                     &T.a4, &T.b4, &T.c4, &T.dragx, &T.dragy, &T.gD};
    for (auto q : tp)
        if ((e = adj_alloc(a, q, F)) != hipSuccess) return e;
    double *mk = nullptr;
    if ((e = adj_alloc(a, &mk, F / 4 + 1)) != hipSuccess) return e;
    T.wx = (uint8_t *)mk;
    if ((e = adj_alloc(a, &mk, F / 4 + 1)) != hipSuccess) return e;
    T.wy = (uint8_t *)mk;
    T.f = shared_f;
    return hipSuccess;
}

// Allocates the context for a Run of N steps with `slots` state buffers.
int adj_setup(m70h *h, int N, int slots, bool tape_all)
{
    if (h->adj && h->adj->N == N && h->adj->tape_all == tape_all && h->adj->slots_arg == slots) {
        // same Run length and schedule: keep the buffers (every one is rewritten before use)
        AdjCtx *a = h->adj;
        for (auto &q : a->slot_step) q = -1;
        a->fwd_steps = a->taped_steps = 0;
        return M70H_OK;
    }
    adj_release(h);
    AdjCtx *a = new AdjCtx();
    a->slots_arg = slots;
    h->adj = a;
    a->N = N;
    a->tape_all = tape_all;
    a->S = slots - 1 < N - 1 ? slots - 1 : N - 1;
    if (a->S < 0) a->S = 0;
    if (tape_all) a->S = N;   // states 0..N
    const long F = h->F, Fp = h->Fp, SB = 6 * F + 4 * Fp;
    hipError_t e;
    a->slot.assign(a->S + 1, nullptr);
    a->slot_step.assign(a->S + 1, -1);
    for (auto &b : a->slot)
        if ((e = adj_alloc(a, &b, SB)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc checkpoints");
    if ((e = adj_alloc(a, &a->s0buf, SB)) != hipSuccess || (e = adj_alloc(a, &a->final_buf, SB)) != hipSuccess)
        return hip_fail(h, e, "adjoint: hipMalloc state");
    double *fbuf = nullptr;
    if ((e = adj_alloc(a, &fbuf, F)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc tape");
    if ((e = adj_alloc_tape(a, a->T, F, fbuf)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc tape");
    if (tape_all) {
        a->tapes.resize(N);
        for (int t = 0; t < N; ++t)
            if ((e = adj_alloc_tape(a, a->tapes[t], F, fbuf)) != hipSuccess)
                return hip_fail(h, e, "adjoint: hipMalloc per-step tapes (use checkpoints instead)");
    }
    for (int q = 0; q < 2; ++q) {
        AdjFields &A = a->A[q];
        double **fp[] = {&A.Zs, &A.D, &A.Uc, &A.Vc, &A.dQx, &A.dQy};
        for (auto p : fp)
            if ((e = adj_alloc(a, p, F)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc adjoint");
        double **pp[] = {&A.Qx, &A.Qy, &A.U, &A.V};
        for (auto p : pp)
            if ((e = adj_alloc(a, p, Fp)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc adjoint");
    }
    AdjWork &W = a->W;
    double **wp[] = {&W.rhs, &W.lam, &W.dragx, &W.dragy, &W.UcN, &W.dQxN, &W.gD, &W.dQyN, &W.qyv, &W.qU};
    for (auto p : wp)
        if ((e = adj_alloc(a, p, F)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc work");
    if ((e = adj_alloc(a, &W.UN, Fp)) != hipSuccess || (e = adj_alloc(a, &W.qy0v, h->ny)) != hipSuccess)
        return hip_fail(h, e, "adjoint: hipMalloc work");
    AdjRun &R = a->R;
    double **rp[] = {&R.cD_x, &R.nuD_x, &R.dnuD_dx, &R.dnuDx_dy, &R.cD_y, &R.nuD_y, &R.dnuD_dy, &R.dnuDy_dx,
                     &R.cD, &R.nuD};
    for (auto p : rp)
        if ((e = adj_alloc(a, p, F)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc run adjoints");
    if ((e = adj_alloc(a, &a->params, 4)) != hipSuccess) return hip_fail(h, e, "adjoint: hipMalloc params");
    if (tape_all) return M70H_OK;
    // schedule: cost(l, s) = fewest untaped steps to reverse l steps from a stored start
    // with s free buffers; split(l, s) = the first advance that achieves it
    const int S = a->S;
    std::vector<long> cost((size_t)(N + 1) * (S + 1), 0);
    a->split.assign((size_t)(N + 1) * (S + 1), 1);
    auto C = [&](int l, int s) -> long & { return cost[(size_t)l * (S + 1) + s]; };
    for (int l = 2; l <= N; ++l) {
        C(l, 0) = (long)l * (l - 1) / 2;
        a->split[(size_t)l * (S + 1)] = l - 1;
    }
    for (int s = 1; s <= S; ++s)
        for (int l = 2; l <= N; ++l) {
            long best = -1;
            int bj = 1;
            for (int j = 1; j < l; ++j) {
                const long c = j + C(l - j, s - 1) + C(j, s);
                if (best < 0 || c < best) { best = c; bj = j; }
            }
            C(l, s) = best;
            a->split[(size_t)l * (S + 1) + s] = bj;
        }
    return M70H_OK;
}

int adj_store(m70h *h, int step)
{
    AdjCtx *a = h->adj;
    for (size_t q = 0; q < a->slot.size(); ++q)
        if (a->slot_step[q] < 0) {
            HIPCHK(h, copy_state(h, buf_state(h, a->slot[q]), h->s));
            a->slot_step[q] = step;
            return M70H_OK;
        }
    return fail(h, M70H_ESTATE, "adjoint: no free checkpoint buffer (schedule error)");
}

int adj_find(m70h *h, int step)
{
    AdjCtx *a = h->adj;
    for (size_t q = 0; q < a->slot.size(); ++q)
        if (a->slot_step[q] == step) return (int)q;
    return -1;
}

// Advances h->s from step `from` by n untaped steps.
int adj_advance(m70h *h, int from, int n)
{
    for (int t = 0; t < n; ++t) {
        h->f0_pos = h->adj->f0_pos0 + from + t;
        int rc = run_step(h);
        if (rc) return rc;
        h->adj->fwd_steps++;
    }
    return M70H_OK;
}

// One step from h->s with its intermediates taped (assembly kernels + partition solves).
int adj_taped_step(m70h *h, int t, const StepTape &T)
{
    AdjCtx *a = h->adj;
    const double f0 = f0_at(h, a->f0_pos0 + t);
    const int nx = h->nx, ny = h->ny;
    const Consts &k = h->k;
    State &s = h->s;
    hipStream_t st = h->stream;
    const long nF = (long)(nx - 1) * ny;
    Work w = h->w;
    w.gD = T.gD;
    Work wx = w, wy = w;
    wx.w = T.wx; wx.drag = T.dragx;
    wy.w = T.wy; wy.drag = T.dragy;
    hipError_t e;
    wx.a = T.a1; wx.b = T.b1; wx.c = T.c1; wx.f = T.f;
    if ((e = launch_qxx_assemble(st, k, h->m, s, h->rf, wx, f0)) != hipSuccess) return hip_fail(h, e, "adjoint tape qxx");
    // the predictors solve straight onto the tape (Qx*, Qy*: unknown i is row i+1); the
    // projector assemblies read them there through a state whose Qx (Qy) points one row
    // before the tape, since they read rows 1..nx-1 only
    (void)nF;
    if ((e = launch_partition_cols(st, T.a1, T.b1, T.c1, T.f, T.qxs, nx - 1, ny, ny, h->tile_cols)) != hipSuccess)
        return hip_fail(h, e, "adjoint tape thomas x");
    wx.a = T.a2; wx.b = T.b2; wx.c = T.c2;
    State sx = s;
    sx.Qx = T.qxs - ny;
    if ((e = launch_qxy_assemble(st, k, sx, h->rf, wx)) != hipSuccess) return hip_fail(h, e, "adjoint tape qxy");
    if ((e = launch_partition_rows(st, T.a2, T.b2, T.c2, T.f, ny, s.Qx + ny, ny, nx - 1, ny, true, h->tile_rows)) !=
        hipSuccess)
        return hip_fail(h, e, "adjoint tape periodic x");
    {   // the tail writes the Qy predictor's IsWater: onto the y tape
        Work wt = wx;
        wt.w = T.wy;
        if ((e = launch_qx_tail(st, k, h->m, s, wt, f0)) != hipSuccess) return hip_fail(h, e, "adjoint tape qx tail");
    }
    wy.a = T.a3; wy.b = T.b3; wy.c = T.c3; wy.f = T.f;
    if ((e = launch_qyy_assemble(st, k, h->m, s, h->rf, wy)) != hipSuccess) return hip_fail(h, e, "adjoint tape qyy");
    if ((e = launch_partition_rows(st, T.a3, T.b3, T.c3, T.f, ny, T.qys, ny, nx - 1, ny, true, h->tile_rows)) !=
        hipSuccess)
        return hip_fail(h, e, "adjoint tape periodic y");
    wy.a = T.a4; wy.b = T.b4; wy.c = T.c4;
    State sy = s;
    sy.Qy = T.qys - ny;
    if ((e = launch_qyx_assemble(st, k, h->m, sy, h->rf, wy)) != hipSuccess) return hip_fail(h, e, "adjoint tape qyx");
    if ((e = launch_partition_cols(st, T.a4, T.b4, T.c4, T.f, s.Qy + ny, nx - 1, ny, ny, h->tile_cols)) != hipSuccess)
        return hip_fail(h, e, "adjoint tape thomas y");
    if ((e = launch_qy_tail(st, k, s)) != hipSuccess) return hip_fail(h, e, "adjoint tape qy tail");
    std::swap(s.D, s.Dn);
    a->taped_steps++;
    return M70H_OK;
}

// Reverse kernels of step t: S0/S1 = states before/after it, T = its tape.
int adj_reverse_kernels(m70h *h, int t, const State &S0, const State &S1, const StepTape &T)
{
    AdjCtx *a = h->adj;
    const int out = 1 - a->cur;
    const AdjFields &Ain = a->A[a->cur], &Aout = a->A[out];
    const AdjWork &W = a->W;
    HIPCHK(h, launch_adj_zero(h->stream, Aout, W, h->F, h->Fp));
    hipError_t e;
    if ((e = launch_adj_y(h->stream, h->k, h->m, S0, S1, h->rf, T, Ain, Aout, W, a->R, h->tile_cols)) != hipSuccess)
        return hip_fail(h, e, "adjoint y half-step");
    const double f0 = f0_at(h, a->f0_pos0 + t);
    if ((e = launch_adj_x(h->stream, h->k, h->m, S0, S1, h->rf, T, Ain, Aout, W, a->R, f0, h->tile_cols)) != hipSuccess)
        return hip_fail(h, e, "adjoint x half-step");
    a->cur = out;
    return M70H_OK;
}

// Reverses step t: h->s holds s_t, S0 is a buffer holding s_t that stays untouched.
int adj_reverse_one(m70h *h, int t, const State &S0)
{
    int rc = adj_taped_step(h, t, h->adj->T);
    if (rc) return rc;
    return adj_reverse_kernels(h, t, S0, h->s, h->adj->T);
}

// Reverses steps a0..b0-1; the state of step a0 is held by a checkpoint buffer and s free
// buffers remain.
int adj_reverse_range(m70h *h, int a0, int b0, int s)
{
    AdjCtx *a = h->adj;
    const int qa = adj_find(h, a0);
    if (qa < 0) return fail(h, M70H_ESTATE, "adjoint: missing checkpoint");
    const State Sa = buf_state(h, a->slot[qa]);
    int rc;
    if (b0 - a0 == 1) {
        HIPCHK(h, copy_state(h, h->s, Sa));
        return adj_reverse_one(h, a0, Sa);
    }
    if (s == 0) {
        for (int t = b0 - 1; t >= a0; --t) {
            HIPCHK(h, copy_state(h, h->s, Sa));
            if (t == a0) {
                if ((rc = adj_reverse_one(h, t, Sa))) return rc;
            } else {
                if ((rc = adj_advance(h, a0, t - a0))) return rc;
                const State S0 = buf_state(h, a->s0buf);
                HIPCHK(h, copy_state(h, S0, h->s));
                if ((rc = adj_reverse_one(h, t, S0))) return rc;
            }
        }
        return M70H_OK;
    }
    const int j = a->split[(size_t)(b0 - a0) * (a->S + 1) + s];
    const int m = a0 + j;
    if (adj_find(h, m) < 0) {
        HIPCHK(h, copy_state(h, h->s, Sa));
        if ((rc = adj_advance(h, a0, j))) return rc;
        if ((rc = adj_store(h, m))) return rc;
    }
    if ((rc = adj_reverse_range(h, m, b0, s - 1))) return rc;
    a->slot_step[adj_find(h, m)] = -1;
    return adj_reverse_range(h, a0, m, s);
}

}  // namespace

extern "C" {

int m70h_adjoint_forward(m70h *h, int nIter, int slots)
{
    if (!h) return M70H_EINVAL;
    if (nIter < 1) return fail(h, M70H_EINVAL, "adjoint: nIter must be >= 1");
    if (slots < 0) return fail(h, M70H_EINVAL, "adjoint: slots must be >= 1, or 0 to tape every step");
    if (!h->forcing_set) return fail(h, M70H_ESTATE, "forcing not set");
    const bool tape_all = slots == 0;
    if (!tape_all && (double)nIter * nIter * (slots < nIter ? slots : nIter) > 4e9)
        return fail(h, M70H_EINVAL, "adjoint: schedule too large (split the Run or use fewer slots)");
    int rc = adj_setup(h, nIter, slots, tape_all);
    if (rc) { adj_release(h); return rc; }
    AdjCtx *a = h->adj;
    h->run_gen++;
    a->gen = h->run_gen;
    a->f0_pos0 = h->f0_pos;
    if ((rc = run_preloop(h))) return rc;
    if (tape_all) {   // every step taped now; the reverse sweep recomputes nothing
        for (int t = 0; t < nIter; ++t) {
            HIPCHK(h, copy_state(h, buf_state(h, a->slot[t]), h->s));
            a->slot_step[t] = t;
            if ((rc = adj_taped_step(h, t, a->tapes[t]))) return rc;
        }
        HIPCHK(h, copy_state(h, buf_state(h, a->slot[nIter]), h->s));
        a->slot_step[nIter] = nIter;
        HIPCHK(h, copy_state(h, buf_state(h, a->final_buf), h->s));
        h->f0_pos = a->f0_pos0 + nIter;
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return M70H_OK;
    }
    if ((rc = adj_store(h, 0))) return rc;
    // first sweep: store the chain of checkpoints the schedule starts with
    int pos = 0, l = nIter, s = a->S;
    while (l > 1 && s > 0) {
        const int j = a->split[(size_t)l * (a->S + 1) + s];
        if ((rc = adj_advance(h, pos, j))) return rc;
        pos += j;
        if ((rc = adj_store(h, pos))) return rc;
        l -= j;
        --s;
    }
    if ((rc = adj_advance(h, pos, nIter - pos))) return rc;
    HIPCHK(h, copy_state(h, buf_state(h, a->final_buf), h->s));
    h->f0_pos = a->f0_pos0 + nIter;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return M70H_OK;
}

int m70h_adjoint_reverse(m70h *h, const m70h_fields *seed, const m70h_fields *grad, double *dparams)
{
    if (!h) return M70H_EINVAL;
    AdjCtx *a = h->adj;
    if (!a) return fail(h, M70H_ESTATE, "adjoint: no recorded Run (call m70h_adjoint_forward)");
    if (a->gen != h->run_gen) return fail(h, M70H_ESTATE, "adjoint: the handle ran or changed forcing since the recording");
    const long F = h->F, Fp = h->Fp;
    a->cur = 0;
    HIPCHK(h, zero_fields(h, a->A[0]));
    if (seed) {
        const double *src[10] = {seed->Zs, seed->D, seed->Qx, seed->Qy, seed->U, seed->V, seed->U_c, seed->V_c,
                                 seed->dQx_dx, seed->dQy_dy};
        const AdjFields &A = a->A[0];
        double *dst[10] = {A.Zs, A.D, A.Qx, A.Qy, A.U, A.V, A.Uc, A.Vc, A.dQx, A.dQy};
        const long n[10] = {F, F, Fp, Fp, Fp, Fp, F, F, F, F};
        for (int q = 0; q < 10; ++q)
            if (src[q])
                HIPCHK(h, hipMemcpyAsync(dst[q], src[q], (size_t)n[q] * sizeof(double), hipMemcpyDeviceToDevice,
                                         h->stream));
    }
    const AdjRun &R = a->R;
    double *rz[] = {R.cD_x, R.nuD_x, R.dnuD_dx, R.dnuDx_dy, R.cD_y, R.nuD_y, R.dnuD_dy, R.dnuDy_dx, R.cD, R.nuD};
    for (double *p : rz) HIPCHK(h, hipMemsetAsync(p, 0, (size_t)F * sizeof(double), h->stream));
    HIPCHK(h, hipMemsetAsync(a->params, 0, 4 * sizeof(double), h->stream));
    a->fwd_steps = 0;
    a->taped_steps = 0;
    int rc;
    if (a->tape_all) {
        for (int t = a->N - 1; t >= 0; --t)
            if ((rc = adj_reverse_kernels(h, t, buf_state(h, a->slot[t]), buf_state(h, a->slot[t + 1]), a->tapes[t])))
                return rc;
    } else if ((rc = adj_reverse_range(h, 0, a->N, a->S))) {   // the first sweep's checkpoints are still held
        return rc;
    }
    // per-Run pre-loop: drag/viscosity fields were formed from the depth at step 0
    const int q0 = adj_find(h, 0);
    if (q0 < 0) return fail(h, M70H_ESTATE, "adjoint: missing start checkpoint");
    const State S0 = buf_state(h, a->slot[q0]);
    const AdjFields &L = a->A[a->cur];
    hipError_t e;
    if ((e = launch_adj_preloop(h->stream, h->k, h->m, S0.D, h->fw, R, L.D, a->params)) != hipSuccess)
        return hip_fail(h, e, "adjoint pre-loop");
    if (grad) {
        double *dst[10] = {grad->Zs, grad->D, grad->Qx, grad->Qy, grad->U, grad->V, grad->U_c, grad->V_c,
                           grad->dQx_dx, grad->dQy_dy};
        const double *src[10] = {L.Zs, L.D, L.Qx, L.Qy, L.U, L.V, L.Uc, L.Vc, L.dQx, L.dQy};
        const long n[10] = {F, F, Fp, Fp, Fp, Fp, F, F, F, F};
        for (int q = 0; q < 10; ++q)
            if (dst[q])
                HIPCHK(h, hipMemcpyAsync(dst[q], src[q], (size_t)n[q] * sizeof(double), hipMemcpyDeviceToDevice,
                                         h->stream));
    }
    double hp[4] = {0, 0, 0, 0};
    HIPCHK(h, hipMemcpyAsync(hp, a->params, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    // leave the handle at the state after the Run, as a plain Run would
    HIPCHK(h, copy_state(h, h->s, buf_state(h, a->final_buf)));
    h->f0_pos = a->f0_pos0 + a->N;
    // the first sweep's chain is consumed; a second reverse recomputes it from step 0
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (dparams) { dparams[0] = hp[0]; dparams[1] = hp[1]; dparams[2] = hp[2]; }
    return M70H_OK;
}

int m70h_adjoint_stats(const m70h *h, long *forward_steps, long *taped_steps, int *buffers)
{
    if (!h || !h->adj) return M70H_EINVAL;
    if (forward_steps) *forward_steps = h->adj->fwd_steps;
    if (taped_steps) *taped_steps = h->adj->taped_steps;
    if (buffers) *buffers = (int)h->adj->slot.size();
    return M70H_OK;
}

int m70h_adjoint_release(m70h *h)
{
    if (!h) return M70H_EINVAL;
    adj_release(h);
    return M70H_OK;
}

}  // extern "C"

// =====================================================================================
// Grouped adjoint of many handles (m70h_adjoint_batch_*): every handle keeps its own
// checkpoints, tape and adjoint buffers (its AdjCtx), all handles share one Run length and
// checkpoint budget and so one binomial schedule, and every operation of that schedule
// (state copies, untaped steps, taped steps, reverse steps) is one launch per stage over
// all handles. Untaped steps and the per-Run pre-loop are the forward batch's grouped
// launches; the rest read a table of per-handle descriptors (AdjScen) that a pinned ring
// stages to the device ahead of each operation.
// =====================================================================================
namespace {
enum { GA_NXP1 = 0, GA_NX = 1, GA_NX1 = 2, GA_COLS = 3, GA_ROWS = 4, GA_QXR = 5, GA_COUNT = 6 };
}

struct m70h_adjoint_batch {
    static constexpr int RING = 64;
    std::vector<m70h *> hs;
    m70h_batch *fb = nullptr;        // forward batch: per-Run pre-loop and untaped steps
    hipStream_t stream = nullptr;    // the forward batch's stream
    int rb = 0;                      // row-solver block size
    int *d_start = nullptr;          // [GA_COUNT][n+1] first block of each handle per stage
    int blocks[GA_COUNT] = {};
    AdjScen *d_ring = nullptr, *h_ring = nullptr;   // RING tables of n descriptors
    hipEvent_t ev[RING] = {};
    bool ev_used[RING] = {};
    long ring_pos = 0;
    int N = 0, S = 0;
    int parity = 0;                  // D/Dn swaps since the forward batch's tables were uploaded
    int cur = 0;                     // A[cur] of every handle holds the adjoint of the current state
    bool recorded = false;
    std::string err;
};

namespace {

int abfail(m70h_adjoint_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg; else g_create_error = msg;
    return code;
}

#define ABCHK(b, expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return abfail(b, M70H_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Which state a copy reads or writes.
enum { SEL_STATE, SEL_SLOT, SEL_S0BUF, SEL_FINAL };
State ab_state(m70h *h, int sel, int q)
{
    switch (sel) {
    case SEL_SLOT: return buf_state(h, h->adj->slot[q]);
    case SEL_S0BUF: return buf_state(h, h->adj->s0buf);
    case SEL_FINAL: return buf_state(h, h->adj->final_buf);
    default: return h->s;
    }
}

// Stages a table of descriptors (fill(handle index, descriptor) on top of the defaults).
template <class Fill>
int ab_table(m70h_adjoint_batch *b, Fill fill, const AdjScen **out)
{
    const int n = (int)b->hs.size();
    const int r = (int)(b->ring_pos++ % m70h_adjoint_batch::RING);
    if (b->ev_used[r]) ABCHK(b, hipEventSynchronize(b->ev[r]));   // its previous copy has left the host buffer
    AdjScen *host = b->h_ring + (size_t)r * n;
    for (int q = 0; q < n; ++q) {
        m70h *h = b->hs[q];
        AdjCtx *a = h->adj;
        AdjScen &d = host[q];
        d = AdjScen{};
        d.k = h->k; d.m = h->m; d.rf = h->rf;
        d.s = h->s; d.S0 = h->s; d.dst = h->s; d.src = h->s;
        d.T = a->T; d.w = h->w; d.W = a->W; d.R = a->R;
        d.Ain = a->A[b->cur]; d.Aout = a->A[1 - b->cur];
        d.f0 = h->f0;
        d.S_cols = adj_cols_segments(h->nx - 1);
        d.L_cols = d.S_cols > 0 ? 1024 / d.S_cols : 0;
        d.S_rows = adj_rows_segments(h->ny, b->rb);
        fill(q, d);
    }
    AdjScen *dev = b->d_ring + (size_t)r * n;
    ABCHK(b, hipMemcpyAsync(dev, host, (size_t)n * sizeof(AdjScen), hipMemcpyHostToDevice, b->stream));
    ABCHK(b, hipEventRecord(b->ev[r], b->stream));
    b->ev_used[r] = true;
    *out = dev;
    return M70H_OK;
}

const int *ab_start(m70h_adjoint_batch *b, int stage) { return b->d_start + (size_t)stage * (b->hs.size() + 1); }

int ab_cells(m70h_adjoint_batch *b, int op, const AdjScen *D)
{
    const int stage = op == AGC_COPY || op == AGC_ZERO ? GA_NXP1 : (agc_rows(op, 4) == 4 ? GA_NX : GA_NX1);
    ABCHK(b, launch_adj_cells_grouped(b->stream, op, D, ab_start(b, stage), (int)b->hs.size(), b->blocks[stage]));
    return M70H_OK;
}
int ab_cols(m70h_adjoint_batch *b, int sys, int trans, const AdjScen *D)
{
    ABCHK(b, launch_adj_cols_grouped(b->stream, sys, trans, D, ab_start(b, GA_COLS), (int)b->hs.size(),
                                     b->blocks[GA_COLS]));
    return M70H_OK;
}
int ab_qxx_rows(m70h_adjoint_batch *b, const AdjScen *D)
{
    ABCHK(b, launch_adj_qxx_rows_grouped(b->stream, D, ab_start(b, GA_QXR), (int)b->hs.size(), b->blocks[GA_QXR]));
    return M70H_OK;
}
int ab_rows(m70h_adjoint_batch *b, int sys, int trans, const AdjScen *D)
{
    ABCHK(b, launch_adj_rows_grouped(b->stream, b->rb, sys, trans, D, ab_start(b, GA_ROWS), (int)b->hs.size(),
                                     b->blocks[GA_ROWS]));
    return M70H_OK;
}

// dst <- src for every handle.
int ab_copy(m70h_adjoint_batch *b, int dsel, int dq, int ssel, int sq)
{
    const AdjScen *D;
    int rc = ab_table(b, [&](int q, AdjScen &d) {
        d.dst = ab_state(b->hs[q], dsel, dq);
        d.src = ab_state(b->hs[q], ssel, sq);
    }, &D);
    if (rc) return rc;
    return ab_cells(b, AGC_COPY, D);
}

void ab_swap_depths(m70h_adjoint_batch *b)
{
    for (m70h *h : b->hs) std::swap(h->s.D, h->s.Dn);
    b->parity ^= 1;
}

// Uploads the forward batch's tables for the handles' current states (parity 0 from here).
int ab_sync_forward_tables(m70h_adjoint_batch *b)
{
    int rc = batch_upload(b->fb);
    if (rc) return abfail(b, rc, b->fb->err);
    b->parity = 0;
    return M70H_OK;
}

// n untaped steps of every handle (the forward batch's grouped step).
int ab_advance(m70h_adjoint_batch *b, int n)
{
    // one Run of n steps: its first step reads V_c as stored, the steps after it form V_c
    // from V, and only its last stores U_c and V_c (the same bits as n one-step Runs)
    for (int t = 0; t < n; ++t) {
        ABCHK(b, batch_step(b->fb, b->parity, t == 0, t == n - 1));
        ab_swap_depths(b);
        for (m70h *h : b->hs) h->adj->fwd_steps++;
    }
    return M70H_OK;
}

// The checkpoint buffer every handle holds `step` in (all handles keep the same layout).
int ab_find(m70h_adjoint_batch *b, int step) { return adj_find(b->hs[0], step); }

int ab_store(m70h_adjoint_batch *b, int step)
{
    AdjCtx *a0 = b->hs[0]->adj;
    for (size_t q = 0; q < a0->slot.size(); ++q)
        if (a0->slot_step[q] < 0) {
            int rc = ab_copy(b, SEL_SLOT, (int)q, SEL_STATE, 0);
            if (rc) return rc;
            for (m70h *h : b->hs) h->adj->slot_step[q] = step;
            return M70H_OK;
        }
    return abfail(b, M70H_ESTATE, "adjoint batch: no free checkpoint buffer (schedule error)");
}

// One taped step t of every handle from its current state (as adj_taped_step).
int ab_taped_step(m70h_adjoint_batch *b, int t)
{
    const AdjScen *D;
    int rc = ab_table(b, [&](int q, AdjScen &d) {
        m70h *h = b->hs[q];
        d.f0 = f0_at(h, h->adj->f0_pos0 + t);
    }, &D);
    if (rc) return rc;
    if ((rc = ab_cells(b, AGC_QXX_ASM, D)) || (rc = ab_cols(b, 1, 0, D)) || (rc = ab_cells(b, AGC_QXY_ASM, D)) ||
        (rc = ab_rows(b, 2, 0, D)) || (rc = ab_cells(b, AGC_QX_TAIL, D)) || (rc = ab_cells(b, AGC_QYY_ASM, D)) ||
        (rc = ab_rows(b, 3, 0, D)) || (rc = ab_cells(b, AGC_QYX_ASM, D)) || (rc = ab_cols(b, 4, 0, D)) ||
        (rc = ab_cells(b, AGC_QY_TAIL, D)))
        return rc;
    ab_swap_depths(b);
    for (m70h *h : b->hs) h->adj->taped_steps++;
    return M70H_OK;
}

// Reverse kernels of step t (as adj_reverse_kernels): the current states hold s_{t+1}, the
// buffers selected by s0sel/s0q hold s_t.
int ab_reverse_kernels(m70h_adjoint_batch *b, int t, int s0sel, int s0q)
{
    const AdjScen *D;
    int rc = ab_table(b, [&](int q, AdjScen &d) {
        m70h *h = b->hs[q];
        d.S0 = ab_state(h, s0sel, s0q);
        d.f0 = f0_at(h, h->adj->f0_pos0 + t);
    }, &D);
    if (rc) return rc;
    if ((rc = ab_cells(b, AGC_ZERO, D)) || (rc = ab_cells(b, AGC_QY_TAIL_A, D)) || (rc = ab_cells(b, AGC_QY_TAIL_B, D)) ||
        (rc = ab_cols(b, 4, 1, D)) || (rc = ab_cells(b, AGC_QYX, D)) || (rc = ab_rows(b, 3, 1, D)) ||
        (rc = ab_cells(b, AGC_QYY, D)) || (rc = ab_cells(b, AGC_QX_TAIL_A, D)) || (rc = ab_cells(b, AGC_QX_TAIL_B, D)) ||
        (rc = ab_rows(b, 2, 1, D)) || (rc = ab_cells(b, AGC_QXY, D)) || (rc = ab_cols(b, 1, 1, D)) ||
        (rc = ab_qxx_rows(b, D)))
        return rc;
    b->cur = 1 - b->cur;
    return M70H_OK;
}

int ab_reverse_one(m70h_adjoint_batch *b, int t, int s0sel, int s0q)
{
    int rc = ab_taped_step(b, t);
    if (rc) return rc;
    return ab_reverse_kernels(b, t, s0sel, s0q);
}

// Steps a0..b0-1 in reverse (as adj_reverse_range): step a0's state is in a checkpoint
// buffer and s buffers are free.
int ab_reverse_range(m70h_adjoint_batch *b, int a0, int b0, int s)
{
    const int qa = ab_find(b, a0);
    if (qa < 0) return abfail(b, M70H_ESTATE, "adjoint batch: missing checkpoint");
    int rc;
    if (b0 - a0 == 1) {
        if ((rc = ab_copy(b, SEL_STATE, 0, SEL_SLOT, qa))) return rc;
        return ab_reverse_one(b, a0, SEL_SLOT, qa);
    }
    if (s == 0) {
        for (int t = b0 - 1; t >= a0; --t) {
            if ((rc = ab_copy(b, SEL_STATE, 0, SEL_SLOT, qa))) return rc;
            if (t == a0) {
                if ((rc = ab_reverse_one(b, t, SEL_SLOT, qa))) return rc;
            } else {
                if ((rc = ab_advance(b, t - a0))) return rc;
                if ((rc = ab_copy(b, SEL_S0BUF, 0, SEL_STATE, 0))) return rc;
                if ((rc = ab_reverse_one(b, t, SEL_S0BUF, 0))) return rc;
            }
        }
        return M70H_OK;
    }
    AdjCtx *a0c = b->hs[0]->adj;
    const int j = a0c->split[(size_t)(b0 - a0) * (a0c->S + 1) + s];
    const int m = a0 + j;
    if (ab_find(b, m) < 0) {
        if ((rc = ab_copy(b, SEL_STATE, 0, SEL_SLOT, qa))) return rc;
        if ((rc = ab_advance(b, j))) return rc;
        if ((rc = ab_store(b, m))) return rc;
    }
    if ((rc = ab_reverse_range(b, m, b0, s - 1))) return rc;
    const int qm = ab_find(b, m);
    for (m70h *h : b->hs) h->adj->slot_step[qm] = -1;
    return ab_reverse_range(b, a0, m, s);
}

int ab_sync_handles(m70h_adjoint_batch *b)
{
    for (m70h *h : b->hs) ABCHK(b, hipStreamSynchronize(h->stream));
    return M70H_OK;
}

}  // namespace

extern "C" {

const char *m70h_adjoint_batch_last_error(const m70h_adjoint_batch *b)
{
    return b ? b->err.c_str() : g_create_error.c_str();
}

void m70h_adjoint_batch_destroy(m70h_adjoint_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->d_ring) (void)hipFree(b->d_ring);
    if (b->h_ring) (void)hipHostFree(b->h_ring);
    if (b->d_start) (void)hipFree(b->d_start);
    m70h_batch_destroy(b->fb);
    delete b;
}

int m70h_adjoint_batch_create(int nscen, m70h *const *handles, m70h_adjoint_batch **out)
{
    if (!out) return abfail(nullptr, M70H_EINVAL, "out is NULL");
    *out = nullptr;
    if (nscen < 1 || !handles) return abfail(nullptr, M70H_EINVAL, "need at least one handle");
    for (int q = 0; q < nscen; ++q) {
        m70h *h = handles[q];
        if (!h) return abfail(nullptr, M70H_EINVAL, "NULL handle");
        if (!h->f0_series.empty())
            return abfail(nullptr, M70H_EINVAL, "adjoint batch: handles with an offshore offset series are not supported");
        if (h->tile_cols != 0 || h->tile_rows != 0)
            return abfail(nullptr, M70H_EINVAL, "adjoint batch: handles must use the default solver tiles");
        if (adj_cols_segments(h->nx - 1) <= 0)
            return abfail(nullptr, M70H_EINVAL, "adjoint batch: nx-1 must be at most 2048");
    }
    m70h_adjoint_batch *b = new m70h_adjoint_batch();
    auto bail = [&](int rc) { g_create_error = b->err; m70h_adjoint_batch_destroy(b); return rc; };
    int rc = m70h_batch_create(nscen, handles, &b->fb);
    if (rc) return bail(abfail(b, rc, g_create_error));
    b->stream = b->fb->stream;
    b->rb = b->fb->rows_block;
    b->hs.assign(handles, handles + nscen);
    std::vector<int> starts((size_t)GA_COUNT * (nscen + 1), 0);
    for (int q = 0; q < nscen; ++q) {
        m70h *h = handles[q];
        const int gx = (h->ny + 255) / 256;
        const int Sc = adj_cols_segments(h->nx - 1), Lc = 1024 / Sc;
        if (adj_rows_segments(h->ny, b->rb) <= 0) return bail(abfail(b, M70H_EINVAL, "adjoint batch: ny too large"));
        const int per[GA_COUNT] = {gx * (h->nx + 1), gx * h->nx, gx * (h->nx - 1), (h->ny + Lc - 1) / Lc, h->nx - 1,
                                   adj_qxx_rows_blocks(h->nx, h->ny)};
        for (int gs = 0; gs < GA_COUNT; ++gs)
            starts[(size_t)gs * (nscen + 1) + q + 1] = starts[(size_t)gs * (nscen + 1) + q] + per[gs];
    }
    for (int gs = 0; gs < GA_COUNT; ++gs) b->blocks[gs] = starts[(size_t)gs * (nscen + 1) + nscen];
    hipError_t e;
    if ((e = hipMalloc(&b->d_start, starts.size() * sizeof(int))) != hipSuccess ||
        (e = hipMemcpy(b->d_start, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMalloc(&b->d_ring, (size_t)m70h_adjoint_batch::RING * nscen * sizeof(AdjScen))) != hipSuccess ||
        (e = hipHostMalloc((void **)&b->h_ring, (size_t)m70h_adjoint_batch::RING * nscen * sizeof(AdjScen),
                           hipHostMallocDefault)) != hipSuccess)
        return bail(abfail(b, M70H_EHIP, std::string("adjoint batch allocation: ") + hipGetErrorString(e)));
    for (auto &ev : b->ev)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
            return bail(abfail(b, M70H_EHIP, std::string("adjoint batch events: ") + hipGetErrorString(e)));
    *out = b;
    return M70H_OK;
}

int m70h_adjoint_batch_forward(m70h_adjoint_batch *b, int nIter, int slots)
{
    if (!b) return M70H_EINVAL;
    if (nIter < 1) return abfail(b, M70H_EINVAL, "adjoint batch: nIter must be >= 1");
    if (slots < 1) return abfail(b, M70H_EINVAL, "adjoint batch: slots must be >= 1");
    if ((double)nIter * nIter * (slots < nIter ? slots : nIter) > 4e9)
        return abfail(b, M70H_EINVAL, "adjoint batch: schedule too large (split the Run or use fewer slots)");
    int rc = ab_sync_handles(b);
    if (rc) return rc;
    for (m70h *h : b->hs) {
        if (!h->forcing_set) return abfail(b, M70H_ESTATE, "forcing not set on a handle");
        if ((rc = adj_setup(h, nIter, slots, false))) {
            std::string msg = h->err;
            adj_release(h);
            return abfail(b, rc, msg);
        }
        AdjCtx *a = h->adj;
        h->run_gen++;
        a->gen = h->run_gen;
        a->f0_pos0 = h->f0_pos;
    }
    b->N = nIter;
    b->S = b->hs[0]->adj->S;
    b->recorded = false;
    if ((rc = ab_sync_forward_tables(b))) return rc;
    const int n = (int)b->hs.size();
    // per-Run pre-loop of every handle (src/Hydro.cpp:175-192)
    ABCHK(b, launch_cells_grouped(b->stream, 0, b->fb->d_desc, b->fb->d_start + GS_CELLS * (n + 1), n,
                                  b->fb->blocks[GS_CELLS]));
    ABCHK(b, launch_cells_grouped(b->stream, 1, b->fb->d_desc, b->fb->d_start + GS_CELLS * (n + 1), n,
                                  b->fb->blocks[GS_CELLS]));
    if ((rc = ab_store(b, 0))) return rc;
    // first sweep: the chain of checkpoints the schedule starts with
    AdjCtx *a0 = b->hs[0]->adj;
    int pos = 0, l = nIter, s = b->S;
    while (l > 1 && s > 0) {
        const int j = a0->split[(size_t)l * (a0->S + 1) + s];
        if ((rc = ab_advance(b, j))) return rc;
        pos += j;
        if ((rc = ab_store(b, pos))) return rc;
        l -= j;
        --s;
    }
    if ((rc = ab_advance(b, nIter - pos))) return rc;
    if ((rc = ab_copy(b, SEL_FINAL, 0, SEL_STATE, 0))) return rc;
    for (m70h *h : b->hs) h->f0_pos = h->adj->f0_pos0 + nIter;
    ABCHK(b, hipStreamSynchronize(b->stream));
    b->recorded = true;
    return M70H_OK;
}

int m70h_adjoint_batch_reverse(m70h_adjoint_batch *b, const m70h_fields *seeds, const m70h_fields *grads,
                               double *dparams)
{
    if (!b) return M70H_EINVAL;
    if (!b->recorded) return abfail(b, M70H_ESTATE, "adjoint batch: no recorded Run (call m70h_adjoint_batch_forward)");
    for (m70h *h : b->hs)
        if (!h->adj || h->adj->gen != h->run_gen || h->adj->N != b->N)
            return abfail(b, M70H_ESTATE, "adjoint batch: a handle ran or changed since the recording");
    int rc = ab_sync_handles(b);
    if (rc) return rc;
    const int n = (int)b->hs.size();
    hipStream_t st = b->stream;
    // adjoint of the final states, zeroed per-Run accumulators
    for (int q = 0; q < n; ++q) {
        m70h *h = b->hs[q];
        AdjCtx *a = h->adj;
        const long F = h->F, Fp = h->Fp;
        const AdjFields &A = a->A[0];
        double *dst[10] = {A.Zs, A.D, A.Qx, A.Qy, A.U, A.V, A.Uc, A.Vc, A.dQx, A.dQy};
        const long len[10] = {F, F, Fp, Fp, Fp, Fp, F, F, F, F};
        const double *src[10] = {};
        if (seeds) {
            const m70h_fields &sd = seeds[q];
            const double *s10[10] = {sd.Zs, sd.D, sd.Qx, sd.Qy, sd.U, sd.V, sd.U_c, sd.V_c, sd.dQx_dx, sd.dQy_dy};
            for (int f = 0; f < 10; ++f) src[f] = s10[f];
        }
        for (int f = 0; f < 10; ++f) {
            if (src[f]) ABCHK(b, hipMemcpyAsync(dst[f], src[f], (size_t)len[f] * sizeof(double), hipMemcpyDeviceToDevice, st));
            else ABCHK(b, hipMemsetAsync(dst[f], 0, (size_t)len[f] * sizeof(double), st));
        }
        const AdjRun &R = a->R;
        double *rz[] = {R.cD_x, R.nuD_x, R.dnuD_dx, R.dnuDx_dy, R.cD_y, R.nuD_y, R.dnuD_dy, R.dnuDy_dx, R.cD, R.nuD};
        for (double *p : rz) ABCHK(b, hipMemsetAsync(p, 0, (size_t)F * sizeof(double), st));
        ABCHK(b, hipMemsetAsync(a->params, 0, 4 * sizeof(double), st));
        a->fwd_steps = 0;
        a->taped_steps = 0;
    }
    b->cur = 0;
    if ((rc = ab_sync_forward_tables(b))) return rc;
    if ((rc = ab_reverse_range(b, 0, b->N, b->S))) return rc;
    // per-Run pre-loop reverse, gradients out, handles back at their final states
    std::vector<double> hp(4 * (size_t)n, 0.0);
    double *hpp = nullptr;
    ABCHK(b, hipHostMalloc((void **)&hpp, 4 * (size_t)n * sizeof(double), hipHostMallocDefault));
    for (int q = 0; q < n; ++q) {
        m70h *h = b->hs[q];
        AdjCtx *a = h->adj;
        a->cur = b->cur;
        const int q0 = adj_find(h, 0);
        if (q0 < 0) { (void)hipHostFree(hpp); return abfail(b, M70H_ESTATE, "adjoint batch: missing start checkpoint"); }
        const State S0 = buf_state(h, a->slot[q0]);
        const AdjFields &L = a->A[a->cur];
        hipError_t e = launch_adj_preloop(st, h->k, h->m, S0.D, h->fw, a->R, L.D, a->params);
        if (e != hipSuccess) { (void)hipHostFree(hpp); return abfail(b, M70H_EHIP, std::string("adjoint pre-loop: ") + hipGetErrorString(e)); }
        if (grads) {
            const m70h_fields &gd = grads[q];
            double *dst[10] = {gd.Zs, gd.D, gd.Qx, gd.Qy, gd.U, gd.V, gd.U_c, gd.V_c, gd.dQx_dx, gd.dQy_dy};
            const double *src[10] = {L.Zs, L.D, L.Qx, L.Qy, L.U, L.V, L.Uc, L.Vc, L.dQx, L.dQy};
            const long F = h->F, Fp = h->Fp;
            const long len[10] = {F, F, Fp, Fp, Fp, Fp, F, F, F, F};
            for (int f = 0; f < 10; ++f)
                if (dst[f]) {
                    e = hipMemcpyAsync(dst[f], src[f], (size_t)len[f] * sizeof(double), hipMemcpyDeviceToDevice, st);
                    if (e != hipSuccess) { (void)hipHostFree(hpp); return abfail(b, M70H_EHIP, hipGetErrorString(e)); }
                }
        }
        e = hipMemcpyAsync(hpp + 4 * q, a->params, 4 * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) { (void)hipHostFree(hpp); return abfail(b, M70H_EHIP, hipGetErrorString(e)); }
        h->f0_pos = a->f0_pos0 + a->N;
    }
    if ((rc = ab_copy(b, SEL_STATE, 0, SEL_FINAL, 0))) { (void)hipHostFree(hpp); return rc; }
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && dparams)
        for (int q = 0; q < n; ++q)
            for (int k = 0; k < 3; ++k) dparams[3 * q + k] = hpp[4 * q + k];
    (void)hipHostFree(hpp);
    if (e != hipSuccess) return abfail(b, M70H_EHIP, std::string("adjoint batch: ") + hipGetErrorString(e));
    b->recorded = false;   // the first sweep's chain is consumed
    return M70H_OK;
}

}  // extern "C"

// =====================================================================================
// Sediment transport and bed update (class Sediment, src/Sediment.cpp)
// =====================================================================================
struct m70s {
    m70h *h = nullptr;
    m70s_params p{};
    std::string err;
    std::vector<void *> allocs;
    double *Alpha = nullptr, *gamma = nullptr, *Gamma = nullptr, *qx = nullptr, *qy = nullptr, *dh_dt = nullptr;
    double *bed_h = nullptr, *Zb = nullptr, *maxDw = nullptr, *Zs_sh = nullptr;
    double *part_max = nullptr;
    int *part_last = nullptr;
};

namespace {

int sfail(m70s *s, int code, const std::string &msg)
{
    if (s) s->err = msg; else g_create_error = msg;
    return code;
}

#define SHIPCHK(s, expr)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return sfail(s, M70H_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

SedArgs sed_args(m70s *s)
{
    const m70h *h = s->h;
    SedArgs a{};
    a.nx = h->nx;
    a.ny = h->ny;
    a.Dmin = h->p.Dmin;
    a.inv_dy = 1 / h->dy;
    a.inv_dx = h->m.inv_dx;
    a.inv_dxx = h->m.inv_dxx;
    a.p = s->p;
    a.D = h->s.D; a.U = h->s.U; a.V = h->s.V; a.Dw = h->fw.Dw; a.Zs = h->s.Zs;
    a.Alpha = s->Alpha; a.gamma = s->gamma; a.Gamma = s->Gamma; a.qx = s->qx; a.qy = s->qy; a.dh_dt = s->dh_dt;
    a.h = s->bed_h; a.Zb = s->Zb; a.maxDw = s->maxDw; a.Zs_sh = s->Zs_sh;
    a.part_max = s->part_max; a.part_last = s->part_last;
    return a;
}

}  // namespace

extern "C" {

const char *m70s_last_error(const m70s *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int m70s_create(m70h *hydro, const double *hb, const double *Zb, const m70s_params *p, m70s **out)
{
    if (!out) return sfail(nullptr, M70H_EINVAL, "out is NULL");
    *out = nullptr;
    if (!hydro || !hb || !Zb || !p) return sfail(nullptr, M70H_EINVAL, "hydro, h, Zb and p must be non-NULL");
    if (!(p->p < 1.0) || !(p->D_alpha0_min != p->D_alpha0_max))
        return sfail(nullptr, M70H_EINVAL, "need p < 1 and D_alpha0_min != D_alpha0_max");
    m70s *s = new m70s();
    s->h = hydro;
    s->p = *p;
    const long F = hydro->F;
    double **fp[] = {&s->Alpha, &s->gamma, &s->Gamma, &s->qx, &s->qy, &s->dh_dt, &s->bed_h, &s->Zb};
    for (auto q : fp) {
        void *m = nullptr;
        hipError_t e = hipMalloc(&m, (size_t)F * sizeof(double));
        if (e == hipSuccess) e = hipMemset(m, 0, (size_t)F * sizeof(double));
        if (e != hipSuccess) {
            g_create_error = std::string("m70s_create: ") + hipGetErrorString(e);
            m70s_destroy(s);
            return M70H_EHIP;
        }
        s->allocs.push_back(m);
        *q = (double *)m;
    }
    for (double **q : {&s->maxDw, &s->Zs_sh}) {
        void *m = nullptr;
        hipError_t e = hipMalloc(&m, (size_t)hydro->ny * sizeof(double));
        if (e == hipSuccess) e = hipMemset(m, 0, (size_t)hydro->ny * sizeof(double));
        if (e != hipSuccess) {
            g_create_error = std::string("m70s_create: ") + hipGetErrorString(e);
            m70s_destroy(s);
            return M70H_EHIP;
        }
        s->allocs.push_back(m);
        *q = (double *)m;
    }
    for (void **q : {(void **)&s->part_max, (void **)&s->part_last}) {
        void *m = nullptr;
        const size_t nb = (size_t)SED_MAX_CHUNKS * hydro->ny * (q == (void **)&s->part_max ? sizeof(double) : sizeof(int));
        hipError_t e = hipMalloc(&m, nb);
        if (e != hipSuccess) {
            g_create_error = std::string("m70s_create: ") + hipGetErrorString(e);
            m70s_destroy(s);
            return M70H_EHIP;
        }
        s->allocs.push_back(m);
        *q = m;
    }
    int rc = m70s_set_bed(s, hb, Zb);
    if (rc) {
        g_create_error = s->err;
        m70s_destroy(s);
        return rc;
    }
    *out = s;
    return M70H_OK;
}

void m70s_destroy(m70s *s)
{
    if (!s) return;
    if (s->h && s->h->stream) (void)hipStreamSynchronize(s->h->stream);
    for (void *q : s->allocs) (void)hipFree(q);
    delete s;
}

int m70s_set_bed(m70s *s, const double *hb, const double *Zb)
{
    if (!s) return M70H_EINVAL;
    const size_t nb = (size_t)s->h->F * sizeof(double);
    if (hb) SHIPCHK(s, hipMemcpyAsync(s->bed_h, hb, nb, hipMemcpyHostToDevice, s->h->stream));
    if (Zb) SHIPCHK(s, hipMemcpyAsync(s->Zb, Zb, nb, hipMemcpyHostToDevice, s->h->stream));
    SHIPCHK(s, hipStreamSynchronize(s->h->stream));
    return M70H_OK;
}

int m70s_calculate_fluxes_async(m70s *s)
{
    if (!s) return M70H_EINVAL;
    m70h *h = s->h;
    if (!h->forcing_set) return sfail(s, M70H_ESTATE, "the hydro handle has no forcing (Dw)");
    const SedArgs a = sed_args(s);
    int rc;
    if ((rc = timed(h, K_SED_COLUMNS, [&] { return launch_sed_columns(h->stream, a); })) ||
        (rc = timed(h, K_SED_FLUXES, [&] { return launch_sed_coeffs_fluxes(h->stream, a); })))
        return sfail(s, rc, h->err);
    return M70H_OK;
}

int m70s_update_bottom_async(m70s *s, double dt)
{
    if (!s) return M70H_EINVAL;
    m70h *h = s->h;
    const SedArgs a = sed_args(s);
    int rc = timed(h, K_SED_UPDATE, [&] { return launch_sed_update(h->stream, a, dt); });
    if (rc) return sfail(s, rc, h->err);
    h->run_gen++;   // Zs moved with the bed
    return M70H_OK;
}

int m70s_calculate_fluxes(m70s *s)
{
    int rc = m70s_calculate_fluxes_async(s);
    if (rc) return rc;
    if ((rc = m70h_synchronize(s->h))) return sfail(s, rc, s->h->err);
    return M70H_OK;
}

int m70s_update_bottom(m70s *s, double dt)
{
    int rc = m70s_update_bottom_async(s, dt);
    if (rc) return rc;
    if ((rc = m70h_synchronize(s->h))) return sfail(s, rc, s->h->err);
    return M70H_OK;
}

int m70s_set_fluxes(m70s *s, const double *qx, const double *qy)
{
    if (!s || !qx || !qy) return M70H_EINVAL;
    const size_t nb = (size_t)s->h->F * sizeof(double);
    SHIPCHK(s, hipMemcpyAsync(s->qx, qx, nb, hipMemcpyHostToDevice, s->h->stream));
    SHIPCHK(s, hipMemcpyAsync(s->qy, qy, nb, hipMemcpyHostToDevice, s->h->stream));
    SHIPCHK(s, launch_sed_dhdt(s->h->stream, sed_args(s)));
    SHIPCHK(s, hipStreamSynchronize(s->h->stream));
    return M70H_OK;
}

int m70s_get(m70s *s, double *Alpha, double *gamma, double *Gamma, double *qx, double *qy, double *dh_dt, double *hb,
             double *Zb)
{
    if (!s) return M70H_EINVAL;
    double *dst[8] = {Alpha, gamma, Gamma, qx, qy, dh_dt, hb, Zb};
    const double *src[8] = {s->Alpha, s->gamma, s->Gamma, s->qx, s->qy, s->dh_dt, s->bed_h, s->Zb};
    const size_t nb = (size_t)s->h->F * sizeof(double);
    for (int q = 0; q < 8; ++q)
        if (dst[q]) SHIPCHK(s, hipMemcpyAsync(dst[q], src[q], nb, hipMemcpyDeviceToHost, s->h->stream));
    SHIPCHK(s, hipStreamSynchronize(s->h->stream));
    return M70H_OK;
}

}  // extern "C"

// =====================================================================================
// Wave and rollers driver (class Wave, src/Wave.cpp; class WaveRow)
// =====================================================================================
struct m70w {
    int nx = 0, ny = 0;
    double dy = 0.0;
    m70w_params p{};
    std::string err;
    std::vector<void *> allocs;
    hipStream_t own_stream = nullptr, stream = nullptr;
    WaveArgs a{};
    double *D = nullptr, *U = nullptr, *V = nullptr, *U_w = nullptr;   // owned inputs
    WaveCtlBufs cb{};
    double T = 0.0;      // wave period of the pending field
    int next_s = 0;      // next marching launch of the pending field (launch 0 is the first row)
    long last_steps = -1;
    bool pending = false;
    m70h *hydro = nullptr;
};

namespace {

int wfail(m70w *w, int code, const std::string &msg)
{
    if (w) w->err = msg; else g_create_error = msg;
    return code;
}

#define WHIPCHK(w, expr)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return wfail(w, M70H_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

hipError_t walloc(m70w *w, double **p, size_t n)
{
    void *m = nullptr;
    hipError_t e = hipMalloc(&m, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(m, 0, n * sizeof(double));
    if (e != hipSuccess) return e;
    w->allocs.push_back(m);
    *p = (double *)m;
    return hipSuccess;
}

// Completes the pending field: reads the marching state after the last launched step
// and launches more steps (intermediate rows) until the last mesh row is reached, then
// the field-wide launch that forms the derived arrays.
int wave_finish(m70w *w)
{
    if (!w->pending) return M70H_OK;
    const long limit = 1000L * w->nx + 1000 + 64;
    for (;;) {
        WaveCtl c{};   // written by the last launch
        WHIPCHK(w, hipMemcpyAsync(&c, w->cb.ctl + ((w->next_s - 1) & 1), sizeof(c), hipMemcpyDeviceToHost,
                                  w->stream));
        WHIPCHK(w, hipStreamSynchronize(w->stream));
        if (c.done) {
            w->pending = false;
            w->last_steps = c.steps;
            if (c.err) return wfail(w, M70H_ENONFINITE, "wave field: the march took more than 1000*nx steps");
            // the arrays storageRow derives from the marched ones, for the whole field
            WHIPCHK(w, launch_wave_field_terms(w->stream, w->a));
            return M70H_OK;
        }
        if (w->next_s > limit) {
            w->pending = false;
            return wfail(w, M70H_ENONFINITE, "wave field: the march did not finish");
        }
        WHIPCHK(w, launch_wave_steps(w->stream, w->a, w->T, w->next_s, 16, w->cb));
        w->next_s += 16;
    }
}

}  // namespace

extern "C" {

const char *m70w_last_error(const m70w *w) { return w ? w->err.c_str() : g_create_error.c_str(); }

int m70w_create(int nx, int ny, const double *x, double dy, const m70w_params *p, m70w **out)
{
    if (!out) return wfail(nullptr, M70H_EINVAL, "out is NULL");
    *out = nullptr;
    if (nx < 2 || ny < 3) return wfail(nullptr, M70H_EINVAL, "need nx >= 2 and ny >= 3");
    if (!x || !p || !(dy > 0)) return wfail(nullptr, M70H_EINVAL, "x and p must be non-NULL and dy > 0");
    if (p->breaking_model < M70W_BREAK_CT || p->breaking_model > M70W_BREAK_REGULAR)
        return wfail(nullptr, M70H_EINVAL, "breaking_model must be ChurchThornton, ThorntonGuza or RegularWaves");
    if (p->friction_model != M70W_FRIC_CONSTANT && p->friction_model != M70W_FRIC_SOULSBY97)
        return wfail(nullptr, M70H_EINVAL, "friction_model must be Constant or Soulsby97");
    for (int i = 0; i + 1 < nx; ++i)
        if (!(x[i + 1] > x[i])) return wfail(nullptr, M70H_EINVAL, "x must increase");
    m70w *w = new m70w();
    w->nx = nx;
    w->ny = ny;
    w->dy = dy;
    w->p = *p;
    auto bail = [&](hipError_t e) {
        g_create_error = std::string("m70w_create: ") + hipGetErrorString(e);
        m70w_destroy(w);
        return M70H_EHIP;
    };
    hipError_t e = hipStreamCreate(&w->own_stream);
    if (e != hipSuccess) return bail(e);
    w->stream = w->own_stream;
    const size_t F = (size_t)nx * ny;
    double *xd = nullptr, *dxd = nullptr;
    if ((e = walloc(w, &xd, nx)) != hipSuccess || (e = walloc(w, &dxd, nx)) != hipSuccess) return bail(e);
    std::vector<double> dx(nx, 0.0);
    for (int i = 0; i + 1 < nx; ++i) dx[i] = x[i + 1] - x[i];   // src/Grid.cpp:32-34
    if ((e = hipMemcpy(xd, x, nx * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return bail(e);
    if ((e = hipMemcpy(dxd, dx.data(), nx * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return bail(e);
    for (double **q : {&w->D, &w->U, &w->V, &w->U_w})
        if ((e = walloc(w, q, F)) != hipSuccess) return bail(e);
    double **fld = &w->a.f.K;
    for (int q = 0; q < WFIELD_ARRAYS; ++q)
        if ((e = walloc(w, &fld[q], F)) != hipSuccess) return bail(e);
    for (double **q : {&w->a.disp.lK, &w->a.disp.Sigma, &w->a.disp.C, &w->a.disp.Cg, &w->a.disp.Urms_H})
        if ((e = walloc(w, q, F)) != hipSuccess) return bail(e);
    for (int r = 0; r < 2; ++r)
        if ((e = walloc(w, &w->a.rows[r], (size_t)WKMAX * WROW_ARRAYS * ny)) != hipSuccess) return bail(e);
    void *cbm = nullptr;
    const size_t cb_bytes = 2 * sizeof(WaveCtl) + 3 * WACC_SLOTS * sizeof(unsigned long long);
    if ((e = hipMalloc(&cbm, cb_bytes)) != hipSuccess) return bail(e);
    w->allocs.push_back(cbm);
    w->cb.ctl = (WaveCtl *)cbm;
    w->cb.acc = (unsigned long long *)((char *)cbm + 2 * sizeof(WaveCtl));
    WaveArgs &a = w->a;
    a.nx = nx;
    a.ny = ny;
    a.dy = dy;
    a.inv_2dy = 0.5 / dy;   // WaveRow::WaveRow (src/WaveRow.cpp:63)
    a.x = xd;
    a.dx = dxd;
    a.p = *p;
    a.D = w->D;
    a.U = w->U;
    a.V = w->V;
    a.U_w = w->U_w;
    *out = w;
    return M70H_OK;
}

void m70w_destroy(m70w *w)
{
    if (!w) return;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    for (void *q : w->allocs) (void)hipFree(q);
    if (w->own_stream) (void)hipStreamDestroy(w->own_stream);
    delete w;
}

int m70w_set_stream(m70w *w, void *hip_stream)
{
    if (!w) return M70H_EINVAL;
    int rc = wave_finish(w);
    if (rc) return rc;
    WHIPCHK(w, hipStreamSynchronize(w->stream));
    w->stream = hip_stream ? (hipStream_t)hip_stream : w->own_stream;
    return M70H_OK;
}

int m70w_set_depth(m70w *w, const double *D)
{
    if (!w || !D) return w ? wfail(w, M70H_EINVAL, "D is NULL") : M70H_EINVAL;
    int rc = wave_finish(w);
    if (rc) return rc;
    WHIPCHK(w, hipMemcpyAsync(w->D, D, (size_t)w->nx * w->ny * sizeof(double), hipMemcpyHostToDevice, w->stream));
    WHIPCHK(w, hipStreamSynchronize(w->stream));
    w->a.D = w->D;
    return M70H_OK;
}

int m70w_set_uv(m70w *w, const double *U, const double *V)
{
    if (!w || !U || !V) return w ? wfail(w, M70H_EINVAL, "U and V are required") : M70H_EINVAL;
    int rc = wave_finish(w);
    if (rc) return rc;
    const long F = (long)w->nx * w->ny;
    double *tU = nullptr, *tV = nullptr;
    WHIPCHK(w, hipMalloc((void **)&tU, (size_t)F * sizeof(double)));
    hipError_t e = hipMalloc((void **)&tV, (size_t)F * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(tU, U, (size_t)F * sizeof(double), hipMemcpyHostToDevice, w->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tV, V, (size_t)F * sizeof(double), hipMemcpyHostToDevice, w->stream);
    if (e == hipSuccess) e = launch_wave_set_uv(w->stream, w->a, w->U, w->V, w->U_w, tU, tV, F);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    (void)hipFree(tU);
    if (tV) (void)hipFree(tV);
    if (e != hipSuccess) return wfail(w, M70H_EHIP, std::string("m70w_set_uv: ") + hipGetErrorString(e));
    return M70H_OK;
}

int m70w_calculate_async(m70w *w, double hOff, double thetaOff, double T)
{
    if (!w) return M70H_EINVAL;
    if (!(T > 0) || !(hOff > 0)) return wfail(w, M70H_EINVAL, "need T > 0 and hOff > 0");
    int rc = wave_finish(w);
    if (rc) return rc;
    if (w->hydro) w->a.D = w->hydro->s.D;
    // every row a step reads was written by the step before it (the first row by launch 0),
    // so the row buffers need no clearing; the accumulators do
    WHIPCHK(w, hipMemsetAsync(w->cb.acc, 0, 3 * WACC_SLOTS * sizeof(unsigned long long), w->stream));
    w->T = T;
    WHIPCHK(w, launch_wave_dispersion(w->stream, w->a, T));
    WHIPCHK(w, launch_wave_first(w->stream, w->a, hOff, thetaOff, T, w->cb));
    // the launches a field of mesh rows only takes; wave_finish adds what intermediate
    // rows need
    const int n0 = wave_march_launches(w->nx);
    WHIPCHK(w, launch_wave_steps(w->stream, w->a, T, 1, n0, w->cb));
    w->next_s = 1 + n0;
    w->pending = true;
    return M70H_OK;
}

int m70w_calculate(m70w *w, double hOff, double thetaOff, double T)
{
    int rc = m70w_calculate_async(w, hOff, thetaOff, T);
    if (rc) return rc;
    return wave_finish(w);
}

int m70w_get(m70w *w, const char *name, double *out)
{
    if (!w || !name || !out) return w ? wfail(w, M70H_EINVAL, "name and out are required") : M70H_EINVAL;
    int rc = wave_finish(w);
    if (rc) return rc;
    static const char *const names[WFIELD_ARRAYS] = {"K", "Theta", "Cg", "H", "E", "Urms", "Sxx_w", "Sxy_w",
                                                     "Syy_w", "Dw", "E_r", "D_r", "Sxx_r", "Sxy_r", "Syy_r", "Sxx",
                                                     "Sxy", "Syy", "CosTheta", "SinTheta"};
    const double *src = nullptr;
    for (int q = 0; q < WFIELD_ARRAYS; ++q)
        if (std::strcmp(name, names[q]) == 0) src = (&w->a.f.K)[q];
    if (std::strcmp(name, "U_w") == 0) src = w->U_w;
    if (!src) return wfail(w, M70H_EINVAL, std::string("unknown wave field ") + name);
    WHIPCHK(w, hipMemcpyAsync(out, src, (size_t)w->nx * w->ny * sizeof(double), hipMemcpyDeviceToHost, w->stream));
    WHIPCHK(w, hipStreamSynchronize(w->stream));
    return M70H_OK;
}

int m70w_steps(m70w *w, long *steps)
{
    if (!w || !steps) return M70H_EINVAL;
    int rc = wave_finish(w);
    if (rc) return rc;
    *steps = w->last_steps;
    return M70H_OK;
}

int m70w_couple_hydro(m70w *w, m70h *h)
{
    if (!w || !h) return w ? wfail(w, M70H_EINVAL, "hydro is NULL") : M70H_EINVAL;
    if (h->nx != w->nx || h->ny != w->ny) return wfail(w, M70H_EINVAL, "the hydro handle's grid differs");
    int rc = wave_finish(w);
    if (rc) return rc;
    WHIPCHK(w, hipStreamSynchronize(w->stream));
    w->hydro = h;
    w->stream = h->stream;
    w->a.D = h->s.D;
    return M70H_OK;
}

int m70w_set_uv_from_hydro(m70w *w)
{
    if (!w) return M70H_EINVAL;
    if (!w->hydro) return wfail(w, M70H_ESTATE, "no hydro handle coupled");
    const long F = (long)w->nx * w->ny;
    WHIPCHK(w, launch_wave_set_uv(w->stream, w->a, w->U, w->V, w->U_w, w->hydro->s.Uc, w->hydro->s.Vc, F));
    return M70H_OK;
}

int m70w_push_forcing(m70w *w)
{
    if (!w) return M70H_EINVAL;
    m70h *h = w->hydro;
    if (!h) return wfail(w, M70H_ESTATE, "no hydro handle coupled");
    int rc = wave_finish(w);   // a field with intermediate rows needs more steps than nx-1
    if (rc) return rc;
    const WField &f = w->a.f;
    const double *src[7] = {f.Sxx, f.Sxy, f.Syy, f.Urms, f.Dw, f.D_r, f.H};
    for (int q = 0; q < 7; ++q)
        WHIPCHK(w, hipMemcpyAsync(h->fw_buf[q], src[q], (size_t)h->F * sizeof(double), hipMemcpyDeviceToDevice,
                                  w->stream));
    h->forcing_set = true;
    h->run_gen++;
    return M70H_OK;
}

}  // extern "C"

// =====================================================================================
// Morphodynamic outer loop on the device (Morfo70::Run, src/Morfo70.cpp:131-218)
// =====================================================================================
namespace {

struct SnapField {
    const char *name;
    bool plus;   // (nx+1)*ny
};
// The arrays saveTimeStep writes (src/Morfo70.cpp:225-276), full-size where the file splits
// them into a row and the rest (U0/U_x, Qx0/Qx_x, V0/V_y, Qy0/Qy_y).
const SnapField kSnapFields[] = {
    {"Zs", false}, {"D", false}, {"U", false}, {"V", false}, {"U_full", true}, {"V_full", true},
    {"Qx_full", true}, {"Qy_full", true}, {"dQx_dx", false}, {"dQy_dy", false}, {"H", false}, {"E", false},
    {"Cg", false}, {"lK", false}, {"Theta", false}, {"Dw", false}, {"E_r", false}, {"qx", false},
    {"qy", false}, {"Gamma_Swash", false}, {"Alpha", false}, {"gamma", false}, {"h", false}};
constexpr int kSnapCount = sizeof(kSnapFields) / sizeof(kSnapFields[0]);

}  // namespace

struct m70o {
    m70h *h = nullptr;
    m70w *w = nullptr;
    m70s *s = nullptr;
    m70o_params p{};
    double H_off0 = 0.01;
    std::string err;
    double time = 0.0, lastSavedTime = -1;
    std::vector<void *> dev;
    double *prev[6] = {};
    unsigned long long *dout = nullptr;
    double *rec = nullptr;   // pinned host, 8 per step
    long rec_cap = 0, rec_n = 0;
    hipStream_t copy_stream = nullptr;
    int max_snap = 0;
    struct Slot {
        double time = 0.0;
        long index = -1;
        double *dev = nullptr, *host = nullptr;
        hipEvent_t copied = nullptr;
    };
    std::vector<Slot> slots;
    long snap_total = 0;
    size_t snap_doubles = 0;
};

namespace {

int ofail(m70o *o, int code, const std::string &msg)
{
    if (o) o->err = msg; else g_create_error = msg;
    return code;
}

#define OHIPCHK(o, expr)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return ofail(o, M70H_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Device sources of the snapshot fields, in kSnapFields order (NULL: zero).
void snap_sources(m70o *o, const double **src)
{
    const State &st = o->h->s;
    const WField &f = o->w->a.f;
    const double *hs[] = {st.Zs, st.D, st.Uc, st.Vc, st.U, st.V, st.Qx, st.Qy, st.dQx, st.dQy,
                          f.H, f.E, f.Cg, f.K, f.Theta, f.Dw, f.E_r};
    int q = 0;
    for (const double *p : hs) src[q++] = p;
    const m70s *s = o->s;
    src[q++] = s ? s->qx : nullptr;
    src[q++] = s ? s->qy : nullptr;
    src[q++] = s ? s->Gamma : nullptr;
    src[q++] = s ? s->Alpha : nullptr;
    src[q++] = s ? s->gamma : nullptr;
    src[q++] = s ? s->bed_h : nullptr;
}

int take_snapshot(m70o *o)
{
    const long F = o->h->F, Fp = o->h->Fp;
    const int k = (int)(o->snap_total % o->max_snap);
    m70o::Slot &sl = o->slots[k];
    if (!sl.dev) {
        OHIPCHK(o, hipMalloc((void **)&sl.dev, o->snap_doubles * sizeof(double)));
        OHIPCHK(o, hipHostMalloc((void **)&sl.host, o->snap_doubles * sizeof(double), hipHostMallocDefault));
        OHIPCHK(o, hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    } else {
        OHIPCHK(o, hipStreamWaitEvent(o->h->stream, sl.copied, 0));   // the slot's last copy is out
    }
    const double *src[kSnapCount];
    snap_sources(o, src);
    size_t off = 0;
    for (int q = 0; q < kSnapCount; ++q) {
        const size_t n = (size_t)(kSnapFields[q].plus ? Fp : F);
        if (src[q])
            OHIPCHK(o, hipMemcpyAsync(sl.dev + off, src[q], n * sizeof(double), hipMemcpyDeviceToDevice, o->h->stream));
        else
            OHIPCHK(o, hipMemsetAsync(sl.dev + off, 0, n * sizeof(double), o->h->stream));
        off += n;
    }
    hipEvent_t ready;
    OHIPCHK(o, hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    OHIPCHK(o, hipEventRecord(ready, o->h->stream));
    OHIPCHK(o, hipStreamWaitEvent(o->copy_stream, ready, 0));
    OHIPCHK(o, hipMemcpyAsync(sl.host, sl.dev, o->snap_doubles * sizeof(double), hipMemcpyDeviceToHost,
                              o->copy_stream));
    OHIPCHK(o, hipEventRecord(sl.copied, o->copy_stream));
    OHIPCHK(o, hipEventDestroy(ready));
    sl.time = o->time;
    sl.index = o->snap_total++;
    return M70H_OK;
}

int ensure_records(m70o *o, long need)
{
    if (need <= o->rec_cap) return M70H_OK;
    long cap = o->rec_cap ? o->rec_cap : 64;
    while (cap < need) cap *= 2;
    OHIPCHK(o, hipStreamSynchronize(o->h->stream));   // pending record copies land first
    double *nrec = nullptr;
    OHIPCHK(o, hipHostMalloc((void **)&nrec, (size_t)cap * 8 * sizeof(double), hipHostMallocDefault));
    if (o->rec) {
        std::memcpy(nrec, o->rec, (size_t)o->rec_n * 8 * sizeof(double));
        (void)hipHostFree(o->rec);
    }
    o->rec = nrec;
    o->rec_cap = cap;
    return M70H_OK;
}

// The fields the records compare, current device pointers.
void diff_sources(m70o *o, const double **cur)
{
    cur[0] = o->w->a.f.E;
    cur[1] = o->w->a.f.E_r;
    cur[2] = o->h->s.Uc;
    cur[3] = o->h->s.Vc;
    cur[4] = o->h->s.Zs;
    cur[5] = o->s ? o->s->bed_h : nullptr;
}

}  // namespace

extern "C" {

const char *m70o_last_error(const m70o *o) { return o ? o->err.c_str() : g_create_error.c_str(); }

int m70o_create(m70h *h, m70w *w, m70s *s, const m70o_params *p, double time0, int max_snapshots, m70o **out)
{
    if (!out) return ofail(nullptr, M70H_EINVAL, "out is NULL");
    *out = nullptr;
    if (!h || !w || !p) return ofail(nullptr, M70H_EINVAL, "hydro, wave and params are required");
    if (w->hydro != h) return ofail(nullptr, M70H_EINVAL, "the wave handle is not coupled to this hydro handle");
    if (s && s->h != h) return ofail(nullptr, M70H_EINVAL, "the sediment handle belongs to another hydro handle");
    if (!(p->dt > 0) || !(p->T_off > 0)) return ofail(nullptr, M70H_EINVAL, "need dt > 0 and T_off > 0");
    if (p->runIterations < 0 || p->runIterationsInBedEvolution < 0)
        return ofail(nullptr, M70H_EINVAL, "run iterations must be >= 0");
    m70o *o = new m70o();
    o->h = h;
    o->w = w;
    o->s = s;
    o->p = *p;
    o->time = time0;
    o->lastSavedTime = -1;
    if (time0 > 0) {   // resumed run (src/Morfo70.cpp:104-110)
        o->lastSavedTime = std::floor(time0);
    }
    o->H_off0 = (p->warmTime <= 0) ? p->H_off : 0.01;   // :96-98
    o->max_snap = max_snapshots > 0 ? max_snapshots : 1;
    o->slots.resize(o->max_snap);
    o->snap_doubles = 0;
    for (const SnapField &f : kSnapFields) o->snap_doubles += (size_t)(f.plus ? h->Fp : h->F);
    auto bail = [&](hipError_t e) {
        g_create_error = std::string("m70o_create: ") + hipGetErrorString(e);
        m70o_destroy(o);
        return M70H_EHIP;
    };
    hipError_t e = hipStreamCreate(&o->copy_stream);
    if (e != hipSuccess) return bail(e);
    const long F = h->F;
    for (int q = 0; q < 6; ++q) {
        void *m = nullptr;
        if ((e = hipMalloc(&m, (size_t)F * sizeof(double))) != hipSuccess) return bail(e);
        o->dev.push_back(m);
        o->prev[q] = (double *)m;
    }
    void *m = nullptr;
    if ((e = hipMalloc(&m, 6 * sizeof(unsigned long long))) != hipSuccess) return bail(e);
    o->dev.push_back(m);
    o->dout = (unsigned long long *)m;
    // the previous step's fields of the first record: the current ones (:120-127)
    int rc = wave_finish(w);
    if (rc) {
        g_create_error = w->err;
        m70o_destroy(o);
        return rc;
    }
    const double *cur[6];
    diff_sources(o, cur);
    for (int q = 0; q < 6; ++q) {
        if (cur[q]) e = hipMemcpyAsync(o->prev[q], cur[q], (size_t)F * sizeof(double), hipMemcpyDeviceToDevice, h->stream);
        else e = hipMemsetAsync(o->prev[q], 0, (size_t)F * sizeof(double), h->stream);
        if (e != hipSuccess) return bail(e);
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bail(e);
    *out = o;
    return M70H_OK;
}

void m70o_destroy(m70o *o)
{
    if (!o) return;
    if (o->h && o->h->stream) (void)hipStreamSynchronize(o->h->stream);
    if (o->copy_stream) (void)hipStreamSynchronize(o->copy_stream);
    for (auto &sl : o->slots) {
        if (sl.dev) (void)hipFree(sl.dev);
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.copied) (void)hipEventDestroy(sl.copied);
    }
    for (void *q : o->dev) (void)hipFree(q);
    if (o->rec) (void)hipHostFree(o->rec);
    if (o->copy_stream) (void)hipStreamDestroy(o->copy_stream);
    delete o;
}

int m70o_run(m70o *o, double stopTime, long *steps)
{
    if (!o) return M70H_EINVAL;
    const m70o_params &p = o->p;
    // the step count, from the same time arithmetic (time = time + dt while time < stopTime)
    long n = 0;
    for (double t = o->time; t < stopTime; t = t + p.dt) ++n;
    int rc = ensure_records(o, o->rec_n + n);
    if (rc) return rc;
    m70h *h = o->h;
    for (long it = 0; it < n; ++it) {
        const double time = o->time;
        // warm-up ramp of the offshore height (:133-139)
        double waveFactor = (1 + std::tanh((5 / p.warmTime) * (time - p.warmTime / 2))) / 2;
        if (p.warmTime <= 0) waveFactor = 1;
        const double a = o->H_off0, b = waveFactor * p.H_off;
        const double H_off_t = (a < b) ? b : a;   // std::max
        if ((rc = m70w_calculate_async(o->w, H_off_t, p.Theta_off, p.T_off))) return ofail(o, rc, o->w->err);
        const bool updateBottom = (time >= p.bedStartTime && time < p.bedStopTime) && p.calculate_bed_evolution;
        const int hydroIterations = updateBottom ? p.runIterationsInBedEvolution : p.runIterations;
        // the forcing also carries Dw to the sediment, which reads wave->Dw in the reference
        if ((rc = m70w_push_forcing(o->w))) return ofail(o, rc, o->w->err);
        if (p.calculate_hydro) {
            if ((rc = run_enqueue(h, hydroIterations, false))) return ofail(o, rc, h->err);
            if (p.wave_current_interaction && (rc = m70w_set_uv_from_hydro(o->w))) return ofail(o, rc, o->w->err);
        }
        if (p.calculate_sediment && o->s && (rc = m70s_calculate_fluxes_async(o->s))) return ofail(o, rc, o->s->err);
        double saveInterval = p.saveIntervalHydro;
        if (updateBottom) {
            if (o->s && (rc = m70s_update_bottom_async(o->s, p.dt))) return ofail(o, rc, o->s->err);
            saveInterval = p.saveIntervalMorfo;
        }
        o->time = time + p.dt;
        // largest changes since the previous step (:184-199), on the device
        OuterDiffArgs da{};
        diff_sources(o, da.cur);
        for (int q = 0; q < 6; ++q) da.prev[q] = o->prev[q];
        da.n = h->F;
        da.out = o->dout;
        OHIPCHK(o, launch_outer_diff(h->stream, da));
        double *r = o->rec + 8 * o->rec_n;
        r[0] = o->time;
        r[1] = H_off_t;
        OHIPCHK(o, hipMemcpyAsync(r + 2, o->dout, 6 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        o->rec_n++;
        // snapshots at the save times (:205-214)
        if (saveInterval > 0 && saveInterval * std::floor(o->time / saveInterval) > o->lastSavedTime) {
            if ((rc = take_snapshot(o))) return rc;
            o->lastSavedTime = saveInterval * std::floor(o->time / saveInterval);
        }
    }
    if (steps) *steps = n;
    return M70H_OK;
}

int m70o_time(m70o *o, double *time)
{
    if (!o || !time) return M70H_EINVAL;
    *time = o->time;
    return M70H_OK;
}

int m70o_records(m70o *o, long *n, double *out)
{
    if (!o || !n) return M70H_EINVAL;
    OHIPCHK(o, hipStreamSynchronize(o->h->stream));
    if (!out) {
        *n = o->rec_n;
        return M70H_OK;
    }
    const long k = *n < o->rec_n ? *n : o->rec_n;
    std::memcpy(out, o->rec, (size_t)k * 8 * sizeof(double));
    *n = k;
    return M70H_OK;
}

int m70o_snapshot_count(m70o *o, long *total)
{
    if (!o || !total) return M70H_EINVAL;
    *total = o->snap_total;
    return M70H_OK;
}

int m70o_snapshot(m70o *o, long i, double *time, const char *field, double *out)
{
    if (!o || !field) return M70H_EINVAL;
    if (i < 0 || i >= o->snap_total || i < o->snap_total - o->max_snap)
        return ofail(o, M70H_EINVAL, "snapshot index out of the kept range");
    m70o::Slot &sl = o->slots[(int)(i % o->max_snap)];
    OHIPCHK(o, hipEventSynchronize(sl.copied));
    if (time) *time = sl.time;
    size_t off = 0;
    for (const SnapField &f : kSnapFields) {
        const size_t len = (size_t)(f.plus ? o->h->Fp : o->h->F);
        if (std::strcmp(f.name, field) == 0) {
            if (out) std::memcpy(out, sl.host + off, len * sizeof(double));
            return M70H_OK;
        }
        off += len;
    }
    return ofail(o, M70H_EINVAL, std::string("unknown snapshot field ") + field);
}

}  // extern "C"

// =====================================================================================
// Files (class MorfoIO::Netcdf): reading, writing, snapshots, resume
// =====================================================================================
namespace {

thread_local std::string g_nc_error;

int ncfail(const std::string &msg)
{
    g_nc_error = msg;
    return M70H_EINVAL;
}

bool nc_var(const m70nc::File &f, const char *name, size_t n, std::vector<double> &out)
{
    std::string err;
    if (!m70nc::read_var(f, name, out, err)) {
        g_nc_error = err;
        return false;
    }
    if (out.size() != n) {
        g_nc_error = f.path + ": variable " + name + " has " + std::to_string(out.size()) + " values, expected " +
                     std::to_string(n);
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

const char *m70nc_last_error(void) { return g_nc_error.c_str(); }

int m70nc_dims(const char *path, int *nx, int *ny)
{
    if (!path || !nx || !ny) return ncfail("path, nx and ny are required");
    m70nc::File f;
    std::string err;
    if (!m70nc::read_header(path, f, err)) return ncfail(err);
    long x = -1, y = -1;
    for (auto &d : f.dims) {
        if (d.first == "x") x = (long)d.second;
        if (d.first == "y") y = (long)d.second;
    }
    if (x < 0 || y < 0) return ncfail(std::string(path) + ": dimensions x and y are required");
    *nx = (int)x;
    *ny = (int)y;
    return M70H_OK;
}

int m70nc_read(const char *path, const char *var, double *out, long n)
{
    if (!path || !var || !out) return ncfail("path, var and out are required");
    m70nc::File f;
    std::string err;
    if (!m70nc::read_header(path, f, err)) return ncfail(err);
    std::vector<double> v;
    if (!nc_var(f, var, (size_t)n, v)) return M70H_EINVAL;
    std::memcpy(out, v.data(), (size_t)n * sizeof(double));
    return M70H_OK;
}

int m70nc_read_att(const char *path, const char *att, double *value)
{
    if (!path || !att || !value) return ncfail("path, att and value are required");
    m70nc::File f;
    std::string err;
    if (!m70nc::read_header(path, f, err) || !m70nc::read_att(f, att, *value, err)) return ncfail(err);
    return M70H_OK;
}

int m70nc_write(const char *path, int nx, int ny, const double *x, const double *y, int nvars,
                const char *const *names, const double *const *data, const int *rows, int natts,
                const char *const *att_names, const double *att_values)
{
    if (!path || nx < 1 || ny < 1 || !x || !y) return ncfail("path, nx, ny, x and y are required");
    if (nvars > 0 && (!names || !data || !rows)) return ncfail("names, data and rows are required");
    if (natts > 0 && (!att_names || !att_values)) return ncfail("attribute names and values are required");
    std::vector<std::pair<std::string, uint64_t>> dims = {{"x", (uint64_t)nx}, {"y", (uint64_t)ny}};
    std::vector<m70nc::OutVar> vars = {{"x", {0}, x}, {"y", {1}, y}};
    for (int q = 0; q < nvars; ++q) {
        if (!names[q] || !data[q]) return ncfail("NULL variable name or data");
        if (rows[q] == nx) vars.push_back({names[q], {0, 1}, data[q]});
        else if (rows[q] == 1) vars.push_back({names[q], {1}, data[q]});
        else return ncfail(std::string("variable ") + names[q] + ": rows must be nx or 1");
    }
    std::vector<std::pair<std::string, double>> atts;
    for (int q = 0; q < natts; ++q) atts.emplace_back(att_names[q], att_values[q]);
    std::string err;
    if (!m70nc::write_file(path, dims, atts, vars, false, err)) return ncfail(err);
    return M70H_OK;
}

int m70o_save_snapshot(m70o *o, long i, const char *path, const double *x, const double *y, const double *Z)
{
    if (!o || !path || !x || !y || !Z) return ncfail("o, path, x, y and Z are required");
    const int nx = o->h->nx, ny = o->h->ny;
    const size_t F = (size_t)nx * ny, Fp = F + ny;
    auto get = [&](const char *name, size_t n, std::vector<double> &v) {
        v.resize(n);
        return m70o_snapshot(o, i, nullptr, name, v.data());
    };
    double t = 0.0;
    int rc = m70o_snapshot(o, i, &t, "Zs", nullptr);
    if (rc) return ncfail(o->err);
    const char *fnames[] = {"h", "D", "H", "E", "Cg", "lK", "dQx_dx", "dQy_dy", "Theta", "Dw", "E_r", "U", "V",
                            "Zs", "qx", "qy", "Gamma_Swash", "Alpha", "gamma"};
    std::vector<std::vector<double>> fv(sizeof(fnames) / sizeof(fnames[0]));
    for (size_t q = 0; q < fv.size(); ++q)
        if ((rc = get(fnames[q], F, fv[q]))) return ncfail(o->err);
    std::vector<double> Uf, Vf, Qxf, Qyf;
    if ((rc = get("U_full", Fp, Uf)) || (rc = get("V_full", Fp, Vf)) || (rc = get("Qx_full", Fp, Qxf)) ||
        (rc = get("Qy_full", Fp, Qyf)))
        return ncfail(o->err);
    auto f = [&](const char *n) -> const double * {
        for (size_t q = 0; q < fv.size(); ++q)
            if (std::strcmp(fnames[q], n) == 0) return fv[q].data();
        return nullptr;
    };
    // Morfo70::saveTimeStep's variables, in its order
    std::vector<m70nc::OutVar> vars = {{"x", {0}, x}, {"y", {1}, y}};
    auto v2 = [&](const char *n, const double *p) { vars.push_back({n, {0, 1}, p}); };
    auto v1 = [&](const char *n, const double *p) { vars.push_back({n, {1}, p}); };
    v2("Z", Z);
    v2("h", f("h"));
    v2("D", f("D"));
    v2("H", f("H"));
    v2("E", f("E"));
    v2("Cg", f("Cg"));
    v2("lK", f("lK"));
    v2("dQx_dx", f("dQx_dx"));
    v2("dQy_dy", f("dQy_dy"));
    v2("Theta", f("Theta"));
    v2("Dw", f("Dw"));
    v2("E_r", f("E_r"));
    v2("U", f("U"));
    v1("U0", Uf.data());
    v1("Qx0", Qxf.data());
    v2("U_x", Uf.data() + ny);
    v2("Qx_x", Qxf.data() + ny);
    v2("V", f("V"));
    v1("V0", Vf.data());
    v1("Qy0", Qyf.data());
    v2("V_y", Vf.data() + ny);
    v2("Qy_y", Qyf.data() + ny);
    v2("Zs", f("Zs"));
    v2("qx", f("qx"));
    v2("qy", f("qy"));
    v2("Gamma_Swash", f("Gamma_Swash"));
    v2("Alpha", f("Alpha"));
    v2("gamma", f("gamma"));
    std::string err;
    if (!m70nc::write_file(path, {{"x", (uint64_t)nx}, {"y", (uint64_t)ny}}, {{"time", t}}, vars, false, err))
        return ncfail(err);
    return M70H_OK;
}

int m70_resume(const char *path, m70h *h, m70w *w, m70s *s, int wave_current_interaction, double *time)
{
    if (!path || !h) return ncfail("path and hydro are required");
    m70nc::File f;
    std::string err;
    if (!m70nc::read_header(path, f, err)) return ncfail(err);
    const int nx = h->nx, ny = h->ny;
    const size_t F = (size_t)nx * ny, Fp = F + ny;
    std::vector<double> D, hb, Uc, Vc, Zs, dqx, dqy, U0, Qx0, V0, Qy0, Ux, Qxx, Vy, Qyy;
    if (!nc_var(f, "D", F, D) || !nc_var(f, "U", F, Uc) || !nc_var(f, "V", F, Vc) || !nc_var(f, "Zs", F, Zs) ||
        !nc_var(f, "dQx_dx", F, dqx) || !nc_var(f, "dQy_dy", F, dqy) || !nc_var(f, "U0", ny, U0) ||
        !nc_var(f, "Qx0", ny, Qx0) || !nc_var(f, "V0", ny, V0) || !nc_var(f, "Qy0", ny, Qy0) ||
        !nc_var(f, "U_x", F, Ux) || !nc_var(f, "Qx_x", F, Qxx) || !nc_var(f, "V_y", F, Vy) ||
        !nc_var(f, "Qy_y", F, Qyy))
        return M70H_EINVAL;
    // Hydro::setHydroFromResumeFile (src/Hydro.cpp:204-227): row 0 and rows 1..nx apart
    auto join = [&](const std::vector<double> &r0, const std::vector<double> &rest) {
        std::vector<double> v(Fp);
        std::memcpy(v.data(), r0.data(), ny * sizeof(double));
        std::memcpy(v.data() + ny, rest.data(), F * sizeof(double));
        return v;
    };
    std::vector<double> U = join(U0, Ux), Qx = join(Qx0, Qxx), V = join(V0, Vy), Qy = join(Qy0, Qyy);
    int rc = m70h_set_state(h, Zs.data(), D.data(), Qx.data(), Qy.data(), U.data(), V.data(), Uc.data(), Vc.data(),
                            dqx.data(), dqy.data());
    if (rc) return ncfail(h->err);
    if (s) {   // Bathymetry::setBathyFromResumeFile (h; Zb stays) and the sediment fluxes
        std::vector<double> qx, qy;
        if (!nc_var(f, "h", F, hb) || !nc_var(f, "qx", F, qx) || !nc_var(f, "qy", F, qy)) return M70H_EINVAL;
        if ((rc = m70s_set_bed(s, hb.data(), nullptr)) || (rc = m70s_set_fluxes(s, qx.data(), qy.data())))
            return ncfail(s->err);
    }
    if (w) {   // Wave::setWaveEnergyFromResumeFile, ::setUVFromResumeFile (src/Wave.cpp:305-336)
        std::vector<double> E, H, Th;
        if (!nc_var(f, "E", F, E) || !nc_var(f, "H", F, H) || !nc_var(f, "Theta", F, Th)) return M70H_EINVAL;
        if ((rc = wave_finish(w))) return ncfail(w->err);
        const size_t nb = F * sizeof(double);
        hipError_t e = hipMemcpyAsync(w->a.f.E, E.data(), nb, hipMemcpyHostToDevice, w->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(w->a.f.H, H.data(), nb, hipMemcpyHostToDevice, w->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(w->a.f.Theta, Th.data(), nb, hipMemcpyHostToDevice, w->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
        if (e != hipSuccess) return ncfail(std::string("m70_resume: ") + hipGetErrorString(e));
        if (wave_current_interaction) {
            std::vector<double> Uw(F);
            for (size_t q = 0; q < F; ++q) Uw[q] = std::cos(Th[q]) * Uc[q] + std::sin(Th[q]) * Vc[q];
            e = hipMemcpyAsync(w->U, Uc.data(), nb, hipMemcpyHostToDevice, w->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(w->V, Vc.data(), nb, hipMemcpyHostToDevice, w->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(w->U_w, Uw.data(), nb, hipMemcpyHostToDevice, w->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
            if (e != hipSuccess) return ncfail(std::string("m70_resume: ") + hipGetErrorString(e));
        }
    }
    double t = 0.0;
    if (!m70nc::read_att(f, "time", t, err)) return ncfail(err);
    if (time) *time = t;
    return M70H_OK;
}

}  // extern "C"
