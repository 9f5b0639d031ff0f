/*
 * m70h.h -- C ABI of the MI355X-native Morfo70 Hydro ADI step (libm70h.so).
 *
 * The reference exposes the step as a plain C++ class with no FFI: `class Hydro`
 * (include/Hydro.h:24-478) with ctor Hydro(Grid*,Bathymetry*,Wave*) (:103), ~Hydro()
 * (:109), setHydroFromResumeFile(string) (:117), Run(int nIter) (:124) and public state
 * arrays U, U_c, V, V_c, Zs, Qx, Qy, dQx_dx, dQy_dy (:26-34). Each entry point below
 * replaces one of those members; the C++ facade (morfo70_amd/host/Hydro.h) rebuilds the
 * reference class on top of them.
 *
 * Conventions:
 *   - Arrays are row-major float64, index i*ny + j; i cross-shore (0 = offshore), j
 *     alongshore (periodic). F = nx*ny; "F+" fields (Qx, Qy, U, V) have nx+1 rows: row 0
 *     is the offshore ghost, row nx the onshore wall.
 *   - Every call returns 0 (M70H_OK) or an M70H_E* code; m70h_last_error() has the text.
 *   - A handle owns its device buffers and stream; it is not re-entrant. All device work
 *     is enqueued on the handle's stream (m70h_set_stream), default: its own stream.
 *   - Pointers named *_dev are device pointers on the handle's device; all others are
 *     host pointers. Any state pointer may be NULL to leave that field unchanged.
 */
#ifndef M70H_H_
#define M70H_H_

#ifdef __cplusplus
extern "C" {
#endif

#define M70H_OK 0
#define M70H_EINVAL 1     /* bad argument (dims < 3, NULL handle, bad enum, ...) */
#define M70H_EHIP 2       /* HIP runtime error; text in m70h_last_error */
#define M70H_ENONFINITE 3 /* NaN/Inf in the state after a Run, or found by m70h_check_finite */
#define M70H_ESTATE 4     /* call out of order (e.g. run before forcing is set) */

/* Options::Hydro enums, include/OptionsParametersConstants.h:50-64 */
#define M70H_DRAG_CONSTANT 0
#define M70H_DRAG_LOGDEPTH 1
#define M70H_DRAG_MANNING 2
#define M70H_VISC_CONSTANT 0
#define M70H_VISC_ROLLERS 1
#define M70H_VISC_DEPTH 2
#define M70H_VISC_HRMS 3
#define M70H_BC_SPONGE 0
#define M70H_BC_CONSTANT_ZS 1

/* Line-solver families (m70h_set_solver). EXACT runs the reference's own Thomas and
 * El-Mikkawy recurrences (src/Math.cpp:15-118) one line per lane, in the reference's
 * operation order: results are bitwise equal to the reference's serial CPU step. */
#define M70H_SOLVER_EXACT 0
/* PARTITION splits every line over the threads of a workgroup: two-sided elimination of
 * per-thread segments in registers, cyclic reduction of the segment-boundary system in
 * LDS, back-substitution. Same systems, different rounding: results agree with EXACT to
 * a few ulps per solve. Line lengths up to 8192. */
#define M70H_SOLVER_PARTITION 1
/* FUSED keeps the coefficients out of HBM where it can: the y-sweeps (lines along rows)
 * are assembled and solved in one launch with the row staged through LDS, and the
 * x-sweeps assemble and eliminate 8-row column segments in registers, solve the reduced
 * boundary systems and back-substitute, every access coalesced across 64 columns.
 * Same partition algorithm as PARTITION, different segment sizes on the columns. nx-1 and
 * ny up to 8192. */
#define M70H_SOLVER_FUSED 2

/* Parameters read by Hydro at call time in the reference (Parameters::Hydro::dt, Dmin,
 * alpha_f, cD0, nu0, M, ka; Parameters::Bottom::z0; Parameters::Rollers::alpha_r;
 * Options::Hydro::DragModel/ViscosityModel/OffshoreBC), passed explicitly here. */
typedef struct m70h_params {
    int drag_model;
    int visc_model;
    int offshore_bc;
    double dt;
    double Dmin;
    double alpha_f;
    double cD0;
    double nu0;
    double M;
    double ka;
    double z0;
    double alpha_r;
} m70h_params;

typedef struct m70h m70h;

/* Replaces Hydro::Hydro (include/Hydro.h:103; src/Hydro.cpp:17-133) together with the
 * Grid metric set-up it depends on (src/Grid.cpp:32-92): x[nx] are the cross-shore node
 * coordinates, dy the alongshore spacing. Allocates all device buffers and sets the
 * offshore boundary constants (src/Hydro.cpp:44-60). State starts at zero. */
int m70h_create(int nx, int ny, const double *x, double dy, const m70h_params *p, m70h **out);

/* Replaces Hydro::~Hydro (include/Hydro.h:109). */
void m70h_destroy(m70h *h);

/* Text of the last error on this handle (or of the last failed m70h_create if h is NULL). */
const char *m70h_last_error(const m70h *h);

/* Replaces the reference's call-time reads of Parameters::Hydro::* and Options::Hydro::*
 * (e.g. src/HydroQx.cpp:102): new parameters for the following runs. The offshore
 * boundary constants keep the values fixed at construction (src/Hydro.cpp:44-60). */
int m70h_set_params(m70h *h, const m70h_params *p);

/* Use `hip_stream` (a hipStream_t) for all device work of this handle; NULL restores
 * the handle's own stream. */
int m70h_set_stream(m70h *h, void *hip_stream);

/* Initial state of the reference ctor: D = bathy->D, Zs = D - Z + h (src/Hydro.cpp:35-37),
 * all other state fields zero. D, Z, hb are host arrays of F doubles. */
int m70h_init_from_bathymetry(m70h *h, const double *D, const double *Z, const double *hb);

/* Full state exchange; replaces Hydro::setHydroFromResumeFile (include/Hydro.h:117) and
 * the public arrays (include/Hydro.h:26-34) plus Bathymetry::D, which Hydro mutates
 * (src/Hydro.cpp:267). Zs, D, U_c, V_c, dQx_dx, dQy_dy: F; Qx, Qy, U, V: F+. */
int m70h_set_state(m70h *h, const double *Zs, const double *D, const double *Qx, const double *Qy,
                   const double *U, const double *V, const double *U_c, const double *V_c,
                   const double *dQx_dx, const double *dQy_dy);
int m70h_get_state(m70h *h, double *Zs, double *D, double *Qx, double *Qy, double *U, double *V,
                   double *U_c, double *V_c, double *dQx_dx, double *dQy_dy);
int m70h_set_state_dev(m70h *h, const double *Zs_dev, const double *D_dev, const double *Qx_dev,
                       const double *Qy_dev, const double *U_dev, const double *V_dev,
                       const double *U_c_dev, const double *V_c_dev, const double *dQx_dx_dev,
                       const double *dQy_dy_dev);
int m70h_get_state_dev(m70h *h, double *Zs_dev, double *D_dev, double *Qx_dev, double *Qy_dev,
                       double *U_dev, double *V_dev, double *U_c_dev, double *V_c_dev,
                       double *dQx_dx_dev, double *dQy_dy_dev);

/* Wave forcing read by Run's pre-loop: Wave::Sxx, Sxy, Syy, Urms, Dw, D_r, H
 * (include/Wave.h:36-53; read at src/Hydro.cpp:175-181). F doubles each. */
int m70h_set_forcing(m70h *h, const double *Sxx, const double *Sxy, const double *Syy,
                     const double *Urms, const double *Dw, const double *D_r, const double *H);
int m70h_set_forcing_dev(m70h *h, const double *Sxx_dev, const double *Sxy_dev,
                         const double *Syy_dev, const double *Urms_dev, const double *Dw_dev,
                         const double *D_r_dev, const double *H_dev);

/* Offshore flux hook Qx0 = f1*Qx1 + f0 (include/Hydro.h:89-90, used at
 * src/HydroBoundaryConditions.cpp:13-39). The reference fixes f0 = 0
 * (src/Hydro.cpp:52,56); a series f0[k] gives a time-varying boundary (tide): the k-th
 * step run after the call uses f0[k] (the last value repeats past the end). */
int m70h_set_offshore_f0(m70h *h, double f0);
int m70h_set_offshore_f0_series(m70h *h, const double *f0, long n);
int m70h_bc_constants(const m70h *h, double *f1, double *f0, double *bcQy);

/* Replaces Hydro::Run (include/Hydro.h:124; src/Hydro.cpp:170-202): per-call drag,
 * viscosity and forcing fields from the current D and forcing, then nIter ADI steps.
 * m70h_run returns when the steps are done; m70h_run_async only enqueues them on the
 * handle's stream (m70h_synchronize waits and reports errors). After the steps, the state
 * is scanned for NaN/Inf on the device (about 0.1 ms at 2048 x 4096): m70h_run, or the
 * m70h_synchronize after m70h_run_async, returns M70H_ENONFINITE when it holds any, with
 * the state left as the steps produced it. The reference lets such a state pass silently
 * (e.g. Urms = 0 in the drag term, src/HydroParametrizations.cpp:25). */
int m70h_run(m70h *h, int nIter);
int m70h_run_async(m70h *h, int nIter);
int m70h_synchronize(m70h *h);

/* Line-solver family, M70H_SOLVER_* (default EXACT). */
int m70h_set_solver(m70h *h, int solver);

/* PARTITION solver tiles: unknowns per thread (8 or 16; 0 = default) for the lines along
 * x (columns) and along y (rows). This is the knob the autotuner sets. */
int m70h_set_solver_tiles(m70h *h, int tile_cols, int tile_rows);

/* FUSED row-solver block: 256 threads (ny up to 2048; four blocks per CU) or 512 (ny up
 * to 4096; two blocks per CU); 0 picks the smallest that fits. Rows longer than 4096 (up
 * to 8192) always take 1024-thread blocks. An autotuner knob. */
int m70h_set_rows_block(m70h *h, int threads);

/* Batched tridiagonal solve on device arrays, the counterpart of Math::TridiagonalSolver
 * (src/Math.cpp:15-65, lines strided as in Hydro's x-sweeps) and
 * Math::TridiagonalSolverPeriodic (src/Math.cpp:68-186, contiguous lines as in Hydro's
 * y-sweeps). Line L, element k lives at L*ld + k when lines_are_rows, else at k*ld + L.
 * a = diagonal, b = lower (b of element 0 is the periodic corner), c = upper (c of element
 * n-1 is the corner), f = right-hand side; y receives the solution with stride ldy.
 * EXACT supports rows+periodic (scratch >= 3*n*roundup(nlines,64) doubles) and
 * columns+non-periodic (ldy == ld, scratch >= n*ld doubles); PARTITION supports rows
 * (periodic or not) and columns (non-periodic), n <= 8192, no scratch. Enqueued on
 * hip_stream (a hipStream_t); returns without waiting. */
int m70h_tridiag_solve(void *hip_stream, int solver, int lines_are_rows, int periodic, int n, int nlines, long ld,
                       const double *a_dev, const double *b_dev, const double *c_dev, const double *f_dev,
                       double *y_dev, long ldy, double *scratch_dev, long scratch_len, int tile);
/* m70h_tridiag_solve without the workspace argument, and without any device allocation:
 * no counterpart in the reference. It is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The EXACT family's scratch lives in a static device array of the library,
 * M70H_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) on every GPU the library runs on, held
 * until the process ends; its address is looked up once per device on that device's
 * first call. Sizes whose scratch would exceed it (EXACT rows+periodic:
 * 3*n*roundup(nlines,64); EXACT columns: n*ld) return M70H_EINVAL with a message in
 * m70h_last_error(NULL). All calls on one GPU share that GPU's array: calls that overlap
 * on one GPU (different streams or threads without ordering between them) race. The
 * array keeps the last call's contents, and no result depends on them. PARTITION needs no
 * scratch. Same results as m70h_tridiag_solve. */
#define M70H_NOALLOC_SCRATCH_DOUBLES 4194304L
int m70h_tridiag_solve_noalloc(void *hip_stream, int solver, int lines_are_rows, int periodic, int n, int nlines,
                               long ld, const double *a_dev, const double *b_dev, const double *c_dev,
                               const double *f_dev, double *y_dev, long ldy, int tile);


/* Scans the state for NaN/Inf on the device; *finite = 1 if none. Returns
 * M70H_ENONFINITE when a non-finite value is found. */
int m70h_check_finite(m70h *h, int *finite);

/* Per-kernel timing with HIP events on the handle's stream (no counterpart in the
 * reference; measurement only). When enabled, run records an event pair around every
 * launch; m70h_profile reports kernel k's name, summed milliseconds and launch count
 * (returns M70H_EINVAL past the last kernel). */
int m70h_set_profiling(m70h *h, int enable);
int m70h_profile(m70h *h, int k, const char **name, double *ms, long *launches);
/* Profiling of the launches of one kernel only (its name as m70h_profile reports it;
 * NULL for every launch): an event pair around each launch of that kernel alone, so
 * timing one kernel inside a timed region disturbs the step less. M70H_EINVAL for an
 * unknown name. */
int m70h_set_profiling_only(m70h *h, const char *kernel);

/* Device pointer of the current state field `name` ("Zs","D","Qx","Qy","U","V","U_c",
 * "V_c","dQx_dx","dQy_dy"), valid until the next run. */
int m70h_state_ptr(m70h *h, const char *name, double **dev);

/* ---------------------------------------------------------------------------------
 * Batches. The reference advances one beach per process (src/main.cpp:17-29); a batch
 * advances many independent handles (uneven grids, each with its own state, forcing and
 * parameters) together: every stage of the step is one launch covering all of them, and
 * with the graph on (default) the two-step launch sequence is captured once in a hipGraph
 * and replayed. Results are bitwise those of m70h_run with M70H_SOLVER_FUSED on each
 * handle, a time series of the offshore offset (m70h_set_offshore_f0_series) included: each
 * scenario's step reads its own series value on the device. The handles keep owning their
 * buffers; get/set their state as usual between batch runs. Requirements: ny and nx-1 up
 * to 4096.
 * --------------------------------------------------------------------------------- */
typedef struct m70h_batch m70h_batch;
int m70h_batch_create(int nscen, m70h *const *handles, m70h_batch **out);
void m70h_batch_destroy(m70h_batch *b);
const char *m70h_batch_last_error(const m70h_batch *b);
int m70h_batch_set_graph(m70h_batch *b, int use_graph);
/* Row-solver block size (256 or 512) for the whole batch; default the smallest that fits
 * the largest ny. An autotuner knob. */
int m70h_batch_set_rows_block(m70h_batch *b, int threads);
/* The batch's hipStream_t (for events around m70h_batch_run_async). */
int m70h_batch_stream(m70h_batch *b, void **hip_stream);
/* Run(nIter) of every handle: per-Run pre-loop, then nIter steps. */
int m70h_batch_run(m70h_batch *b, int nIter);
int m70h_batch_run_async(m70h_batch *b, int nIter);
int m70h_batch_synchronize(m70h_batch *b);

/* ------------------------------------------------------------------------------------
 * Sediment transport and bed update: class Sediment (include/Sediment.h:17-115,
 * src/Sediment.cpp). A sediment handle is attached to a hydro handle and works on its
 * device state: it reads Zs, D, U, V and the wave dissipation Dw of the hydro's forcing,
 * owns the bed fields h and Zb (Bathymetry::h, ::Zb, which the reference's Sediment
 * mutates) and the sediment fields, and its bed update moves the hydro's Zs with the bed
 * (src/Sediment.cpp:131), as the reference's does. Destroy it before its hydro handle.
 * ------------------------------------------------------------------------------------ */
/* Parameters::Bottom::p, Parameters::Sediment::CWS::* and ::Swash::*
 * (include/OptionsParametersConstants.h:84-138). */
typedef struct m70s_params {
    double p;
    double alpha0_max, alpha0_min, D_alpha0_max, D_alpha0_min, gamma0;
    double Gamma0, A_swash;
} m70s_params;

typedef struct m70s m70s;

/* Replaces Sediment::Sediment (include/Sediment.h:66; src/Sediment.cpp:13-31): zeroed
 * sediment fields; h and Zb (host, F doubles) are the bathymetry's bed fields. */
int m70s_create(m70h *hydro, const double *h, const double *Zb, const m70s_params *p, m70s **out);
/* Replaces Sediment::~Sediment (include/Sediment.h:71). */
void m70s_destroy(m70s *s);
const char *m70s_last_error(const m70s *s);
/* Replaces Sediment::calculateSedimentFluxes (include/Sediment.h:76; src/Sediment.cpp:47-65). */
int m70s_calculate_fluxes(m70s *s);
/* Replaces Sediment::updateBottom (include/Sediment.h:81; src/Sediment.cpp:67-75,105-151);
 * dt is Parameters::Run::dt, read at call time in the reference. */
int m70s_update_bottom(m70s *s, double dt);
/* The two calls above without the wait: enqueued on the hydro handle's stream after its
 * pending work (m70h_run_async), completed by m70h_synchronize or any m70s call that
 * copies to or from the host. */
int m70s_calculate_fluxes_async(m70s *s);
int m70s_update_bottom_async(m70s *s, double dt);
/* Replaces Sediment::setSedimentFluxesFromResumeFile's arithmetic (include/Sediment.h:87;
 * src/Sediment.cpp:77-103): sets qx, qy (host, F) and the dh_dt they imply. */
int m70s_set_fluxes(m70s *s, const double *qx, const double *qy);
/* The public arrays Alpha, gamma, Gamma, qx, qy, dh_dt (include/Sediment.h:21-26) and the
 * bed fields h, Zb; host pointers of F doubles, NULL = skip. */
int m70s_get(m70s *s, double *Alpha, double *gamma, double *Gamma, double *qx, double *qy, double *dh_dt, double *h,
             double *Zb);
/* Callers' writes to Bathymetry::h / ::Zb between calls (host, F; NULL = unchanged). */
int m70s_set_bed(m70s *s, const double *h, const double *Zb);

/* ------------------------------------------------------------------------------------
 * Wave and rollers driver: class Wave (include/Wave.h:16-128, src/Wave.cpp) with
 * class WaveRow (include/WaveRow.h, src/WaveRow.cpp, src/WaveDir.cpp, src/WaveEnergy.cpp,
 * src/WaveParametrizations.cpp, src/WaveRollers.cpp). It produces the hydro forcing
 * (Sxx, Sxy, Syy, Urms, Dw, D_r, H) every outer step (src/Morfo70.cpp:141).
 * Semantics differ from the reference's code in two places where that code cannot be
 * what was meant: the two row buffers swap after every marching step (the brace-less
 * `else` at src/Wave.cpp:158-163 aliases them after a mesh row), and bottom friction is
 * evaluated per node being solved (src/WaveParametrizations.cpp:357-378 index the first
 * nNodesToSolve nodes). With one OpenMP thread and friction off (the defaults) the
 * reference computes the same thing.
 * ------------------------------------------------------------------------------------ */
enum { M70W_BREAK_CT = 1, M70W_BREAK_TG = 2, M70W_BREAK_REGULAR = 3 };   /* include/OptionsParametersConstants.h:31-36 */
enum { M70W_FRIC_CONSTANT = 0, M70W_FRIC_SOULSBY97 = 1 };                /* :37-40 */

/* Options::Wave::*, Options::Rollers::CalculateRollers, Parameters::Wave::*,
 * Parameters::Rollers::*, Parameters::Bottom::z0 (include/OptionsParametersConstants.h).
 * theta_max in radians. BreakingModel None is rejected (the reference calls through a
 * NULL member pointer for it, src/WaveRow.cpp:68-80). */
typedef struct m70w_params {
    int breaking_model, friction_model, bottom_friction, rollers, max_iterations;
    double tolerance, B3, gamma_b, m_mon, n_mon, theta_max, fw0, z0, sinBeta_r, alpha_r;
} m70w_params;

typedef struct m70w m70w;

/* Replaces Wave::Wave (src/Wave.cpp:13-56): grid x[nx] (offshore first), ny, dy. */
int m70w_create(int nx, int ny, const double *x, double dy, const m70w_params *p, m70w **out);
void m70w_destroy(m70w *w);
const char *m70w_last_error(const m70w *w);
int m70w_set_stream(m70w *w, void *hip_stream);
/* bathy->D, the depth the field is computed on (host, nx*ny). */
int m70w_set_depth(m70w *w, const double *D);
/* Replaces Wave::setUV (src/Wave.cpp:293-302): currents (host, nx*ny) and their
 * projection on the stored wave direction. */
int m70w_set_uv(m70w *w, const double *U, const double *V);
/* Replaces Wave::calculateWaveField (src/Wave.cpp:80-172). */
int m70w_calculate(m70w *w, double hOff, double thetaOff, double T);
int m70w_calculate_async(m70w *w, double hOff, double thetaOff, double T);
/* A public array of Wave (include/Wave.h:33-53) by name: K (log of the wave number),
 * Theta, Cg, H, E, Urms, Sxx_w, Sxy_w, Syy_w, Dw, E_r, D_r, Sxx_r, Sxy_r, Syy_r, Sxx, Sxy,
 * Syy, CosTheta, SinTheta, U_w. out: host, nx*ny doubles. */
int m70w_get(m70w *w, const char *name, double *out);
/* Marching steps (mesh rows plus intermediate rows) of the last field. */
int m70w_steps(m70w *w, long *steps);
/* Device coupling with a hydro handle of the same grid (no counterpart in the reference,
 * whose objects share host arrays): the field is computed on the hydro's depth D (which
 * Hydro::Run updates, src/Hydro.cpp:267) and on the hydro's stream. */
int m70w_couple_hydro(m70w *w, m70h *h);
/* Wave::setUV(hydro->U_c, hydro->V_c) (src/Morfo70.cpp:151-154) on the device. */
int m70w_set_uv_from_hydro(m70w *w);
/* The forcing Hydro::Run reads (Sxx, Sxy, Syy, Urms, Dw, D_r, H) copied device to device
 * into the coupled hydro handle, in place of m70h_set_forcing. */
int m70w_push_forcing(m70w *w);

/* ------------------------------------------------------------------------------------
 * Morphodynamic outer loop on the device: Morfo70::Run (src/Morfo70.cpp:131-218).
 * Each outer step computes the wave field on the hydro's depth, hands it to the hydro as
 * forcing, runs the hydro, optionally projects its currents on the wave direction, computes
 * the sediment fluxes and, inside the bed evolution window, updates the bed; then the
 * largest changes of E, E_r, U_c, V_c, Zs and h since the previous step (the line the
 * reference prints, :184-199) are reduced on the device into a per-step record, and a
 * snapshot of the saved fields (saveTimeStep, :225-276) is taken at the save times by a
 * device copy followed by an asynchronous copy to pinned host memory. The fields stay
 * resident; the host only evaluates the scalar schedule (time, wave ramp, bed window,
 * save times) in the reference's arithmetic. The wave handle must be coupled to the
 * hydro handle (m70w_couple_hydro); the sediment handle (NULL = none) must be attached to
 * it. Destroy the outer-loop handle before the others.
 * ------------------------------------------------------------------------------------ */
/* Morfo70::H_off, T_off, Theta_off (radians), Parameters::Run::warmTime, bedStartTime,
 * bedStopTime, dt, saveIntervalHydro, saveIntervalMorfo, Parameters::Hydro::runIterations,
 * runIterationsInBedEvolution, Options::Hydro::CalculateHydro,
 * Options::Wave::WaveCurrentInteraction, Options::Sediment::CalculateSedimentTransport,
 * Options::Bottom::CalulateBedEvolution (include/OptionsParametersConstants.h). The
 * minimum offshore height is 0.01 m, or H_off when warmTime <= 0 (src/Morfo70.cpp:96-98). */
typedef struct m70o_params {
    double H_off, T_off, Theta_off;
    double warmTime, bedStartTime, bedStopTime, dt, saveIntervalHydro, saveIntervalMorfo;
    int runIterations, runIterationsInBedEvolution;
    int calculate_hydro, wave_current_interaction, calculate_sediment, calculate_bed_evolution;
} m70o_params;

typedef struct m70o m70o;

/* time0: the start time (0, or a resume file's time attribute, :104-110). The previous
 * step's fields for the first record are the handles' current E, E_r, U_c, V_c, Zs, h
 * (:120-127). max_snapshots: device and pinned host slots kept (the oldest is reused). */
int m70o_create(m70h *h, m70w *w, m70s *s, const m70o_params *p, double time0, int max_snapshots, m70o **out);
void m70o_destroy(m70o *o);
const char *m70o_last_error(const m70o *o);
/* Outer steps while time < stopTime (:131-218). *steps (may be NULL): steps run. */
int m70o_run(m70o *o, double stopTime, long *steps);
int m70o_time(m70o *o, double *time);
/* Per-step records, 8 doubles each: time, offshore height, dE, dR, dU, dV, dZs, dh
 * (:194). out = NULL returns the count in *n; otherwise *n records are copied. */
int m70o_records(m70o *o, long *n, double *out);
/* Snapshots taken so far (save times, :205-214), and a field of snapshot i by the name it
 * has in the reference's m70_*.nc files: Zs, D, U (= U_c), V (= V_c), U_full, V_full,
 * Qx_full, Qy_full ((nx+1)*ny, row 0 the offshore ghost), dQx_dx, dQy_dy, H, E, Cg, lK,
 * Theta, Dw, E_r, qx, qy, Gamma_Swash, Alpha, gamma, h. Waits for that snapshot's copy. */
int m70o_snapshot_count(m70o *o, long *total);
int m70o_snapshot(m70o *o, long i, double *time, const char *field, double *out);

/* ------------------------------------------------------------------------------------
 * Files: NetCDF classic I/O in place of class MorfoIO::Netcdf (include/Netcdf.h,
 * src/Netcdf.cpp), which wraps netcdf-c. Files hold dimensions x (nx) and y (ny), a
 * global attribute `time`, and double variables over (x, y) or (y). The reader also takes
 * the other numeric classic types; record (unlimited) variables are not supported.
 * ------------------------------------------------------------------------------------ */
/* Netcdf::ReadDims (src/Netcdf.cpp:39-50). */
int m70nc_dims(const char *path, int *nx, int *ny);
/* Netcdf::ReadVar / ReadVarY / ReadXY (:52-97): variable `var` as doubles; n must be its
 * element count (nx*ny, ny or nx). */
int m70nc_read(const char *path, const char *var, double *out, long n);
/* Netcdf::ReadAtt (:79-82): a global attribute's first value. */
int m70nc_read_att(const char *path, const char *att, double *value);
/* Netcdf::WriteXY / WriteVar / WriteVarY / WriteAtt (:99-140) in one call: dimensions and
 * variables x, y, then nvars variables over (x, y) when rows[q] == nx or (y) when
 * rows[q] == 1, then the global attributes. The file is CDF-1, or CDF-2 (64-bit offsets)
 * when it passes 2 GiB. */
int m70nc_write(const char *path, int nx, int ny, const double *x, const double *y, int nvars,
                const char *const *names, const double *const *data, const int *rows, int natts,
                const char *const *att_names, const double *att_values);
/* Message of the last failed m70nc_* or m70o_save_snapshot / m70_resume call on this thread. */
const char *m70nc_last_error(void);
/* Morfo70::saveTimeStep (src/Morfo70.cpp:225-276) of snapshot i of an outer loop: the same
 * variables in the same order; x, y, Z (host) are the grid's and the bathymetry's. */
int m70o_save_snapshot(m70o *o, long i, const char *path, const double *x, const double *y, const double *Z);
/* Morfo70::resume's reads (src/Morfo70.cpp:298-312) from one m70_*.nc file into the
 * handles: bathymetry D and h (Bathymetry::setBathyFromResumeFile; Zb is not re-read,
 * as there), wave E, H, Theta and, with the wave-current interaction, U, V and U_w
 * (Wave::setWaveEnergyFromResumeFile, ::setUVFromResumeFile), the hydro state
 * (Hydro::setHydroFromResumeFile), the sediment fluxes and their dh_dt
 * (Sediment::setSedimentFluxesFromResumeFile); *time = the time attribute. w and s may
 * be NULL. */
int m70_resume(const char *path, m70h *h, m70w *w, m70s *s, int wave_current_interaction, double *time);

/* ------------------------------------------------------------------------------------
 * Adjoint of Run. No counterpart in the reference, which has no derivatives: these
 * entry points serve callers that fit bathymetry or friction to observations (gradient
 * of a loss of the final state with respect to the initial state and the drag and
 * viscosity parameters).
 *
 * m70h_adjoint_forward runs Run(nIter) like m70h_run (same pre-loop, same steps, same
 * final state) and keeps `slots` state checkpoints (>= 1; at most nIter are used), placed
 * by the optimal binomial (Revolve) schedule. slots = 0 instead tapes every step during
 * the forward sweep (about 30*nx*ny doubles per step of device memory), so that the
 * reverse sweep recomputes nothing. m70h_adjoint_reverse then takes the
 * adjoint of the final state (seed, device pointers, NULL fields = zero) and returns the
 * adjoint of the state before the Run (grad, device pointers, NULL = not returned) and,
 * in dparams[0..2] (host), the adjoints of the parameters cD0, nu0 and M. The depth
 * adjoint includes the per-Run drag and viscosity fields' dependence on that depth.
 * Steps are recomputed from checkpoints with the intermediates taped; wet/dry masks are
 * held fixed (their derivative is zero almost everywhere). Afterwards the handle's state
 * is the final state again. Any Run or forcing change in between makes the reverse fail
 * with M70H_ESTATE. m70h_adjoint_release frees the checkpoints and work buffers.
 * ------------------------------------------------------------------------------------ */
typedef struct m70h_fields {
    double *Zs, *D, *Qx, *Qy, *U, *V, *U_c, *V_c, *dQx_dx, *dQy_dy;
} m70h_fields;

int m70h_adjoint_forward(m70h *h, int nIter, int slots);
int m70h_adjoint_reverse(m70h *h, const m70h_fields *seed_dev, const m70h_fields *grad_dev, double *dparams);
/* Untaped and taped step evaluations of the last reverse sweep, and buffers in use. */
int m70h_adjoint_stats(const m70h *h, long *forward_steps, long *taped_steps, int *buffers);
int m70h_adjoint_release(m70h *h);

/* The adjoint of Run for many handles at once (no counterpart in the reference): every
 * operation of the checkpointed forward and reverse sweeps -- state copies, untaped steps,
 * taped steps and reverse steps -- is one launch per stage over all handles, as the
 * forward batch (m70h_batch_*) advances them. The handles share the Run length and the
 * checkpoint budget (so one binomial schedule), and each keeps its own checkpoints and
 * adjoint buffers, as with m70h_adjoint_forward. Gradients equal each handle's own
 * m70h_adjoint_reverse to rounding (the reverse adds through atomics). Requirements: the
 * batch requirements, nx-1 up to 2048, default solver tiles, no offshore offset series,
 * slots >= 1. seeds_dev and grads_dev hold one m70h_fields per handle (NULL: zero seeds /
 * no gradients); dparams (host, may be NULL) receives 3 doubles per handle. Destroy the
 * batch before its handles; the handles' buffers stay with them (m70h_adjoint_release). */
typedef struct m70h_adjoint_batch m70h_adjoint_batch;
int m70h_adjoint_batch_create(int nscen, m70h *const *handles, m70h_adjoint_batch **out);
void m70h_adjoint_batch_destroy(m70h_adjoint_batch *b);
const char *m70h_adjoint_batch_last_error(const m70h_adjoint_batch *b);
int m70h_adjoint_batch_forward(m70h_adjoint_batch *b, int nIter, int slots);
int m70h_adjoint_batch_reverse(m70h_adjoint_batch *b, const m70h_fields *seeds_dev, const m70h_fields *grads_dev,
                               double *dparams);

#ifdef __cplusplus
}
#endif
#endif
