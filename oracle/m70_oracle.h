/*
 * m70_oracle.h -- CPU restatement of the Morfo70 Hydro ADI step (TEST INFRASTRUCTURE).
 *
 * This oracle is the checker for the HIP path. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it. The product library (libm70h.so) never links,
 * loads or calls it.
 *
 * It restates, in plain C99 (+ race-free OpenMP over independent rows/columns), the
 * reference's Hydro class: src/Hydro.cpp, src/HydroQx.cpp, src/HydroQy.cpp,
 * src/HydroBoundaryConditions.cpp, src/HydroParametrizations.cpp,
 * src/HydroMeshOperations.cpp, include/MeshOperations.h, src/Math.cpp:15-118,
 * src/Grid.cpp:32-92,127-145 and src/Bathymetry.cpp:33-39. Every function cites the
 * reference lines it follows. Arithmetic is kept in the reference's operation order;
 * build with -ffp-contract=off so results are bitwise identical to a serial x86-64
 * build of the reference (which has no FMA: compile.sh:41 passes no -march).
 *
 * Layout: row-major, index i*ny+j; i is cross-shore (0 = offshore), j alongshore
 * (periodic). Fields with nx+1 rows (U, Qx, V, Qy) hold the offshore ghost in row 0
 * and the onshore wall in row nx.
 *
 * What pins it (the reference's Hydro sources cannot be compiled here, and the reference
 * ships no expected outputs): its solvers equal the reference's own src/Math.cpp bit for
 * bit (built into oracle/_ref); grid weights, masks, drag, viscosity and forcing stencils
 * match closed forms; the composed step converges at second order to the continuous
 * flux-form shallow-water equations and satisfies the exact semi-implicit gravity
 * identities of its half-steps, and a build with any one of 22 assembly terms scaled by
 * 1.1 (ORC_MUTANT) fails those checks (tests/test_oracle_pins.py,
 * tests/test_oracle_consistency.py). Parity unpinned: the boundary and mask quirks no
 * continuous equation determines -- the row-0 mask in the Qy predictor
 * (src/HydroQy.cpp:140,170), the never-written mask row nx-1 (src/Hydro.cpp:91), the row-0
 * masking asymmetry of the two lower diagonals (src/HydroQx.cpp:214-229,
 * src/HydroQy.cpp:340-362) and the stale depth row nx (src/HydroQx.cpp:164-188).
 */
#ifndef M70_ORACLE_H_
#define M70_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Enumerations mirror include/OptionsParametersConstants.h:50-64. */
enum { ORC_DRAG_CONSTANT = 0, ORC_DRAG_LOGDEPTH = 1, ORC_DRAG_MANNING = 2 };
enum { ORC_VISC_CONSTANT = 0, ORC_VISC_ROLLERS = 1, ORC_VISC_DEPTH = 2, ORC_VISC_HRMS = 3 };
enum { ORC_BC_SPONGE = 0, ORC_BC_CONSTANT_ZS = 1 };

typedef struct orc_params {
    int drag_model;     /* Options::Hydro::DragModel */
    int visc_model;     /* Options::Hydro::ViscosityModel */
    int offshore_bc;    /* Options::Hydro::OffshoreBC */
    double dt;          /* Parameters::Hydro::dt */
    double Dmin;        /* Parameters::Hydro::Dmin */
    double alpha_f;     /* Parameters::Hydro::alpha_f */
    double cD0;         /* Parameters::Hydro::cD0 */
    double nu0;         /* Parameters::Hydro::nu0 */
    double M;           /* Parameters::Hydro::M */
    double ka;          /* Parameters::Hydro::ka */
    double z0;          /* Parameters::Bottom::z0 */
    double alpha_r;     /* Parameters::Rollers::alpha_r */
} orc_params;

typedef struct orc_hydro orc_hydro;

/* Grid metrics exactly as Grid::Grid (src/Grid.cpp:32-92). Output arrays: dx[nx-1],
 * dx_x[nx], difx_x_{up,d,low}[nx-1], dif2x_x_{up,d,low}[nx-1], difx_c_{up,d,low}[nx],
 * dif2x_c_{up,d,low}[nx], x2c_{up,d}[nx], packed in that order into `out`
 * (total 20*nx - 7 doubles). */
void orc_grid_metrics(int nx, const double *x, double *out);

/* Bathymetry ctor clamp (src/Bathymetry.cpp:33-39) + Hydro ctor (src/Hydro.cpp:17-133). */
orc_hydro *orc_create(int nx, int ny, const double *x, double dy, const orc_params *p,
                      const double *Z, const double *h, double zs0);
void orc_destroy(orc_hydro *H);
void orc_set_threads(int n);

/* Forcing arrays read by Run's pre-loop (include/Wave.h public members). Copied. */
void orc_set_forcing(orc_hydro *H, const double *Sxx, const double *Sxy, const double *Syy,
                     const double *Urms, const double *Dw, const double *D_r, const double *Hw);

/* Offshore Qx BC hook Qx0 = f1*Qx1 + f0 (include/Hydro.h:89-90). A series, when set,
 * supplies f0 for the k-th step counted over all Run calls since it was set. */
void orc_set_f0(orc_hydro *H, double f0);
int orc_set_f0_series(orc_hydro *H, const double *f0, long n);
void orc_bc_constants(const orc_hydro *H, double *f1, double *f0, double *bcQy);

/* State: Zs, D [F]; Qx, Qy, U, V [F+]; U_c, V_c, dQx_dx, dQy_dy [F]. Any pointer may be NULL. */
void orc_get_state(const orc_hydro *H, double *Zs, double *D, double *Qx, double *Qy,
                   double *U, double *V, double *U_c, double *V_c, double *dQx_dx,
                   double *dQy_dy);
void orc_set_state(orc_hydro *H, const double *Zs, const double *D, const double *Qx,
                   const double *Qy, const double *U, const double *V, const double *U_c,
                   const double *V_c, const double *dQx_dx, const double *dQy_dy);

/* Hydro::Run (src/Hydro.cpp:170-202). Returns 0. */
int orc_run(orc_hydro *H, int nIter);

/* Per-Run fields after the last Run's pre-loop (for stage tests). name: "cD","nuD",
 * "Forcing_x","Forcing_y","Urms_x","Urms_y","cD_x","cD_y","nuD_x","nuD_y","dnuD_dx",
 * "dnuDx_dy","dnuD_dy","dnuDy_dx". Returns number of doubles copied, or -1. */
long orc_get_field(const orc_hydro *H, const char *name, double *out);

/* Solvers, as Math::TridiagonalSolver (src/Math.cpp:15-65) and
 * Math::TridiagonalSolverPeriodic (src/Math.cpp:68-118). */
void orc_tridiag(double *Y, double *V, const double *A, const double *B, const double *C,
                 const double *F, int n, int m);
void orc_tridiag_periodic(const double *a, const double *b, const double *c, const double *f,
                          int n, double *y);

/* Wet/dry face rule of setWaterNodes_x/_y (src/HydroQx.cpp:245-253). */
int orc_wet_rule(double d, double dup, double zs, double zsup, double s, double sup,
                 double Dmin);

/* ---------------------------------------------------------------------------------
 * Sediment transport and bed update (class Sediment, include/Sediment.h:17-115,
 * src/Sediment.cpp). The sediment object reads the hydro object's Zs, D, U, V and Dw and
 * owns the bed fields h and Zb (copied from the hydro's at creation: h, Z - h).
 * update_bottom mutates the hydro's Zs, as the reference's does (src/Sediment.cpp:131).
 * --------------------------------------------------------------------------------- */
typedef struct orc_sed_params {
    double p;             /* Parameters::Bottom::p */
    double alpha0_max, alpha0_min, D_alpha0_max, D_alpha0_min, gamma0; /* Parameters::Sediment::CWS */
    double Gamma0, A_swash;                                           /* Parameters::Sediment::Swash */
} orc_sed_params;
typedef struct orc_sediment orc_sediment;

orc_sediment *orc_sed_create(orc_hydro *H, const orc_sed_params *sp);
void orc_sed_destroy(orc_sediment *S);
/* Sediment::calculateSedimentFluxes, src/Sediment.cpp:47-65 */
void orc_sed_fluxes(orc_sediment *S);
/* Sediment::updateBottom, src/Sediment.cpp:67-75,105-151; dt = Parameters::Run::dt */
void orc_sed_update_bottom(orc_sediment *S, double dt);
/* Sediment::setSedimentFluxesFromResumeFile's arithmetic, src/Sediment.cpp:77-103 */
void orc_sed_set_fluxes(orc_sediment *S, const double *qx, const double *qy);
/* Copies a field: Alpha, gamma, Gamma, qx, qy, dh_dt, h, Zb (F) or maxDw, Zs_sh (ny);
 * returns its length, -1 for an unknown name; out may be NULL to query the length. */
long orc_sed_get(orc_sediment *S, const char *name, double *out);

#ifdef __cplusplus
}
#endif
#endif
