/*
 * wave_oracle.h -- CPU restatement of the Morfo70 wave and rollers driver (TEST
 * INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it; the product library never does).
 *
 * Restates class Wave and class WaveRow: src/Wave.cpp:80-290, src/WaveRow.cpp,
 * src/WaveDir.cpp, src/WaveEnergy.cpp, src/WaveParametrizations.cpp and
 * src/WaveRollers.cpp, in their serial (one OpenMP thread) semantics, with the
 * reference's operation order (build with -ffp-contract=off). Three places where the
 * reference's code cannot be what was meant are given their evident meaning; the GPU
 * path follows the same definitions:
 *   - src/Wave.cpp:158-163: the `else` without braces takes only `row_tmp = row_m1`, so
 *     after a row that falls on the mesh the two row buffers alias (or the new row is
 *     NULL on the first step). Here the rows swap after every step.
 *   - src/WaveEnergy.cpp:37-40: each OpenMP thread copies G and Gp from index 0 rather
 *     than from its chunk; with one thread (chunk = whole row) that is the identity used
 *     here.
 *   - src/WaveParametrizations.cpp:357-378,394-421: the bottom-friction loops index the
 *     first nNodesToSolve nodes instead of the nodes still being solved; here each
 *     unconverged node updates its own friction terms. With friction off (the default,
 *     src/OptionsParametersConstants.cpp:17) the two agree.
 * BreakingModel None leaves the dissipation member pointer NULL (src/WaveRow.cpp:68-80)
 * and the reference would call through it; it is rejected here.
 *
 * Parity of this restatement is pinned by closed forms and invariants
 * (tests/test_oracle_pins.py): the Doppler dispersion relation, the scheme's discrete
 * energy-flux identity, Snell's law on alongshore-uniform beaches, y-mirror and periodic
 * shift symmetry, and the dissipation closed forms. The reference's wave sources cannot
 * be built here (they include Netcdf.h through Grid.h and libnetcdf is absent), and its
 * tests hold no wave outputs.
 *
 * Layout: row-major i*ny+j, i cross-shore (0 = offshore), j alongshore (periodic).
 */
#ifndef M70_WAVE_ORACLE_H_
#define M70_WAVE_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* include/OptionsParametersConstants.h:31-40 */
enum { ORC_BREAK_CT = 1, ORC_BREAK_TG = 2, ORC_BREAK_REGULAR = 3 };
enum { ORC_FRIC_CONSTANT = 0, ORC_FRIC_SOULSBY97 = 1 };

typedef struct orc_wave_params {
    int breaking_model;      /* Options::Wave::BreakingModel */
    int friction_model;      /* Options::Wave::fwModel */
    int bottom_friction;     /* Options::Wave::BottomFriction */
    int rollers;             /* Options::Rollers::CalculateRollers */
    int max_iterations;      /* Parameters::Wave::SolverMaxIterations */
    double tolerance;        /* Parameters::Wave::SolverTolerance */
    double B3, gamma_b, m_mon, n_mon;   /* Parameters::Wave::* */
    double theta_max;        /* Parameters::Wave::ThetaMax, radians */
    double fw0;              /* Parameters::Wave::fw0 */
    double z0;               /* Parameters::Bottom::z0 */
    double sinBeta_r, alpha_r;          /* Parameters::Rollers::* */
} orc_wave_params;

typedef struct orc_wave orc_wave;

orc_wave *orc_wave_create(int nx, int ny, const double *x, double dy, const orc_wave_params *p);
void orc_wave_destroy(orc_wave *W);
/* Water depth D [nx*ny] (bathy->D, read at every calculation). */
void orc_wave_set_depth(orc_wave *W, const double *D);
/* Wave::setUV (src/Wave.cpp:293-302): currents and their projection on the stored wave
 * direction. */
void orc_wave_set_uv(orc_wave *W, const double *U, const double *V);
/* Wave::calculateWaveField (src/Wave.cpp:80-172). Returns the number of marching steps
 * (mesh rows plus intermediate rows), or -1 for a bad breaking model. */
long orc_wave_calculate(orc_wave *W, double hOff, double thetaOff, double T);
/* Public arrays by name (K = log of the wave number, Theta, Cg, H, E, Urms, Sxx_w, Sxy_w,
 * Syy_w, Dw, E_r, D_r, Sxx_r, Sxy_r, Syy_r, Sxx, Sxy, Syy, CosTheta, SinTheta, U_w).
 * Returns the length, or -1 for an unknown name; out may be NULL. */
long orc_wave_get(const orc_wave *W, const char *name, double *out);

/* Per-node pieces, for the closed-form pins. */
void orc_wave_dispersion(double D, double U_w, double w, double z0, double *lK, double *sigma, double *C,
                         double *Cg, double *Urms_H);
void orc_wave_breaking(const orc_wave_params *p, double H, double sigma, double D, double *dw, double *ddw_dh);

#ifdef __cplusplus
}
#endif
#endif
