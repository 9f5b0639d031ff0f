// ref_wrap.cpp -- C entry points onto the reference's own solver and mesh sources
// (TEST INFRASTRUCTURE, built into oracle/_ref/ by oracle/Makefile).
//
// Compiled against /root/reference/include with /root/reference/src/Math.cpp; nothing of
// the reference is copied. The rest of the reference's Hydro path includes <netcdf.h>
// (include/Grid.h:15 -> include/Netcdf.h:15), which this image lacks, so only these
// self-contained pieces are built.
#include "Math.h"
#include "MeshOperations.h"

extern "C" {
// Math::TridiagonalSolver, src/Math.cpp:15-65
void ref_tridiag(double *Y, double *V, double *A, double *B, double *C, double *F, int n, int m)
{
    Math::TridiagonalSolver(Y, V, A, B, C, F, n, m);
}
// Math::TridiagonalSolverPeriodic (single line), src/Math.cpp:68-118
void ref_tridiag_periodic(double *a, double *b, double *c, double *f, int n, double *y)
{
    Math::TridiagonalSolverPeriodic(a, b, c, f, n, y);
}
// Math::TridiagonalSolverPeriodic (batched overload, unused by Hydro), src/Math.cpp:121-186
void ref_tridiag_periodic_batched(double *Y, double *A, double *B, double *C, double *F, int n, int m)
{
    Math::TridiagonalSolverPeriodic(Y, A, B, C, F, n, m);
}
// MeshOperations::C2X, include/MeshOperations.h:18-27
void ref_c2x(double *var_x, const double *var_c, int nx, int ny) { MeshOperations::C2X(var_x, var_c, nx, ny); }
// MeshOperations::X2C, include/MeshOperations.h:36-47
void ref_x2c(double *var_c, const double *var_x, const double *x2c_d, const double *x2c_up, int nx, int ny)
{
    MeshOperations::X2C(var_c, var_x, x2c_d, x2c_up, nx, ny);
}
}
