/*
 * tlm.cpp -- tangent-linear build of the oracle (TEST INFRASTRUCTURE).
 *
 * Compiles m70_oracle.c unchanged with every `double` replaced by the dual number of
 * tlm_dual.h. The result, build/liboracle_tlm.so, exports the oracle's API with each
 * double argument, array element and parameter widened to a (value, tangent) pair: a
 * run from an initial state and parameters with tangents returns the final state with
 * the exact directional derivative of the discrete step (Jacobian times the tangent).
 * The adjoint tests compare the GPU gradient against these tangents (dot-product test).
 * Values equal the plain oracle's bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "tlm_dual.h"

#define double orc_dual
#include "m70_oracle.c"
#undef double
