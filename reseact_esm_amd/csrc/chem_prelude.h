/* chem_prelude.h -- macros and helpers seen by the generated mechanism code.
 * Compiled at run time by hipRTC for gfx950 (device build) and by g++ for the
 * CPU check of the generator (RS_HOST_BUILD). */
#ifndef RS_CHEM_PRELUDE_H
#define RS_CHEM_PRELUDE_H
#ifdef RS_HOST_BUILD
#include <cmath>
#define RS_FN static inline
#define RS_SQRT std::sqrt
#define RS_FABS std::fabs
#define RS_POW std::pow
#else
#define RS_FN __device__ __attribute__((always_inline)) inline
#define RS_SQRT sqrt
#define RS_FABS fabs
#define RS_POW pow
#endif
/* Pivot reciprocals and back substitution x_i = acc / U_ii.  The device
 * keeps 1/U_ii in the diagonal slot of the factor (one register per species
 * fewer) and multiplies; the exact form keeps U_ii, stores the reciprocals
 * aside and divides, as the reference's blocksolve does
 * (rx_traced_integrator.jl:114, :132). */
#ifdef RS_EXACT_DIV
#define RS_DINV_SET(k, s, v) (dinv[k] = (v))
#define RS_DINV(k, s) (dinv[k])
#define RS_DIAG_SOLVE(acc, w, di) ((acc) / (w))
#else
#define RS_DINV_SET(k, s, v) (W[s] = (v))
#define RS_DINV(k, s) (W[s])
#define RS_DIAG_SOLVE(acc, w, di) ((acc) * (di))
#endif
/* Julia's max: propagates a NaN in the first argument like std::max. */
RS_FN double rs_max(double a, double b) { return (a < b) ? b : a; }
#define RS_ROS23_D (1.0 / (2.0 + 1.4142135623730951))   /* rx_traced_integrator.jl:256 */
#define RS_ROS23_C32 (6.0 + 1.4142135623730951)         /* :257 */
#endif
