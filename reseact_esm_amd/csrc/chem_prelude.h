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
#define RS_ISFINITE std::isfinite
#else
#define RS_FN __device__ __attribute__((always_inline)) inline
#define RS_SQRT sqrt
#define RS_FABS fabs
#define RS_POW pow
#define RS_ISFINITE __builtin_isfinite
#endif
/* Pivots.  RS_PIVOT(w) gives 1/U_kk for the elimination (the reference's
 * piv = 1 ./ W[k,k], rx_traced_integrator.jl:114); RS_DIAG_SOLVE(acc, w)
 * divides by U_kk in the substitutions (:132).  The device build stores the
 * reciprocal in the diagonal slot and multiplies by it; the exact build keeps
 * U_kk and divides, operation for operation like the reference. */
#ifdef RS_EXACT_DIV
#define RS_PIVOT(w) (1.0 / (w))
#define RS_DIAG_SOLVE(acc, w) ((acc) / (w))
#define RS_DIV(a, b) ((a) / (b))
#else
#define RS_PIVOT(w) ((w) = rs_rcp(w))
#define RS_DIAG_SOLVE(acc, w) ((acc) * (w))
/* the error norm's scaled residuals: a times the refined reciprocal of b */
#define RS_DIV(a, b) ((a) * rs_rcp(b))
#endif
/* 1/x from the hardware reciprocal estimate refined by two Newton steps
 * (~1 ulp; a zero pivot gives a NaN, which the controller rejects). */
#ifdef RS_HOST_BUILD
RS_FN double rs_rcp(double x) { return 1.0 / x; }
#else
RS_FN double rs_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
#endif
/* Julia's max: propagates a NaN in the first argument like std::max. */
RS_FN double rs_max(double a, double b) { return (a < b) ? b : a; }
#define RS_ROS23_D (1.0 / (2.0 + 1.4142135623730951))   /* rx_traced_integrator.jl:256 */
#define RS_ROS23_C32 (6.0 + 1.4142135623730951)         /* :257 */
#endif
