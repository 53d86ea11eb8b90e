// Test helper: the runtime-compiled chemistry code built for the CPU
// (RS_HOST_BUILD), looping over cells, so tests can compare the generated
// per-cell ROS23 attempt and VJP with the oracle on identical inputs.
#define RS_HOST_BUILD
#include "chem_prelude.h"
#include "rs_launch.h"
#include "gen_mech.h"
#include "chem_kernels.hip"

extern "C" int hc_ns(void) { return RS_NS; }

// One attempt for every cell; unew raw; returns the sum of per-cell sse in
// cell order.
extern "C" double hc_attempt(long long nc, long long ncol, const double* K, const double* jfac,
                             const double* E, const double* u, double* unew, double dt,
                             double atol, double rtol) {
    double tot = 0.0;
    for (long long c = 0; c < nc; ++c) {
        double sv[3 * RS_NS];
        for (int s = 0; s < RS_NS; ++s) sv[2 * RS_NS + s] = u[s * nc + c];
        const double jf = jfac ? jfac[c] : 1.0;
        const double* Ec = (E && c < ncol) ? E + c : nullptr;
        double W[RS_NNZ];
        tot += rs_ros23_attempt<1, 1, false>(K, jf, Ec, ncol, sv, W, dt, atol, rtol,
                                             [&](int s, double un) { unew[s * nc + c] = un; });
    }
    return tot;
}

extern "C" void hc_vjp(long long nc, long long ncol, const double* K, const double* jfac,
                       const double* E, const double* u, double dt, double* lam, double* gE) {
    for (long long c = 0; c < nc; ++c) {
        double sv[3 * RS_NS], lc[RS_NS];
        for (int s = 0; s < RS_NS; ++s) {
            sv[s] = u[s * nc + c];
            lc[s] = lam[s * nc + c];
        }
        const double jf = jfac ? jfac[c] : 1.0;
        const double* Ec = (E && c < ncol) ? E + c : nullptr;
        double* Gc = (c < ncol) ? gE + c : nullptr;
        double W[RS_NNZ];
        rs_ros23_vjp<1, 1>(K, jf, Ec, ncol, sv, W, dt, [&](int s) { return lc[s]; },
                        [&](int s, double v) { lc[s] = v; }, Gc);
        for (int s = 0; s < RS_NS; ++s) lam[s * nc + c] = lc[s];
    }
}
