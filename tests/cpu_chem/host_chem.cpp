// Test helper: the runtime-compiled chemistry code built for the CPU
// (RS_HOST_BUILD), looping over cells, so tests can compare the generated
// per-cell ROS23 attempt and VJP with the oracle on identical inputs.
#define RS_HOST_BUILD
#include "chem_prelude.h"
#include "rs_launch.h"
#include "gen_mech.h"
#include "chem_kernels.hip"

extern "C" int hc_ns(void) { return RS_NS; }

// Cell c's emissions gathered from the species-major E (stride ncol) into ec,
// the contiguous per-cell form the generated code reads; 0 without any.
static const double* cell_emis(const double* E, long long c, long long ncol, double* ec) {
    if (!E || c >= ncol) return nullptr;
    for (int s = 0; s < RS_NS; ++s) ec[s] = E[s * ncol + c];
    return ec;
}

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
        double ec[RS_NS];
        const double* Ec = cell_emis(E, c, ncol, ec);
        double W[RS_NNZ];
        tot += rs_ros23_attempt<1, 1, false>(K, jf, Ec, sv, W, dt, atol, rtol,
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
        double ec[RS_NS];
        const double* Ec = cell_emis(E, c, ncol, ec);
        double* Gc = (c < ncol) ? gE + c : nullptr;
        double W[RS_NNZ];
        rs_ros23_vjp<1, 1>(K, jf, Ec, ncol, sv, W, dt, [&](int s) { return lc[s]; },
                        [&](int s, double v) { lc[s] = v; }, Gc);
        for (int s = 0; s < RS_NS; ++s) lam[s * nc + c] = lc[s];
    }
}

// The device's storage layout on the CPU: the lanes of one wave share a slab
// laid out [3][NS][64] (stride 64 between a lane's species, lane offset
// within), as the LDS slab of the chemistry kernels.  Lanes run one after the
// other on a slab whose other lanes' slots hold a sentinel, and after each lane
// every slot that is not its own must still hold the sentinel: a slab index
// that strays into a neighbour's slots is reported (returns the number of
// clobbered slots; 0 when clean).  Cells are taken 64 at a time.
extern "C" long long hc_attempt_wave(long long nc, long long ncol, const double* K,
                                     const double* jfac, const double* E, const double* u,
                                     double* unew, double dt, double atol, double rtol,
                                     double* sse_out) {
    static double slab[3 * RS_NS * 64];
    const double sentinel = -1234.5678;
    long long bad = 0;
    double tot = 0.0;
    for (long long c0 = 0; c0 < nc; c0 += 64) {
        for (int lane = 0; lane < 64; ++lane) {
            const long long c = c0 + lane;
            if (c >= nc) break;
            for (int x = 0; x < 3 * RS_NS * 64; ++x) slab[x] = sentinel;
            double* sv = slab + lane;
            for (int s = 0; s < RS_NS; ++s) sv[(2 * RS_NS + s) * 64] = u[s * nc + c];
            const double jf = jfac ? jfac[c] : 1.0;
            double ec[RS_NS];
            const double* Ec = cell_emis(E, c, ncol, ec);
            double W[RS_NNZ];
            tot += rs_ros23_attempt<64, 1, false>(K, jf, Ec, sv, W, dt, atol, rtol,
                                                  [&](int s, double un) { unew[s * nc + c] = un; });
            for (int x = 0; x < 3 * RS_NS * 64; ++x)
                if ((x & 63) != lane && slab[x] != sentinel) ++bad;
        }
    }
    *sse_out = tot;
    return bad;
}
