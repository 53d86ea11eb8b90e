// reseact_cpu.cpp -- the CPU ORACLE of the operator-split step.
//
// TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load libreseact_cpu.so, and only as the
// checker (or as the timed CPU baseline, kind "port").  The product path is
// libreseact_hip.so; nothing in reseact_esm_amd/ loads this library.
//
// It is a plain FP64 restatement of the reference's Julia algebra, behind the
// same C ABI as the HIP library (include/reseact_hip.h), so every parity test is
// "same call, two libraries, compare".  Built with -ffp-contract=off so that
// each statement rounds exactly as the reference's one-operation broadcasts do.
// OpenMP parallelises over cells / rows only; every reduction is done in a fixed
// order, so results do not depend on the thread count.
//
// Function -> reference map (file:line under EarthSciML/reseact.esm):
//   pi_update            PICtrl + adaptive loop body, tools/reactant_handoff/
//                        rx_traced_integrator.jl:86-99, :480-505;
//                        host_adaptive! tools/adjoint_gradient.jl:538-582
//   blocksolve           rx_traced_integrator.jl:110-135 (pivot-free elimination)
//   ssprk43_attempt      ssprk43_step_unrolled rx_traced_integrator.jl:404-437
//   ros23_cell           ros23_step rx_traced_integrator.jl:279-349, jac = :sym
//   macro step           macro_step tools/adjoint_gradient.jl:609-619
//   replay / adjoint     replay_fixed :647-661, forward_pass :671-688,
//                        backward_stage! :731-784, backward_sweep :813-834
//   one-step VJPs        hand-derived adjoints of the two step maps (the
//                        reference obtains them with Enzyme, :610-623)
// The transport RHS (3rd-order upwind, periodic longitude) and the synthetic
// mass-action chemistry are the workloads' own definitions; the reference's
// PPM / SuperFast right-hand sides are not on this path.
//
// Parity pins (tests/): SSPRK43 order 3 and ROS23 order 2 / L-stability on
// linear problems against expm, transport mass conservation and freestream,
// upwind order-3 convergence, analytic exp(Kt) for first-order mechanisms,
// FD-vs-analytic Jacobian, JVP-vs-FD, dot-product identity VJP<->JVP, and
// the reference's own toy problem thresholds (tools/rx_adjoint_toy.jl:17-24).

#include "../include/reseact_hip.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double ROS23_D = 1.0 / (2.0 + std::sqrt(2.0));   // rx_traced_integrator.jl:256
const double ROS23_C32 = 6.0 + std::sqrt(2.0);          // :257
const double C16 = 1.0 / 6.0;

struct Seq { std::vector<std::pair<double, double>> ts; };

}  // namespace

struct rs_ctx {
    int nlon = 0, nlat = 0, nlev = 0, ns = 0, nt = 0;
    size_t nc = 0, ncol = 0;
    std::vector<double> cx, wz, wl;
    bool have_wind = false, have_mech = false, have_state = false;
    // PPM transport (scheme RS_TRANSPORT_PPM)
    int scheme = RS_TRANSPORT_UPWIND3;
    int lon_periodic = 1;
    double dlon = 0.0;
    std::vector<double> coslat_e, dS_lat, dphi_lat;   // nlat+1, nlat, nlat
    std::vector<double> Mx, My, Mz, dp;               // faces / cells, lon fastest
    std::vector<double> qw, qe, qs, qn;               // inflow halos, (nt-1) species
    std::vector<double> dPSdt, msrc;                  // pressure fixer (ncol), air-mass source (nc)
    // PBL mixing of the mixing ratios: layer fraction in the PBL (nc), its
    // weight dp * pbl_f (nc), max(column weight, 1) (ncol), relaxation time
    std::vector<double> pblf, pblw, pblWp;
    double pbl_tau = 0.0;
    bool have_ppm_geom = false, have_ppm_flux = false;
    // mechanism
    int nr = 0;
    std::vector<int> ra, rb, pa, pb, photo;
    std::vector<double> kr;
    std::vector<std::vector<std::pair<int, double>>> net;   // per reaction: (species, nu), nu != 0
    std::vector<double> jfac;   // nc
    std::vector<double> emis;   // ns*ncol, empty = none
    // state
    std::vector<double> u, cand;
    bool have_cand = false;
    std::vector<double> lam, gemis;
    std::vector<double> wobj;   // linear objective weights (loss kind 1), empty = unset
    // time-loop adjoint records (last rs_adjoint_run)
    std::vector<Seq> seqT, seqC;
    // the last rs_adaptive_transport / rs_adaptive_chem call: every attempt as
    // (t, dtc, EEst, outcome: 1 accept, 0 reject, -1 out-of-domain reject),
    // and the accepted (t, dt) per operator (rs_get_step_log)
    std::vector<std::array<double, 4>> alog;
    Seq lastseq[2];
    rs_split_opts last_opts{};
    bool have_record = false;
    std::vector<double> rec_tstops;   // macro stops of the last adjoint run
    double rec_dtC0 = 0.0;            // its per-cell chemistry start dt
    std::string err;
};

namespace {

int fail(rs_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// ----------------------------------------------------------------------------
// PI controller (rx_traced_integrator.jl:480-505; adjoint_gradient.jl:553-577)
// ----------------------------------------------------------------------------
struct PIState {
    double t, dt, qold;
    long long nacc = 0, nrej = 0, iters = 0;
};

// Returns true on accept.  dtc is the attempted step.
bool pi_update(const rs_pictrl& c, PIState& s, double dtc, double eest_raw) {
    const double invqmax = 1.0 / c.qmax, invqmin = 1.0 / c.qmin;
    double EEst = std::isnan(eest_raw) ? 1.0e10 : eest_raw;       // :484
    double q11 = std::pow(std::max(EEst, 1.0e-35), c.beta1);      // :486
    double qden = std::pow(s.qold, c.beta2);                      // :487
    double q = q11 / qden;                                         // :488
    q = std::max(invqmax, std::min(invqmin, q / c.gamma));        // :489
    bool accept = EEst <= 1.0;                                     // :490
    bool insteady = (c.qsteady_min <= q) && (q <= c.qsteady_max); // :492
    double qa = insteady ? 1.0 : q;                                // :493
    if (accept) {
        s.t = s.t + dtc;
        s.dt = dtc / qa;                                           // :494
        s.qold = std::max(EEst, c.qoldinit);                       // :501
        s.nacc += 1;
    } else {
        double qr = std::min(invqmin, q11 / c.gamma);              // :496
        s.dt = dtc / qr;                                           // :497
        s.nrej += 1;
    }
    s.iters += 1;
    return accept;
}

// ----------------------------------------------------------------------------
// Transport operator: 3rd-order upwind flux form, periodic in longitude,
// closed bottom/top, vertical stencil indices clamped to the column.
// ----------------------------------------------------------------------------
inline double face3(double qmm, double qm, double qp) {
    // upwind-biased 3rd-order face value (-q_{m-1} + 5 q_m + 2 q_{m+1}) / 6
    return (5.0 * qm - qmm + 2.0 * qp) * C16;
}

struct TGeom {
    int nlon, nlat, nlev;
    size_t nc;
};

// f = L q for one species block (q, f of length nc).
void transport_rhs_block(const rs_ctx& c, const double* q, double* f) {
    const int nlon = c.nlon, nlat = c.nlat, nlev = c.nlev;
    const size_t rowlen = (size_t)nlon, plane = (size_t)nlon * nlat;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nlev; ++k) {
        for (int j = 0; j < nlat; ++j) {
            const double cxj = c.cx[j];
            const double* qr = q + k * plane + j * rowlen;
            for (int i = 0; i < nlon; ++i) {
                auto Q = [&](int ii) { ii %= nlon; if (ii < 0) ii += nlon; return qr[ii]; };
                double fe, fw;
                if (cxj >= 0.0) {
                    fe = cxj * face3(Q(i - 1), Q(i), Q(i + 1));
                    fw = cxj * face3(Q(i - 2), Q(i - 1), Q(i));
                } else {
                    fe = cxj * face3(Q(i + 2), Q(i + 1), Q(i));
                    fw = cxj * face3(Q(i + 1), Q(i), Q(i - 1));
                }
                auto V = [&](int kk) { return q[(size_t)kk * plane + j * rowlen + i]; };
                auto G = [&](int kf) -> double {   // flux through face kf (between kf-1 and kf)
                    if (kf <= 0 || kf >= nlev) return 0.0;
                    double w = c.wz[kf] * c.wl[i];
                    if (w >= 0.0) {
                        int mm = std::max(kf - 2, 0);
                        return w * face3(V(mm), V(kf - 1), V(kf));
                    } else {
                        int mm = std::min(kf + 1, nlev - 1);
                        return w * face3(V(mm), V(kf), V(kf - 1));
                    }
                };
                double gt = G(k + 1), gb = G(k);
                f[k * plane + j * rowlen + i] = -(fe - fw) - (gt - gb);
            }
        }
    }
}

// g += L^T lam for one species block, by reversing the face fluxes: every
// face flux F leaves its lower/western cell and enters its upper/eastern cell.
void transport_rhs_adj_block(const rs_ctx& c, const double* lam, double* g) {
    const int nlon = c.nlon, nlat = c.nlat, nlev = c.nlev;
    const size_t rowlen = (size_t)nlon, plane = (size_t)nlon * nlat;
    // zonal faces: face (i+1/2) between i (west) and i+1 (east); rows are
    // independent, so parallelise over (k, j).
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nlev; ++k) {
        for (int j = 0; j < nlat; ++j) {
            const double cxj = c.cx[j];
            const size_t b = k * plane + j * rowlen;
            auto W = [&](int ii) { ii %= nlon; if (ii < 0) ii += nlon; return (size_t)ii; };
            for (int i = 0; i < nlon; ++i) {
                // F_{i+1/2}: f_i -= F, f_{i+1} += F  => Fbar = lam_{i+1} - lam_i
                double Fb = lam[b + W(i + 1)] - lam[b + W(i)];
                int mm, m, p;
                if (cxj >= 0.0) { mm = i - 1; m = i; p = i + 1; }
                else { mm = i + 2; m = i + 1; p = i; }
                double a = cxj * Fb * C16;
                g[b + W(mm)] += -a;
                g[b + W(m)] += 5.0 * a;
                g[b + W(p)] += 2.0 * a;
            }
        }
    }
    // vertical faces kf = 1..nlev-1 between kf-1 (below) and kf (above);
    // columns are independent, parallelise over (j, i).
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < nlat; ++j) {
        for (int i = 0; i < nlon; ++i) {
            const size_t b = j * rowlen + i;
            for (int kf = 1; kf < nlev; ++kf) {
                double w = c.wz[kf] * c.wl[i];
                double Gb = lam[b + kf * plane] - lam[b + (kf - 1) * plane];
                int mm, m, p;
                if (w >= 0.0) { mm = std::max(kf - 2, 0); m = kf - 1; p = kf; }
                else { mm = std::min(kf + 1, nlev - 1); m = kf; p = kf - 1; }
                double a = w * Gb * C16;
                g[b + mm * plane] += -a;
                g[b + m * plane] += 5.0 * a;
                g[b + p * plane] += 2.0 * a;
            }
        }
    }
}

// Blocked sum in a fixed order (independent of the OpenMP thread count).
double fixed_sum(const std::vector<double>& v) {
    const size_t B = 4096;
    size_t nb = (v.size() + B - 1) / B;
    std::vector<double> part(nb, 0.0);
#pragma omp parallel for schedule(static)
    for (long long b = 0; b < (long long)nb; ++b) {
        double s = 0.0;
        size_t e = std::min(v.size(), (size_t)(b + 1) * B);
        for (size_t x = (size_t)b * B; x < e; ++x) s += v[x];
        part[b] = s;
    }
    double s = 0.0;
    for (double x : part) s += x;
    return s;
}

// ----------------------------------------------------------------------------
// Monotone PPM in CWC mixing-ratio form (reseact.esm:1536-3100 and
// prototypes/transport_3d/transport_3d.esm; the discretisation rules of
// reseact.esm:61-110 are not in the reference tree).  Restatement of
// Colella & Woodward (1984), method-of-lines form:
//   eq. 1.7-1.8  monotone slope  dm_i = sgn(da) min(|da|, 2|a_i-a_{i-1}|,
//                2|a_{i+1}-a_i|) if (a_{i+1}-a_i)(a_i-a_{i-1}) > 0, else 0,
//                da = (a_{i+1}-a_{i-1})/2 (uniform) or the weighted form;
//   eq. 1.6      edge value a_{i+1/2} = a_i + (a_{i+1}-a_i)/2 - (dm_{i+1}-dm_i)/6
//                (uniform) or the width-weighted form (vertical, widths dp);
//   eq. 1.10     parabola limiter on (aL, aR) of each cell;
// the face value is the limited edge of the upwind cell.  Lateral open walls
// fill 3 halo cells with the prescribed inflow mixing ratio; the vertical
// clamps its stencil at the surface and the lid (no flux there).
// ----------------------------------------------------------------------------
inline double ppm_slope(double am, double a0, double ap) {
    const double dl = a0 - am, dr = ap - a0;
    if (dr * dl > 0.0) {
        const double da = 0.5 * (ap - am);
        const double m = std::min(std::fabs(da), std::min(2.0 * std::fabs(dl), 2.0 * std::fabs(dr)));
        return std::copysign(m, da);
    }
    return 0.0;
}

inline double ppm_edge(double a0, double a1, double d0, double d1) {
    const double e = a0 + 0.5 * (a1 - a0);
    return e - (d1 - d0) * C16;
}

// Width-weighted slope (eq. 1.7) of cell 0 with neighbour widths hm, hp:
// da = h0/(hm+h0+hp) [(2hm+h0)/(hp+h0) dr + (h0+2hp)/(hm+h0) dl], written with
// the two width-only factors cR, cL (a kernel may tabulate them per cell).
inline double ppm_slope_nu(double hm, double h0, double hp, double am, double a0, double ap) {
    const double dl = a0 - am, dr = ap - a0;
    if (dr * dl > 0.0) {
        const double w = h0 / (hm + h0 + hp);
        const double cR = w * ((2.0 * hm + h0) / (hp + h0));
        const double cL = w * ((h0 + 2.0 * hp) / (hm + h0));
        const double da = cR * dr + cL * dl;
        const double m = std::min(std::fabs(da), std::min(2.0 * std::fabs(dl), 2.0 * std::fabs(dr)));
        return std::copysign(m, da);
    }
    return 0.0;
}

// Width-weighted edge between cells 0 and 1 (eq. 1.6); hm, h2 the outer widths:
// a_{1/2} = a0 + h0/(h0+h1) da + S^-1 { 2 h1 h0/(h0+h1) [t1 - t2] da
//           - h0 t1 d1 + h1 (h1+h2)/(h0+2h1) d0 },  S = hm+h0+h1+h2,
// t1 = (hm+h0)/(2h0+h1), t2 = (h2+h1)/(2h1+h0); written with the three
// width-only factors cE, cF, cG (a kernel may tabulate them per edge).
inline double ppm_edge_nu(double hm, double h0, double h1, double h2, double a0, double a1,
                          double d0, double d1) {
    const double cA = h0 / (h0 + h1);
    const double g = (2.0 * h1 * h0) / (h0 + h1);
    const double t1 = (hm + h0) / (2.0 * h0 + h1);
    const double t2 = (h2 + h1) / (2.0 * h1 + h0);
    const double cS = 1.0 / (((hm + h0) + h1) + h2);
    const double cE = cA + (g * (t1 - t2)) * cS;
    const double cF = (h0 * t1) * cS;
    const double cG = (h1 * ((h1 + h2) / (h0 + 2.0 * h1))) * cS;
    const double da = a1 - a0;
    const double e0 = a0 + cE * da;
    return (e0 - cF * d1) + cG * d0;
}

// Parabola limiter (eq. 1.10) of a cell with mean a0 and edges aL, aR.
inline void ppm_limit(double a0, double& aL, double& aR) {
    if ((aR - a0) * (a0 - aL) <= 0.0) {
        aL = a0;
        aR = a0;
        return;
    }
    const double d = aR - aL;
    const double a6 = 6.0 * (a0 - 0.5 * (aL + aR));
    const double dd = d * d;
    const double da6 = d * a6;
    if (da6 > dd) aL = 3.0 * a0 - 2.0 * aR;
    else if (-dd > da6) aR = 3.0 * a0 - 2.0 * aL;
}

// Face value from the six cells v[0..5] = a_{f-3} .. a_{f+2} around face f,
// uniform spacing: the limited right edge of cell f-1 for M >= 0, the limited
// left edge of cell f otherwise.
inline double ppm_face(const double* v, double M) {
    if (M >= 0.0) {
        const double dm2 = ppm_slope(v[0], v[1], v[2]);
        const double dm1 = ppm_slope(v[1], v[2], v[3]);
        const double d0 = ppm_slope(v[2], v[3], v[4]);
        double aL = ppm_edge(v[1], v[2], dm2, dm1);
        double aR = ppm_edge(v[2], v[3], dm1, d0);
        ppm_limit(v[2], aL, aR);
        return aR;
    }
    const double dm1 = ppm_slope(v[1], v[2], v[3]);
    const double d0 = ppm_slope(v[2], v[3], v[4]);
    const double dp1 = ppm_slope(v[3], v[4], v[5]);
    double aL = ppm_edge(v[2], v[3], dm1, d0);
    double aR = ppm_edge(v[3], v[4], d0, dp1);
    ppm_limit(v[3], aL, aR);
    return aL;
}

// The same with widths h[0..5] of the six cells (vertical).
inline double ppm_face_nu(const double* v, const double* h, double M) {
    if (M >= 0.0) {
        const double dm2 = ppm_slope_nu(h[0], h[1], h[2], v[0], v[1], v[2]);
        const double dm1 = ppm_slope_nu(h[1], h[2], h[3], v[1], v[2], v[3]);
        const double d0 = ppm_slope_nu(h[2], h[3], h[4], v[2], v[3], v[4]);
        double aL = ppm_edge_nu(h[0], h[1], h[2], h[3], v[1], v[2], dm2, dm1);
        double aR = ppm_edge_nu(h[1], h[2], h[3], h[4], v[2], v[3], dm1, d0);
        ppm_limit(v[2], aL, aR);
        return aR;
    }
    const double dm1 = ppm_slope_nu(h[1], h[2], h[3], v[1], v[2], v[3]);
    const double d0 = ppm_slope_nu(h[2], h[3], h[4], v[2], v[3], v[4]);
    const double dp1 = ppm_slope_nu(h[3], h[4], h[5], v[3], v[4], v[5]);
    double aL = ppm_edge_nu(h[1], h[2], h[3], h[4], v[2], v[3], dm1, d0);
    double aR = ppm_edge_nu(h[2], h[3], h[4], h[5], v[3], v[4], d0, dp1);
    ppm_limit(v[3], aL, aR);
    return aL;
}

// Flux divergence of one cell from its six face fluxes, in the operator's
// fixed order ((x + y) + z); gx = dphi_lat / (dlon dS_lat), rdS = 1 / dS_lat.
inline double ppm_div(double fw, double fe, double fs, double fn, double fb, double ft,
                      double gx, double ces, double cen, double rdS) {
    const double dx = (fe - fw) * gx;
    const double dy = (cen * fn - ces * fs) * rdS;
    const double dz = ft - fb;
    return (dx + dy) + dz;
}

// PBL means of the mixing ratios of u, per species and column:
// pbl_mean = sum_k (dp pbl_f)_k q_k / max(sum_k (dp pbl_f)_k, 1), sums in level
// order (reseact.esm:7698-7760 pbl_wt, pbl_sum_X; :8455 pbl_mean_X).
static std::vector<double> ppm_pbl_means(const rs_ctx& c, const double* u) {
    const size_t nc = c.nc, plane = (size_t)c.nlon * c.nlat;
    const int nq = c.nt - 1;
    std::vector<double> pm((size_t)std::max(nq, 0) * plane);
    for (int s = 1; s < c.nt; ++s) {
        const double* q = u + (size_t)s * nc;
        for (size_t col = 0; col < plane; ++col) {
            double S = 0.0;
            for (int k = 0; k < c.nlev; ++k) {
                const size_t x = (size_t)k * plane + col;
                S = S + c.pblw[x] * q[x];
            }
            pm[(size_t)(s - 1) * plane + col] = S / c.pblWp[col];
        }
    }
    return pm;
}

// f = L_ppm u over the transported block (species 0 = air mass).  The
// mixing-ratio tendency is the CWC advection [q div(M) - div(M q_f)] / m plus,
// when set, the PBL mixing pbl_f (pbl_mean - q) / tau (reseact.esm:1563-1690);
// with_pbl = false leaves the latter out (the derivatives take the advective
// part for the air-mass row).
void ppm_rhs_all(const rs_ctx& c, const double* u, double* f, bool with_pbl) {
    const int nlon = c.nlon, nlat = c.nlat, nlev = c.nlev, nt = c.nt;
    const size_t nc = c.nc, plane = (size_t)nlon * nlat;
    const double* m = u;
    const bool pbl = with_pbl && !c.pblf.empty();
    const std::vector<double> pm = pbl ? ppm_pbl_means(c, u) : std::vector<double>();
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nlev; ++k) {
        for (int j = 0; j < nlat; ++j) {
            const double gx = c.dphi_lat[j] / (c.dlon * c.dS_lat[j]);
            const double ces = c.coslat_e[j], cen = c.coslat_e[j + 1], rdS = 1.0 / c.dS_lat[j];
            for (int i = 0; i < nlon; ++i) {
                const size_t cell = (size_t)k * plane + (size_t)j * nlon + i;
                const double* mxr = c.Mx.data() + ((size_t)k * nlat + j) * (nlon + 1);
                const double mw = mxr[i], me = mxr[i + 1];
                const double ms = c.My[((size_t)k * (nlat + 1) + j) * nlon + i];
                const double mn = c.My[((size_t)k * (nlat + 1) + j + 1) * nlon + i];
                const double mb = c.Mz[(size_t)k * plane + (size_t)j * nlon + i];
                const double mt = c.Mz[(size_t)(k + 1) * plane + (size_t)j * nlon + i];
                const double divM = ppm_div(mw, me, ms, mn, mb, mt, gx, ces, cen, rdS);
                f[cell] = c.msrc.empty() ? -divM : -divM + c.msrc[cell];
                for (int s = 1; s < nt; ++s) {
                    const double* q = u + (size_t)s * nc;
                    const int sq = s - 1;
                    // lon line (periodic wrap or the w/e inflow halo)
                    auto X = [&](int ii) -> double {
                        if (c.lon_periodic) {
                            ii %= nlon;
                            if (ii < 0) ii += nlon;
                            return q[(size_t)k * plane + (size_t)j * nlon + ii];
                        }
                        if (ii < 0) return c.qw[((size_t)sq * nlev + k) * nlat + j];
                        if (ii >= nlon) return c.qe[((size_t)sq * nlev + k) * nlat + j];
                        return q[(size_t)k * plane + (size_t)j * nlon + ii];
                    };
                    auto Y = [&](int jj) -> double {
                        if (jj < 0) return c.qs[((size_t)sq * nlev + k) * nlon + i];
                        if (jj >= nlat) return c.qn[((size_t)sq * nlev + k) * nlon + i];
                        return q[(size_t)k * plane + (size_t)jj * nlon + i];
                    };
                    auto Zc = [&](int kk) { return std::min(std::max(kk, 0), nlev - 1); };
                    double v[6], h[6];
                    for (int d = 0; d < 6; ++d) v[d] = X(i - 3 + d);
                    const double qw_ = ppm_face(v, mw);
                    for (int d = 0; d < 6; ++d) v[d] = X(i - 2 + d);
                    const double qe_ = ppm_face(v, me);
                    for (int d = 0; d < 6; ++d) v[d] = Y(j - 3 + d);
                    const double qs_ = ppm_face(v, ms);
                    for (int d = 0; d < 6; ++d) v[d] = Y(j - 2 + d);
                    const double qn_ = ppm_face(v, mn);
                    for (int d = 0; d < 6; ++d) {
                        const size_t x = (size_t)Zc(k - 3 + d) * plane + (size_t)j * nlon + i;
                        v[d] = q[x];
                        h[d] = c.dp[x];
                    }
                    const double qb_ = ppm_face_nu(v, h, mb);
                    for (int d = 0; d < 6; ++d) {
                        const size_t x = (size_t)Zc(k - 2 + d) * plane + (size_t)j * nlon + i;
                        v[d] = q[x];
                        h[d] = c.dp[x];
                    }
                    const double qt_ = ppm_face_nu(v, h, mt);
                    const double divF = ppm_div(mw * qw_, me * qe_, ms * qs_, mn * qn_, mb * qb_,
                                                mt * qt_, gx, ces, cen, rdS);
                    const double qc = q[cell];
                    const double num = qc * divM - divF;
                    const double fa = num / m[cell];
                    double fo = fa;
                    if (pbl) {
                        const double t = c.pblf[cell] * (pm[(size_t)sq * plane + (size_t)j * nlon + i] - qc);
                        fo = fa + t / c.pbl_tau;
                    }
                    f[(size_t)s * nc + cell] = fo;
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Derivatives of the PPM tendency, on the branch the forward takes at u.  The
// slope and parabola limiters switch on products of neighbour differences, so
// L_ppm is only piecewise smooth; these are the derivatives of the active
// branch, the derivative almost everywhere (the reference's caveat,
// DIFFERENTIABILITY_PLAN.md:709-736: on an exactly uniform field every switch
// sits on its surface and finite differences see other branches).
// Two independent derivations: a forward-mode tangent (value-and-tangent per
// operation, for the JVP) and the local gradient of each face value with
// respect to its six stencil cells (reverse rules, for the VJP).
// ----------------------------------------------------------------------------
struct Tan { double v, d; };

// slope of a cell from (am, a0, ap): value and tangent; nu selects the
// width-weighted form with widths (hm, h0, hp)
inline Tan ppm_slope_tan(Tan am, Tan a0, Tan ap, bool nu, double hm, double h0, double hp) {
    const double val = nu ? ppm_slope_nu(hm, h0, hp, am.v, a0.v, ap.v) : ppm_slope(am.v, a0.v, ap.v);
    const double dl = a0.v - am.v, dr = ap.v - a0.v;
    if (!(dr * dl > 0.0)) return {val, 0.0};
    double da, dda;
    if (nu) {
        const double w = h0 / (hm + h0 + hp);
        const double cR = w * ((2.0 * hm + h0) / (hp + h0));
        const double cL = w * ((h0 + 2.0 * hp) / (hm + h0));
        da = cR * dr + cL * dl;
        dda = cR * (ap.d - a0.d) + cL * (a0.d - am.d);
    } else {
        da = 0.5 * (ap.v - am.v);
        dda = 0.5 * (ap.d - am.d);
    }
    // value = sgn(da) min(|da|, min(2|dl|, 2|dr|)); dl, dr and da share one sign
    const double x = std::fabs(da), y = 2.0 * std::fabs(dl), z = 2.0 * std::fabs(dr);
    const double in = (z < y) ? z : y;
    if (in < x) return {val, (z < y) ? 2.0 * (ap.d - a0.d) : 2.0 * (a0.d - am.d)};
    return {val, dda};
}

inline Tan ppm_edge_tan(Tan a0, Tan a1, Tan d0, Tan d1, bool nu, double hm, double h0, double h1,
                        double h2) {
    if (!nu) {
        const double val = ppm_edge(a0.v, a1.v, d0.v, d1.v);
        const double t = a0.d + 0.5 * (a1.d - a0.d);
        return {val, t - (d1.d - d0.d) * C16};
    }
    const double val = ppm_edge_nu(hm, h0, h1, h2, a0.v, a1.v, d0.v, d1.v);
    // the edge is linear in (a0, a1, d0, d1): its tangent is the edge of the tangents
    return {val, ppm_edge_nu(hm, h0, h1, h2, a0.d, a1.d, d0.d, d1.d)};
}

inline void ppm_limit_tan(Tan a0, Tan& aL, Tan& aR) {
    double l = aL.v, r = aR.v;
    ppm_limit(a0.v, l, r);
    if ((aR.v - a0.v) * (a0.v - aL.v) <= 0.0) {
        aL = {l, a0.d};
        aR = {r, a0.d};
        return;
    }
    const double d = aR.v - aL.v;
    const double a6 = 6.0 * (a0.v - 0.5 * (aL.v + aR.v));
    const double dd = d * d;
    const double da6 = d * a6;
    if (da6 > dd) aL = {l, 3.0 * a0.d - 2.0 * aR.d};
    else if (-dd > da6) aR = {r, 3.0 * a0.d - 2.0 * aL.d};
}

// tangent of the face value of face f from v[0..5] (cells f-3..f+2), widths h
// (null: uniform)
inline double ppm_face_tan(const double* v, const double* vd, const double* h, double M) {
    const bool nu = h != nullptr;
    auto H = [&](int x) { return nu ? h[x] : 1.0; };
    Tan t[6];
    for (int d = 0; d < 6; ++d) t[d] = {v[d], vd[d]};
    const int c = M >= 0.0 ? 2 : 3;             // upwind cell
    auto sl = [&](int x) { return ppm_slope_tan(t[x - 1], t[x], t[x + 1], nu, H(x - 1), H(x), H(x + 1)); };
    const Tan sm = sl(c - 1), s0 = sl(c), sp = sl(c + 1);
    Tan aL = ppm_edge_tan(t[c - 1], t[c], sm, s0, nu, H(c - 2), H(c - 1), H(c), H(c + 1));
    Tan aR = ppm_edge_tan(t[c], t[c + 1], s0, sp, nu, H(c - 1), H(c), H(c + 1), H(c + 2));
    ppm_limit_tan(t[c], aL, aR);
    return M >= 0.0 ? aR.d : aL.d;
}

// Local gradient g[d] = d(face value)/d v[d] of the active branch (reverse rules).
inline void ppm_slope_grad(const double* v, const double* h, int x, double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    const double am = v[x - 1], a0 = v[x], ap = v[x + 1];
    const double dl = a0 - am, dr = ap - a0;
    if (!(dr * dl > 0.0)) return;
    double da, cR = 0.5, cL = 0.5;
    if (h) {
        const double w = h[x] / (h[x - 1] + h[x] + h[x + 1]);
        cR = w * ((2.0 * h[x - 1] + h[x]) / (h[x + 1] + h[x]));
        cL = w * ((h[x] + 2.0 * h[x + 1]) / (h[x - 1] + h[x]));
        da = cR * dr + cL * dl;
    } else {
        da = 0.5 * (ap - am);
    }
    const double ax = std::fabs(da), y = 2.0 * std::fabs(dl), z = 2.0 * std::fabs(dr);
    const double in = (z < y) ? z : y;
    if (in < ax) {
        if (z < y) { g[2] = 2.0; g[1] = -2.0; }
        else { g[1] = 2.0; g[0] = -2.0; }
        return;
    }
    // da = cR (ap - a0) + cL (a0 - am)
    g[2] = cR;
    g[1] = cL - cR;
    g[0] = -cL;
}

// Edge between cells x and x+1: gradient with respect to (a_x, a_x+1, s_x, s_x+1).
inline void ppm_edge_grad(const double* h, int x, double e[4]) {
    if (!h) {
        e[0] = 0.5; e[1] = 0.5; e[2] = C16; e[3] = -C16;
        return;
    }
    const double hm = h[x - 1], h0 = h[x], h1 = h[x + 1], h2 = h[x + 2];
    const double cA = h0 / (h0 + h1);
    const double gg = (2.0 * h1 * h0) / (h0 + h1);
    const double t1 = (hm + h0) / (2.0 * h0 + h1);
    const double t2 = (h2 + h1) / (2.0 * h1 + h0);
    const double cS = 1.0 / (((hm + h0) + h1) + h2);
    const double cE = cA + (gg * (t1 - t2)) * cS;
    const double cF = (h0 * t1) * cS;
    const double cG = (h1 * ((h1 + h2) / (h0 + 2.0 * h1))) * cS;
    e[0] = 1.0 - cE; e[1] = cE; e[2] = cG; e[3] = -cF;
}

inline void ppm_face_grad(const double* v, const double* h, double M, double g[6]) {
    for (int d = 0; d < 6; ++d) g[d] = 0.0;
    const int c = M >= 0.0 ? 2 : 3;
    double sm[3], s0[3], sp[3], eL[4], eR[4];
    ppm_slope_grad(v, h, c - 1, sm);
    ppm_slope_grad(v, h, c, s0);
    ppm_slope_grad(v, h, c + 1, sp);
    auto slope_val = [&](int x) {
        return h ? ppm_slope_nu(h[x - 1], h[x], h[x + 1], v[x - 1], v[x], v[x + 1])
                 : ppm_slope(v[x - 1], v[x], v[x + 1]);
    };
    auto edge_val = [&](int x, double d0, double d1) {
        return h ? ppm_edge_nu(h[x - 1], h[x], h[x + 1], h[x + 2], v[x], v[x + 1], d0, d1)
                 : ppm_edge(v[x], v[x + 1], d0, d1);
    };
    const double dsm = slope_val(c - 1), ds0 = slope_val(c), dsp = slope_val(c + 1);
    const double aL = edge_val(c - 1, dsm, ds0), aR = edge_val(c, ds0, dsp);
    ppm_edge_grad(h, c - 1, eL);
    ppm_edge_grad(h, c, eR);
    // adjoints of (a_c, aL, aR) from the limiter branch
    double bc = 0.0, bL = 0.0, bR = 0.0;
    const bool right = M >= 0.0;
    if ((aR - v[c]) * (v[c] - aL) <= 0.0) {
        bc = 1.0;
    } else {
        const double d = aR - aL;
        const double a6 = 6.0 * (v[c] - 0.5 * (aL + aR));
        const double dd = d * d;
        const double da6 = d * a6;
        if (da6 > dd) {             // aL <- 3 a_c - 2 aR
            if (right) bR = 1.0;
            else { bc = 3.0; bR = -2.0; }
        } else if (-dd > da6) {     // aR <- 3 a_c - 2 aL
            if (right) { bc = 3.0; bL = -2.0; }
            else bL = 1.0;
        } else {
            if (right) bR = 1.0;
            else bL = 1.0;
        }
    }
    g[c] += bc;
    // aL = edge(c-1): (a_c-1, a_c, s_c-1, s_c); aR = edge(c): (a_c, a_c+1, s_c, s_c+1)
    g[c - 1] += bL * eL[0];
    g[c] += bL * eL[1];
    const double bsm = bL * eL[2];
    double bs0 = bL * eL[3];
    g[c] += bR * eR[0];
    g[c + 1] += bR * eR[1];
    bs0 += bR * eR[2];
    const double bsp = bR * eR[3];
    for (int q = 0; q < 3; ++q) {
        g[c - 2 + q] += bsm * sm[q];
        g[c - 1 + q] += bs0 * s0[q];
        g[c + q] += bsp * sp[q];
    }
}


// The PBL mixing term is linear in the mixing ratios (pbl_f fixed):
//   T_k = pbl_f_k (sum_j w_j q_j / W' - q_k) / tau,  w = dp pbl_f,
//   W' = max(sum_j w_j, 1), per column.
// Its tangent, added to out (mixing ratios only): out_k += T_k(v).
static void ppm_pbl_jvp(const rs_ctx& c, const double* v, double* out) {
    if (c.pblf.empty()) return;
    const size_t nc = c.nc, plane = (size_t)c.nlon * c.nlat;
    for (int s = 1; s < c.nt; ++s) {
        const double* vs = v + (size_t)s * nc;
        double* os = out + (size_t)s * nc;
        for (size_t col = 0; col < plane; ++col) {
            double V = 0.0;
            for (int k = 0; k < c.nlev; ++k) V = V + c.pblw[(size_t)k * plane + col] * vs[(size_t)k * plane + col];
            const double vm = V / c.pblWp[col];
            for (int k = 0; k < c.nlev; ++k) {
                const size_t x = (size_t)k * plane + col;
                os[x] += (c.pblf[x] * (vm - vs[x])) / c.pbl_tau;
            }
        }
    }
}

// Its transpose, added to out: out_j += w_j L / (W' tau) - pbl_f_j lam_j / tau,
// L = sum_k pbl_f_k lam_k over the column.
static void ppm_pbl_vjp(const rs_ctx& c, const double* lam, double* out) {
    if (c.pblf.empty()) return;
    const size_t nc = c.nc, plane = (size_t)c.nlon * c.nlat;
    for (int s = 1; s < c.nt; ++s) {
        const double* ls = lam + (size_t)s * nc;
        double* os = out + (size_t)s * nc;
        for (size_t col = 0; col < plane; ++col) {
            double L = 0.0;
            for (int k = 0; k < c.nlev; ++k) L = L + c.pblf[(size_t)k * plane + col] * ls[(size_t)k * plane + col];
            const double g = L / (c.pblWp[col] * c.pbl_tau);
            for (int k = 0; k < c.nlev; ++k) {
                const size_t x = (size_t)k * plane + col;
                os[x] += c.pblw[x] * g - (c.pblf[x] * ls[x]) / c.pbl_tau;
            }
        }
    }
}

// out = J(u) v over the transported block (tangent of ppm_rhs_all).
void ppm_rhs_jvp_all(const rs_ctx& c, const double* u, const double* vt, double* out) {
    const int nlon = c.nlon, nlat = c.nlat, nlev = c.nlev, nt = c.nt;
    const size_t nc = c.nc, plane = (size_t)nlon * nlat;
    std::vector<double> f((size_t)nt * nc);
    ppm_rhs_all(c, u, f.data(), false);
    const double* m = u;
    const double* vm = vt;
    for (size_t x = 0; x < nc; ++x) out[x] = 0.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nlev; ++k) {
        for (int j = 0; j < nlat; ++j) {
            const double gx = c.dphi_lat[j] / (c.dlon * c.dS_lat[j]);
            const double ces = c.coslat_e[j], cen = c.coslat_e[j + 1], rdS = 1.0 / c.dS_lat[j];
            for (int i = 0; i < nlon; ++i) {
                const size_t cell = (size_t)k * plane + (size_t)j * nlon + i;
                const double* mxr = c.Mx.data() + ((size_t)k * nlat + j) * (nlon + 1);
                const double mw = mxr[i], me = mxr[i + 1];
                const double ms = c.My[((size_t)k * (nlat + 1) + j) * nlon + i];
                const double mn = c.My[((size_t)k * (nlat + 1) + j + 1) * nlon + i];
                const double mb = c.Mz[(size_t)k * plane + (size_t)j * nlon + i];
                const double mt = c.Mz[(size_t)(k + 1) * plane + (size_t)j * nlon + i];
                const double divM = ppm_div(mw, me, ms, mn, mb, mt, gx, ces, cen, rdS);
                for (int s = 1; s < nt; ++s) {
                    const double* q = u + (size_t)s * nc;
                    const double* qd = vt + (size_t)s * nc;
                    const int sq = s - 1;
                    // value and tangent of the lon / lat lines (halo cells: tangent 0)
                    auto X = [&](int ii, double* val, double* tan) {
                        if (c.lon_periodic) {
                            ii %= nlon;
                            if (ii < 0) ii += nlon;
                        } else if (ii < 0 || ii >= nlon) {
                            *val = ii < 0 ? c.qw[((size_t)sq * nlev + k) * nlat + j]
                                          : c.qe[((size_t)sq * nlev + k) * nlat + j];
                            *tan = 0.0;
                            return;
                        }
                        const size_t x = (size_t)k * plane + (size_t)j * nlon + ii;
                        *val = q[x];
                        *tan = qd[x];
                    };
                    auto Y = [&](int jj, double* val, double* tan) {
                        if (jj < 0 || jj >= nlat) {
                            *val = jj < 0 ? c.qs[((size_t)sq * nlev + k) * nlon + i]
                                          : c.qn[((size_t)sq * nlev + k) * nlon + i];
                            *tan = 0.0;
                            return;
                        }
                        const size_t x = (size_t)k * plane + (size_t)jj * nlon + i;
                        *val = q[x];
                        *tan = qd[x];
                    };
                    auto Zc = [&](int kk) { return std::min(std::max(kk, 0), nlev - 1); };
                    double v[6], vd[6], h[6];
                    for (int d = 0; d < 6; ++d) X(i - 3 + d, &v[d], &vd[d]);
                    const double tw = ppm_face_tan(v, vd, nullptr, mw);
                    for (int d = 0; d < 6; ++d) X(i - 2 + d, &v[d], &vd[d]);
                    const double te = ppm_face_tan(v, vd, nullptr, me);
                    for (int d = 0; d < 6; ++d) Y(j - 3 + d, &v[d], &vd[d]);
                    const double ts = ppm_face_tan(v, vd, nullptr, ms);
                    for (int d = 0; d < 6; ++d) Y(j - 2 + d, &v[d], &vd[d]);
                    const double tn = ppm_face_tan(v, vd, nullptr, mn);
                    for (int d = 0; d < 6; ++d) {
                        const size_t x = (size_t)Zc(k - 3 + d) * plane + (size_t)j * nlon + i;
                        v[d] = q[x]; vd[d] = qd[x]; h[d] = c.dp[x];
                    }
                    const double tb = ppm_face_tan(v, vd, h, mb);
                    for (int d = 0; d < 6; ++d) {
                        const size_t x = (size_t)Zc(k - 2 + d) * plane + (size_t)j * nlon + i;
                        v[d] = q[x]; vd[d] = qd[x]; h[d] = c.dp[x];
                    }
                    const double tt = ppm_face_tan(v, vd, h, mt);
                    const double ddivF = ppm_div(mw * tw, me * te, ms * ts, mn * tn, mb * tb,
                                                 mt * tt, gx, ces, cen, rdS);
                    const double dnum = qd[cell] * divM - ddivF;
                    out[(size_t)s * nc + cell] = (dnum - f[(size_t)s * nc + cell] * vm[cell]) / m[cell];
                }
            }
        }
    }
    ppm_pbl_jvp(c, vt, out);
}

// out = J(u)^T lam over the transported block: the air mass gathers
// -lam_s f_s / m over the mixing ratios; each mixing ratio gets
// mu divM (mu = lam / m) and, from every face of every cell, the face's local
// gradient times that cell's coefficient of the face flux, scattered to the
// face's stencil cells (halo cells drop out; clamped vertical indices
// accumulate on the clamped cell).  Tracers run in parallel, each serially.
void ppm_rhs_vjp_all(const rs_ctx& c, const double* u, const double* lam, double* out) {
    const int nlon = c.nlon, nlat = c.nlat, nlev = c.nlev, nt = c.nt;
    const size_t nc = c.nc, plane = (size_t)nlon * nlat;
    std::vector<double> f((size_t)nt * nc);
    ppm_rhs_all(c, u, f.data(), false);
    const double* m = u;
    for (size_t x = 0; x < nc; ++x) {
        double a = 0.0;
        for (int s = 1; s < nt; ++s) {
            const double t = lam[(size_t)s * nc + x] * f[(size_t)s * nc + x];
            a -= t / m[x];
        }
        out[x] = a;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 1; s < nt; ++s) {
        const double* q = u + (size_t)s * nc;
        const double* L = lam + (size_t)s * nc;
        double* o = out + (size_t)s * nc;
        for (size_t x = 0; x < nc; ++x) o[x] = 0.0;
        for (int k = 0; k < nlev; ++k) {
            for (int j = 0; j < nlat; ++j) {
                const double gx = c.dphi_lat[j] / (c.dlon * c.dS_lat[j]);
                const double ces = c.coslat_e[j], cen = c.coslat_e[j + 1], rdS = 1.0 / c.dS_lat[j];
                for (int i = 0; i < nlon; ++i) {
                    const size_t cell = (size_t)k * plane + (size_t)j * nlon + i;
                    const double* mxr = c.Mx.data() + ((size_t)k * nlat + j) * (nlon + 1);
                    const double mw = mxr[i], me = mxr[i + 1];
                    const double ms = c.My[((size_t)k * (nlat + 1) + j) * nlon + i];
                    const double mn = c.My[((size_t)k * (nlat + 1) + j + 1) * nlon + i];
                    const double mb = c.Mz[(size_t)k * plane + (size_t)j * nlon + i];
                    const double mt = c.Mz[(size_t)(k + 1) * plane + (size_t)j * nlon + i];
                    const double divM = ppm_div(mw, me, ms, mn, mb, mt, gx, ces, cen, rdS);
                    const double mu = L[cell] / m[cell];
                    o[cell] += mu * divM;
                    double v[6], h[6], g[6];
                    long long id[6];
                    // f_cell = (q divM - divF) / m: coefficient of each face's value
                    auto scatter = [&](double coef) {
                        for (int d = 0; d < 6; ++d)
                            if (id[d] >= 0) o[id[d]] += coef * g[d];
                    };
                    auto line_x = [&](int i0) {
                        for (int d = 0; d < 6; ++d) {
                            int ii = i0 + d;
                            if (c.lon_periodic) {
                                ii %= nlon;
                                if (ii < 0) ii += nlon;
                            } else if (ii < 0 || ii >= nlon) {
                                id[d] = -1;
                                v[d] = ii < 0 ? c.qw[((size_t)(s - 1) * nlev + k) * nlat + j]
                                              : c.qe[((size_t)(s - 1) * nlev + k) * nlat + j];
                                continue;
                            }
                            id[d] = (long long)((size_t)k * plane + (size_t)j * nlon + ii);
                            v[d] = q[id[d]];
                        }
                    };
                    auto line_y = [&](int j0) {
                        for (int d = 0; d < 6; ++d) {
                            const int jj = j0 + d;
                            if (jj < 0 || jj >= nlat) {
                                id[d] = -1;
                                v[d] = jj < 0 ? c.qs[((size_t)(s - 1) * nlev + k) * nlon + i]
                                              : c.qn[((size_t)(s - 1) * nlev + k) * nlon + i];
                                continue;
                            }
                            id[d] = (long long)((size_t)k * plane + (size_t)jj * nlon + i);
                            v[d] = q[id[d]];
                        }
                    };
                    auto line_z = [&](int k0) {
                        for (int d = 0; d < 6; ++d) {
                            const int kk = std::min(std::max(k0 + d, 0), nlev - 1);
                            id[d] = (long long)((size_t)kk * plane + (size_t)j * nlon + i);
                            v[d] = q[id[d]];
                            h[d] = c.dp[id[d]];
                        }
                    };
                    line_x(i - 3); ppm_face_grad(v, nullptr, mw, g); scatter(mu * (mw * gx));
                    line_x(i - 2); ppm_face_grad(v, nullptr, me, g); scatter(-mu * (me * gx));
                    line_y(j - 3); ppm_face_grad(v, nullptr, ms, g); scatter(mu * (ces * ms * rdS));
                    line_y(j - 2); ppm_face_grad(v, nullptr, mn, g); scatter(-mu * (cen * mn * rdS));
                    line_z(k - 3); ppm_face_grad(v, h, mb, g); scatter(mu * mb);
                    line_z(k - 2); ppm_face_grad(v, h, mt, g); scatter(-mu * mt);
                }
            }
        }
    }
    ppm_pbl_vjp(c, lam, out);
}

// f = L u over the transported block, for either scheme.
void transport_rhs_all(const rs_ctx& c, const double* u, double* f) {
    if (c.scheme == RS_TRANSPORT_PPM) {
        ppm_rhs_all(c, u, f, true);
        return;
    }
    for (int s = 0; s < c.nt; ++s)
        transport_rhs_block(c, u + (size_t)s * c.nc, f + (size_t)s * c.nc);
}

// One SSPRK43 attempt (ssprk43_step_unrolled, rx_traced_integrator.jl:404-437).
// Writes the raw result into unew (all species; non-transported copied), returns EEst.
// Stages run over the whole transported block (the PPM tendency couples the
// mixing ratios to the air mass); every element's arithmetic is per element.
double ssprk43_attempt(const rs_ctx& c, const std::vector<double>& u, std::vector<double>& unew,
                       double dt, double atol, double rtol) {
    const size_t nc = c.nc, nT = (size_t)c.nt * nc;
    std::vector<double> k(nT), u1(nT), u2(nT), u3(nT), u4(nT);
    unew.assign(u.begin(), u.end());
    const double dt_2 = dt / 2;
    std::vector<double> sq(nT);
    const double* us = u.data();
    transport_rhs_all(c, us, k.data());
    for (size_t x = 0; x < nT; ++x) { double du = dt_2 * k[x]; u1[x] = us[x] + du; }
    transport_rhs_all(c, u1.data(), k.data());
    for (size_t x = 0; x < nT; ++x) { double du = dt_2 * k[x]; u2[x] = u1[x] + du; }
    transport_rhs_all(c, u2.data(), k.data());
    for (size_t x = 0; x < nT; ++x) { double du = dt_2 * k[x]; u3[x] = u2[x] + du; }
    for (size_t x = 0; x < nT; ++x) {
        double a = (2.0 / 3.0) * us[x];
        double b = (1.0 / 3.0) * u3[x];
        u4[x] = a + b;
    }
    transport_rhs_all(c, u4.data(), k.data());
    double* un = unew.data();
    for (size_t x = 0; x < nT; ++x) {
        double du = dt_2 * k[x];
        un[x] = u4[x] + du;
        double a = (1.0 / 3.0) * us[x];
        double b = (2.0 / 3.0) * u3[x];
        double ut = a + b;
        double ud = ut - un[x];
        ut = 0.5 * ud;
        double am = std::max(std::fabs(us[x]), std::fabs(un[x]));
        double den = am * rtol;
        den = den + atol;
        double at = ut / den;
        sq[x] = at * at;
    }
    double sse = fixed_sum(sq);
    double N = (double)c.ns * (double)nc;
    return std::sqrt(sse / N);
}

// ----------------------------------------------------------------------------
// Chemistry (mass action), one cell at a time, dense NS x NS blocks.
// ----------------------------------------------------------------------------
struct CellChem {
    const rs_ctx& c;
    size_t cell;
    double jf;        // photolysis factor
    const double* E;  // emission of this cell per species (stride ncol) or null
    size_t estride;
    CellChem(const rs_ctx& cc, size_t cell_) : c(cc), cell(cell_) {
        jf = cc.jfac.empty() ? 1.0 : cc.jfac[cell_];
        E = nullptr; estride = cc.ncol;
        if (!cc.emis.empty() && cell_ < cc.ncol) E = cc.emis.data() + cell_;
    }
    double kc(int r) const { return c.photo[r] ? c.kr[r] * jf : c.kr[r]; }
    double rate(int r, const double* u) const {
        double k = kc(r);
        if (c.rb[r] < 0) return k * u[c.ra[r]];
        return (k * u[c.ra[r]]) * u[c.rb[r]];
    }
    // f = sum_r nu_r rate_r (+ E on level-0 cells), accumulation in reaction order.
    void rhs(const double* u, double* f) const {
        const int ns = c.ns;
        for (int i = 0; i < ns; ++i) f[i] = E ? E[(size_t)i * estride] : 0.0;
        for (int r = 0; r < c.nr; ++r) {
            double rt = rate(r, u);
            for (auto& pr : c.net[r]) f[pr.first] += pr.second * rt;
        }
    }
    // dense analytic Jacobian J[i*ns + a]
    void jac(const double* u, double* J) const {
        const int ns = c.ns;
        std::fill(J, J + (size_t)ns * ns, 0.0);
        for (int r = 0; r < c.nr; ++r) {
            double k = kc(r);
            int a = c.ra[r], b = c.rb[r];
            double da, db = 0.0;
            if (b < 0) da = k;
            else if (a == b) da = (2.0 * k) * u[a];
            else { da = k * u[b]; db = k * u[a]; }
            for (auto& pr : c.net[r]) {
                double nu = pr.second;
                J[(size_t)pr.first * ns + a] += nu * da;
                if (b >= 0 && b != a) J[(size_t)pr.first * ns + b] += nu * db;
            }
        }
    }
};

// blocksolve (rx_traced_integrator.jl:110-135): pivot-free Gaussian
// elimination of W x = b, W row-major NS x NS.  W and b are copied.
void blocksolve(int ns, const double* Win, const double* bin, double* x) {
    thread_local std::vector<double> W, b;   // per-thread copies, reused across calls
    W.assign(Win, Win + (size_t)ns * ns);
    b.assign(bin, bin + ns);
    for (int k = 0; k < ns; ++k) {
        double piv = 1.0 / W[(size_t)k * ns + k];
        for (int i = k + 1; i < ns; ++i) {
            double m = W[(size_t)i * ns + k] * piv;
            for (int j = k + 1; j < ns; ++j) {
                double mw = m * W[(size_t)k * ns + j];
                W[(size_t)i * ns + j] = W[(size_t)i * ns + j] - mw;
            }
            double mb = m * b[k];
            b[i] = b[i] - mb;
        }
    }
    for (int i = ns - 1; i >= 0; --i) {
        double acc = b[i];
        for (int j = i + 1; j < ns; ++j) {
            double wx = W[(size_t)i * ns + j] * x[j];
            acc = acc - wx;
        }
        x[i] = acc / W[(size_t)i * ns + i];
    }
}

// W^T x = b by the same elimination applied to the transpose.
void blocksolve_T(int ns, const double* W, const double* b, double* x) {
    std::vector<double> WT((size_t)ns * ns);
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) WT[(size_t)j * ns + i] = W[(size_t)i * ns + j];
    blocksolve(ns, WT.data(), b, x);
}

struct Ros23Work {
    std::vector<double> f0, f1, f2, k1, k2, k3, u1, J, W, b;
    explicit Ros23Work(int ns)
        : f0(ns), f1(ns), f2(ns), k1(ns), k2(ns), k3(ns), u1(ns),
          J((size_t)ns * ns), W((size_t)ns * ns), b(ns) {}
};

// Build W = I - dt*d*J (rx_traced_integrator.jl:291-296).
void build_W(int ns, double dt, const double* J, double* W) {
    double dtgamma = dt * ROS23_D;
    for (int r = 0; r < ns; ++r)
        for (int s = 0; s < ns; ++s) {
            double tmp = dtgamma * J[(size_t)r * ns + s];
            W[(size_t)r * ns + s] = (r == s) ? 1.0 - tmp : 0.0 - tmp;
        }
}

// One ROS23 attempt on one cell (ros23_step, rx_traced_integrator.jl:279-349).
// u, unew: NS values; returns the cell's sum of squared scaled residuals.
double ros23_cell(const CellChem& cc, const double* u, double* unew, double dt,
                  double atol, double rtol, Ros23Work& w) {
    const int ns = cc.c.ns;
    cc.rhs(u, w.f0.data());                                           // :282
    cc.jac(u, w.J.data());                                            // :287
    build_W(ns, dt, w.J.data(), w.W.data());                          // :291-296
    blocksolve(ns, w.W.data(), w.f0.data(), w.k1.data());             // :298
    double dto2 = dt / 2;
    for (int i = 0; i < ns; ++i) { double du1 = dto2 * w.k1[i]; w.u1[i] = u[i] + du1; }
    cc.rhs(w.u1.data(), w.f1.data());                                 // :303
    for (int i = 0; i < ns; ++i) w.b[i] = w.f1[i] - w.k1[i];          // :306-309
    blocksolve(ns, w.W.data(), w.b.data(), w.k2.data());              // :310
    for (int i = 0; i < ns; ++i) w.k2[i] = w.k2[i] + w.k1[i];         // :311-313
    for (int i = 0; i < ns; ++i) { double du2 = dt * w.k2[i]; unew[i] = u[i] + du2; }
    cc.rhs(unew, w.f2.data());                                        // :318
    for (int i = 0; i < ns; ++i) {                                    // :319-327
        double a = w.k2[i] - w.f1[i];
        a = ROS23_C32 * a;
        double c = w.k1[i] - w.f0[i];
        c = 2.0 * c;
        double d3 = w.f2[i] - a;
        w.b[i] = d3 - c;
    }
    blocksolve(ns, w.W.data(), w.b.data(), w.k3.data());              // :328
    double dto6 = dt / 6, sse = 0.0;
    for (int i = 0; i < ns; ++i) {                                    // :330-346
        double w2 = 2.0 * w.k2[i];
        double ut = w.k1[i] - w2;
        ut = ut + w.k3[i];
        ut = dto6 * ut;
        double am = std::max(std::fabs(u[i]), std::fabs(unew[i]));
        double den = am * rtol;
        den = den + atol;
        double at = ut / den;
        sse += at * at;
    }
    return sse;
}

// Gather / scatter one cell of a species-major vector.
inline void gather(const rs_ctx& c, const double* v, size_t cell, double* out) {
    for (int s = 0; s < c.ns; ++s) out[s] = v[(size_t)s * c.nc + cell];
}
inline void scatter(const rs_ctx& c, const double* in, size_t cell, double* v) {
    for (int s = 0; s < c.ns; ++s) v[(size_t)s * c.nc + cell] = in[s];
}

double ros23_attempt(const rs_ctx& c, const std::vector<double>& u, std::vector<double>& unew,
                     double dt, double atol, double rtol) {
    unew.resize(u.size());
    std::vector<double> sq(c.nc);
#pragma omp parallel
    {
        Ros23Work w(c.ns);
        std::vector<double> uc(c.ns), un(c.ns);
#pragma omp for schedule(static)
        for (long long cell = 0; cell < (long long)c.nc; ++cell) {
            CellChem cc(c, (size_t)cell);
            gather(c, u.data(), cell, uc.data());
            sq[cell] = ros23_cell(cc, uc.data(), un.data(), dt, atol, rtol, w);
            scatter(c, un.data(), cell, unew.data());
        }
    }
    double sse = fixed_sum(sq);
    return std::sqrt(sse / ((double)c.ns * (double)c.nc));
}

// ----------------------------------------------------------------------------
// One-step adjoints and tangents.
// ----------------------------------------------------------------------------
// SSPRK43 at fixed dt, transport linear: unew = P u with
// P = (2/3 I + 1/3 S^3)(I + h/2 L)... reversed stage by stage.
void ssprk43_vjp(const rs_ctx& c, std::vector<double>& lam, double dt) {
    const size_t nc = c.nc;
    const double h2 = dt / 2;
    std::vector<double> g(nc), l4(nc), l3(nc), l2(nc), l1(nc), l0(nc), ub(nc);
    auto apply = [&](const double* in, double* out) {    // out = in + h/2 L^T in
        std::fill(g.begin(), g.end(), 0.0);
        transport_rhs_adj_block(c, in, g.data());
        for (size_t x = 0; x < nc; ++x) out[x] = in[x] + h2 * g[x];
    };
    for (int s = 0; s < c.nt; ++s) {
        double* L = lam.data() + s * nc;
        apply(L, l4.data());                                        // unew = u4 + h/2 L u4
        for (size_t x = 0; x < nc; ++x) { ub[x] = (2.0 / 3.0) * l4[x]; l3[x] = (1.0 / 3.0) * l4[x]; }
        apply(l3.data(), l2.data());                                // u3 = u2 + h/2 L u2
        apply(l2.data(), l1.data());                                // u2 = u1 + h/2 L u1
        apply(l1.data(), l0.data());                                // u1 = u + h/2 L u
        for (size_t x = 0; x < nc; ++x) L[x] = ub[x] + l0[x];
    }
}

// SSPRK43 at fixed dt with the PPM tendency, which is nonlinear (limiters,
// division by the air mass): the stages u1, u2, u4 are recomputed from the
// step input u as in ssprk43_attempt, then reversed with J at each stage state:
//   l4 = lam + h/2 J(u4)^T lam;  ub = 2/3 l4;  l3 = 1/3 l4
//   l2 = l3 + h/2 J(u2)^T l3;  l1 = l2 + h/2 J(u1)^T l2;  l0 = l1 + h/2 J(u)^T l1
//   lam_u = ub + l0
// Species beyond the transported block pass through.
void ssprk43_ppm_stage_states(const rs_ctx& c, const double* u, double dt, std::vector<double>& u1,
                              std::vector<double>& u2, std::vector<double>& u4) {
    const size_t nT = (size_t)c.nt * c.nc;
    const double dt_2 = dt / 2;
    std::vector<double> k(nT), u3(nT);
    u1.resize(nT); u2.resize(nT); u4.resize(nT);
    transport_rhs_all(c, u, k.data());
    for (size_t x = 0; x < nT; ++x) { double du = dt_2 * k[x]; u1[x] = u[x] + du; }
    transport_rhs_all(c, u1.data(), k.data());
    for (size_t x = 0; x < nT; ++x) { double du = dt_2 * k[x]; u2[x] = u1[x] + du; }
    transport_rhs_all(c, u2.data(), k.data());
    for (size_t x = 0; x < nT; ++x) { double du = dt_2 * k[x]; u3[x] = u2[x] + du; }
    for (size_t x = 0; x < nT; ++x) {
        double a = (2.0 / 3.0) * u[x];
        double b = (1.0 / 3.0) * u3[x];
        u4[x] = a + b;
    }
}

void ssprk43_vjp_ppm(const rs_ctx& c, const double* u, std::vector<double>& lam, double dt) {
    const size_t nT = (size_t)c.nt * c.nc;
    const double h2 = dt / 2;
    std::vector<double> u1, u2, u4, g(nT), l4(nT), l3(nT), ub(nT);
    ssprk43_ppm_stage_states(c, u, dt, u1, u2, u4);
    double* L = lam.data();
    auto apply = [&](const double* at, const double* in, double* out) {   // out = in + h/2 J(at)^T in
        ppm_rhs_vjp_all(c, at, in, g.data());
        for (size_t x = 0; x < nT; ++x) out[x] = in[x] + h2 * g[x];
    };
    apply(u4.data(), L, l4.data());
    for (size_t x = 0; x < nT; ++x) { ub[x] = (2.0 / 3.0) * l4[x]; l3[x] = (1.0 / 3.0) * l4[x]; }
    apply(u2.data(), l3.data(), l4.data());     // l2
    apply(u1.data(), l4.data(), l3.data());     // l1
    apply(u, l3.data(), l4.data());             // l0
    for (size_t x = 0; x < nT; ++x) L[x] = ub[x] + l4[x];
}

// Tangent of the same step: v1 = v + h/2 J(u) v, v2 = v1 + h/2 J(u1) v1,
// v3 = v2 + h/2 J(u2) v2, v4 = 2/3 v + 1/3 v3, out = v4 + h/2 J(u4) v4.
void ssprk43_jvp_ppm(const rs_ctx& c, const double* u, const double* v, double* out, double dt) {
    const size_t nT = (size_t)c.nt * c.nc;
    const double h2 = dt / 2;
    std::vector<double> u1, u2, u4, g(nT), v1(nT), v2(nT), v4(nT);
    ssprk43_ppm_stage_states(c, u, dt, u1, u2, u4);
    ppm_rhs_jvp_all(c, u, v, g.data());
    for (size_t x = 0; x < nT; ++x) v1[x] = v[x] + h2 * g[x];
    ppm_rhs_jvp_all(c, u1.data(), v1.data(), g.data());
    for (size_t x = 0; x < nT; ++x) v2[x] = v1[x] + h2 * g[x];
    ppm_rhs_jvp_all(c, u2.data(), v2.data(), g.data());
    for (size_t x = 0; x < nT; ++x) {
        const double v3 = v2[x] + h2 * g[x];
        v4[x] = (2.0 / 3.0) * v[x] + (1.0 / 3.0) * v3;
    }
    ppm_rhs_jvp_all(c, u4.data(), v4.data(), g.data());
    for (size_t x = 0; x < nT; ++x) out[x] = v4[x] + h2 * g[x];
    for (size_t x = nT; x < (size_t)c.ns * c.nc; ++x) out[x] = v[x];
}

void ssprk43_jvp(const rs_ctx& c, const std::vector<double>& v, std::vector<double>& out, double dt) {
    // The transport step is linear in u: its JVP is the step applied to v.
    std::vector<double> tmp;
    ssprk43_attempt(c, v, tmp, dt, 1.0, 0.0);
    out = tmp;
}

// Products of the analytic Jacobian with vectors and its u-derivative terms.
void jac_T_vec(const rs_ctx& c, const CellChem& cc, const double* u, const double* z, double* out) {
    // out += J(u)^T z
    std::vector<double> J((size_t)c.ns * c.ns);
    cc.jac(u, J.data());
    for (int a = 0; a < c.ns; ++a) {
        double s = 0.0;
        for (int i = 0; i < c.ns; ++i) s += J[(size_t)i * c.ns + a] * z[i];
        out[a] += s;
    }
}

// ubar += coef * d/du [ z^T J(u) v ]  (only bimolecular reactions contribute)
void hess_term(const rs_ctx& c, const CellChem& cc, const double* z, const double* v,
               double coef, double* ubar) {
    for (int r = 0; r < c.nr; ++r) {
        int a = c.ra[r], b = c.rb[r];
        if (b < 0) continue;
        double sr = 0.0;
        for (auto& pr : c.net[r]) sr += (double)pr.second * z[pr.first];
        double k = cc.kc(r);
        if (a == b) ubar[a] += coef * (2.0 * k) * v[a] * sr;
        else {
            ubar[a] += coef * k * v[b] * sr;
            ubar[b] += coef * k * v[a] * sr;
        }
    }
}

// Reverse of the part of ros23_step that reaches unew (f2, k3, EEst are dead):
//   f0=f(u); k1=W\f0; u1=u+h/2 k1; f1=f(u1); k2=W\(f1-k1)+k1; unew=u+h k2.
void ros23_vjp_cell(const CellChem& cc, const double* u, double dt, const double* lam_in,
                    double* lam_out, double* ebar /*ns, surface only, may be null*/) {
    const rs_ctx& c = cc.c;
    const int ns = c.ns;
    Ros23Work w(ns);
    cc.rhs(u, w.f0.data());
    cc.jac(u, w.J.data());
    build_W(ns, dt, w.J.data(), w.W.data());
    blocksolve(ns, w.W.data(), w.f0.data(), w.k1.data());
    double dto2 = dt / 2;
    for (int i = 0; i < ns; ++i) w.u1[i] = u[i] + dto2 * w.k1[i];
    cc.rhs(w.u1.data(), w.f1.data());
    for (int i = 0; i < ns; ++i) w.b[i] = w.f1[i] - w.k1[i];
    blocksolve(ns, w.W.data(), w.b.data(), w.k2.data());
    for (int i = 0; i < ns; ++i) w.k2[i] = w.k2[i] + w.k1[i];
    const double hg = dt * ROS23_D;
    std::vector<double> ub(lam_in, lam_in + ns), k2b(ns), z2(ns), k1b(ns), u1b(ns, 0.0), z1(ns), v(ns);
    for (int i = 0; i < ns; ++i) k2b[i] = dt * lam_in[i];
    blocksolve_T(ns, w.W.data(), k2b.data(), z2.data());
    for (int i = 0; i < ns; ++i) k1b[i] = k2b[i] - z2[i];
    jac_T_vec(c, cc, w.u1.data(), z2.data(), u1b.data());
    for (int i = 0; i < ns; ++i) { ub[i] += u1b[i]; k1b[i] += dto2 * u1b[i]; }
    blocksolve_T(ns, w.W.data(), k1b.data(), z1.data());
    jac_T_vec(c, cc, u, z1.data(), ub.data());
    for (int i = 0; i < ns; ++i) v[i] = w.k2[i] - w.k1[i];
    hess_term(c, cc, z2.data(), v.data(), hg, ub.data());
    hess_term(c, cc, z1.data(), w.k1.data(), hg, ub.data());
    for (int i = 0; i < ns; ++i) lam_out[i] = ub[i];
    if (ebar)
        for (int i = 0; i < ns; ++i) ebar[i] += z1[i] + z2[i];
}

// Forward-mode tangent of the same map, derived independently of the VJP.
void ros23_jvp_cell(const CellChem& cc, const double* u, double dt, const double* ud,
                    const double* Ed /*ns or null*/, double* outd) {
    const rs_ctx& c = cc.c;
    const int ns = c.ns;
    Ros23Work w(ns);
    cc.rhs(u, w.f0.data());
    cc.jac(u, w.J.data());
    build_W(ns, dt, w.J.data(), w.W.data());
    blocksolve(ns, w.W.data(), w.f0.data(), w.k1.data());
    double dto2 = dt / 2;
    for (int i = 0; i < ns; ++i) w.u1[i] = u[i] + dto2 * w.k1[i];
    cc.rhs(w.u1.data(), w.f1.data());
    std::vector<double> y(ns);
    for (int i = 0; i < ns; ++i) w.b[i] = w.f1[i] - w.k1[i];
    blocksolve(ns, w.W.data(), w.b.data(), y.data());
    for (int i = 0; i < ns; ++i) w.k2[i] = y[i] + w.k1[i];
    // Jdot = dJ/du[ud]
    std::vector<double> Jd((size_t)ns * ns, 0.0);
    for (int r = 0; r < c.nr; ++r) {
        int a = c.ra[r], b = c.rb[r];
        if (b < 0) continue;
        double k = cc.kc(r);
        for (auto& pr : c.net[r]) {
            double nu = (double)pr.second;
            if (a == b) Jd[(size_t)pr.first * ns + a] += nu * (2.0 * k) * ud[a];
            else {
                Jd[(size_t)pr.first * ns + a] += nu * k * ud[b];
                Jd[(size_t)pr.first * ns + b] += nu * k * ud[a];
            }
        }
    }
    const double hg = dt * ROS23_D;
    auto Wd_mul = [&](const double* x, double* out) {   // out = Wdot x = -hg Jd x
        for (int i = 0; i < ns; ++i) {
            double s = 0.0;
            for (int j = 0; j < ns; ++j) s += Jd[(size_t)i * ns + j] * x[j];
            out[i] = -hg * s;
        }
    };
    auto J_mul = [&](const double* uu, const double* x, double* out) {
        std::vector<double> JJ((size_t)ns * ns);
        cc.jac(uu, JJ.data());
        for (int i = 0; i < ns; ++i) {
            double s = 0.0;
            for (int j = 0; j < ns; ++j) s += JJ[(size_t)i * ns + j] * x[j];
            out[i] = s;
        }
    };
    std::vector<double> f0d(ns), tmp(ns), k1d(ns), u1d(ns), f1d(ns), yd(ns);
    J_mul(u, ud, f0d.data());
    if (Ed) for (int i = 0; i < ns; ++i) f0d[i] += Ed[i];
    Wd_mul(w.k1.data(), tmp.data());
    for (int i = 0; i < ns; ++i) tmp[i] = f0d[i] - tmp[i];
    blocksolve(ns, w.W.data(), tmp.data(), k1d.data());
    for (int i = 0; i < ns; ++i) u1d[i] = ud[i] + dto2 * k1d[i];
    J_mul(w.u1.data(), u1d.data(), f1d.data());
    if (Ed) for (int i = 0; i < ns; ++i) f1d[i] += Ed[i];
    Wd_mul(y.data(), tmp.data());
    for (int i = 0; i < ns; ++i) tmp[i] = f1d[i] - k1d[i] - tmp[i];
    blocksolve(ns, w.W.data(), tmp.data(), yd.data());
    for (int i = 0; i < ns; ++i) outd[i] = ud[i] + dt * (yd[i] + k1d[i]);
}

// Clamp flags: 1 = max(u, 0) on each accepted state (clamp_nonneg,
// rx_traced_integrator.jl:466-470); 2 = positive domain: a trial state with a
// negative entry is rejected and the step retried at half its size (the
// failfactor-2 rule of an out-of-domain step in the ODE solvers the reference
// configures with PositiveDomain, HELPERS.md:396-447), before the PI rule.
bool out_of_domain(const std::vector<double>& x, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (x[i] < 0.0) return true;
    return false;
}

void reject_out_of_domain(PIState& s, double dtc) {
    s.dt = dtc / 2.0;
    s.nrej += 1;
    s.iters += 1;
}

// ----------------------------------------------------------------------------
// Adaptive loops.
// ----------------------------------------------------------------------------
enum class Op { T, C };

int adaptive(rs_ctx* c, Op op, double t0, double t1, double* dt_io, const rs_pictrl& ctrl,
             double atol, double rtol, int clamp, long long maxiters, rs_counts* out,
             Seq* seq, std::vector<std::array<double, 4>>* alog = nullptr) {
    PIState s{t0, *dt_io, ctrl.qoldinit};
    if (alog) alog->clear();
    const double tlim = t1 - 1.0e-9;                              // :479
    std::vector<double> raw;
    while (s.t < tlim && s.iters < maxiters) {                    // :480
        double dtc = std::min(s.dt, t1 - s.t);                     // :482
        double ee = (op == Op::T) ? ssprk43_attempt(*c, c->u, raw, dtc, atol, rtol)
                                  : ros23_attempt(*c, c->u, raw, dtc, atol, rtol);
        double tbefore = s.t;
        // the operator's own entries: transported species, or the whole state
        const size_t nown = (op == Op::T) ? (size_t)c->nt * c->nc : raw.size();
        if ((clamp & 2) && out_of_domain(raw, nown)) {
            reject_out_of_domain(s, dtc);
            if (alog) alog->push_back({tbefore, dtc, ee, -1.0});
            continue;
        }
        const bool acc = pi_update(ctrl, s, dtc, ee);
        if (alog) alog->push_back({tbefore, dtc, ee, acc ? 1.0 : 0.0});
        if (acc) {
            if (clamp & 1) for (double& x : raw) x = std::max(x, 0.0);
            c->u.swap(raw);
            if (seq) seq->ts.emplace_back(tbefore, dtc);
        }
    }
    *dt_io = s.dt;
    if (out) { out->naccept = s.nacc; out->nreject = s.nrej; out->t_end = s.t; out->dt_end = s.dt; }
    if (s.iters >= maxiters && s.t < tlim)
        return fail(c, RS_E_MAXITERS, "adaptive loop hit maxiters at t=" + std::to_string(s.t));
    return RS_OK;
}

// One cell's adaptive ROS23 solve over [t0, t1] from uc (in place): the
// reference loop with NC = 1 (rx_traced_integrator.jl:455-507).  With rec,
// the accepted steps are recorded: rec->st gets the state before each
// accepted step and, last, the end state (ns values each), rec->dt the steps.
struct CellRec { std::vector<double> st, dt; };
PIState percell_cell(const CellChem& cc, double* uc, double* un, double t0, double t1, double dt0,
                     const rs_pictrl& ctrl, double atol, double rtol, int clamp,
                     long long maxiters, Ros23Work& w, CellRec* rec) {
    const int ns = cc.c.ns;
    PIState s{t0, dt0, ctrl.qoldinit};
    const double tlim = t1 - 1.0e-9;
    if (rec) { rec->st.assign(uc, uc + ns); rec->dt.clear(); }
    std::vector<double> unv(ns);
    while (s.t < tlim && s.iters < maxiters) {
        double dtc = std::min(s.dt, t1 - s.t);
        double sse = ros23_cell(cc, uc, un, dtc, atol, rtol, w);
        double ee = std::sqrt(sse / (double)ns);
        unv.assign(un, un + ns);
        if ((clamp & 2) && out_of_domain(unv, (size_t)ns)) {
            reject_out_of_domain(s, dtc);
            continue;
        }
        if (pi_update(ctrl, s, dtc, ee)) {
            for (int i = 0; i < ns; ++i) uc[i] = (clamp & 1) ? std::max(un[i], 0.0) : un[i];
            if (rec) { rec->dt.push_back(dtc); rec->st.insert(rec->st.end(), uc, uc + ns); }
        }
    }
    return s;
}

// Per-cell adaptive chemistry (the reference loop with NC = 1).
int percell(rs_ctx* c, double t0, double t1, double dt0, double* dt_cells, const rs_pictrl& ctrl,
            double atol, double rtol, int clamp, long long maxiters, rs_percell_stats* st) {
    const int ns = c->ns;
    long long att = 0, acc = 0, rej = 0, amax = 0, amin = -1, nfail = 0;
#pragma omp parallel reduction(+ : att, acc, rej, nfail) reduction(max : amax)
    {
        Ros23Work w(ns);
        std::vector<double> uc(ns), un(ns);
        long long lmin = -1;
#pragma omp for schedule(dynamic, 64)
        for (long long cell = 0; cell < (long long)c->nc; ++cell) {
            CellChem cc(*c, (size_t)cell);
            gather(*c, c->u.data(), cell, uc.data());
            PIState s = percell_cell(cc, uc.data(), un.data(), t0, t1,
                                     dt_cells ? dt_cells[cell] : dt0, ctrl, atol, rtol, clamp,
                                     maxiters, w, nullptr);
            if (s.iters >= maxiters && s.t < t1 - 1.0e-9) nfail += 1;
            scatter(*c, uc.data(), cell, c->u.data());
            if (dt_cells) dt_cells[cell] = s.dt;
            att += s.iters; acc += s.nacc; rej += s.nrej;
            amax = std::max(amax, s.iters);
            lmin = (lmin < 0) ? s.iters : std::min(lmin, s.iters);
        }
#pragma omp critical
        { if (lmin >= 0) amin = (amin < 0) ? lmin : std::min(amin, lmin); }
    }
    if (st) {
        st->attempts_total = att; st->accepts_total = acc; st->rejects_total = rej;
        st->attempts_max = amax; st->attempts_min = amin < 0 ? 0 : amin; st->n_failed = nfail;
    }
    if (nfail) return fail(c, RS_E_MAXITERS, std::to_string(nfail) + " cells hit maxiters");
    return RS_OK;
}

// Adjoint of the per-cell adaptive chemistry: the reference's backward sweep
// (adjoint_gradient.jl:731-834) over each cell's own accepted sequence
// (adaptive_solve with NC = 1, rx_traced_integrator.jl:455-507).  Each cell
// is integrated again from u0 with its accepted steps recorded (the same
// operations as percell, so the same trajectory), then swept backwards, last
// step first: lambda .*= 1[u_{j+1} > 0] when clamping (the accepted state is
// max(u, 0), :736), lambda <- VJP of step j at u_j.  c->lam: dJ/du(t1) in,
// dJ/du(t0) out; c->gemis accumulates dJ/dE.  *uend gets the end states,
// *nvjp the step VJPs taken, *ngrad the non-finite emission-gradient
// contributions.
int percell_vjp(rs_ctx* c, const std::vector<double>& u0, double t0, double t1, double dt0,
                const rs_pictrl& ctrl, double atol, double rtol, int clamp, long long maxiters,
                std::vector<double>* uend, long long* nvjp, long long* ngrad,
                rs_percell_stats* st) {
    const int ns = c->ns;
    if (uend) uend->assign(u0.size(), 0.0);
    long long att = 0, acc = 0, rej = 0, amax = 0, amin = -1, nfail = 0, nv = 0, ng = 0;
#pragma omp parallel reduction(+ : att, acc, rej, nfail, nv, ng) reduction(max : amax)
    {
        Ros23Work w(ns);
        CellRec rec;
        std::vector<double> uc(ns), un(ns), lc(ns), lo(ns), eb(ns);
        long long lmin = -1;
#pragma omp for schedule(dynamic, 16)
        for (long long cell = 0; cell < (long long)c->nc; ++cell) {
            CellChem cc(*c, (size_t)cell);
            gather(*c, u0.data(), cell, uc.data());
            PIState s = percell_cell(cc, uc.data(), un.data(), t0, t1, dt0, ctrl, atol, rtol,
                                     clamp, maxiters, w, &rec);
            if (s.iters >= maxiters && s.t < t1 - 1.0e-9) nfail += 1;
            if (uend) scatter(*c, uc.data(), cell, uend->data());
            att += s.iters; acc += s.nacc; rej += s.nrej;
            amax = std::max(amax, s.iters);
            lmin = (lmin < 0) ? s.iters : std::min(lmin, s.iters);
            const bool surf = (size_t)cell < c->ncol;
            gather(*c, c->lam.data(), cell, lc.data());
            for (long long q = (long long)rec.dt.size() - 1; q >= 0; --q) {
                const double* uout = rec.st.data() + (size_t)(q + 1) * ns;
                if (clamp & 1)
                    for (int i = 0; i < ns; ++i) lc[i] = lc[i] * ((uout[i] > 0.0) ? 1.0 : 0.0);
                std::fill(eb.begin(), eb.end(), 0.0);
                ros23_vjp_cell(cc, rec.st.data() + (size_t)q * ns, rec.dt[q], lc.data(), lo.data(),
                               surf ? eb.data() : nullptr);
                lc.swap(lo);
                nv += 1;
                if (surf)
                    for (int i = 0; i < ns; ++i) {
                        ng += !std::isfinite(eb[i]);
                        c->gemis[(size_t)i * c->ncol + cell] += eb[i];
                    }
            }
            scatter(*c, lc.data(), cell, c->lam.data());
        }
#pragma omp critical
        { if (lmin >= 0) amin = (amin < 0) ? lmin : std::min(amin, lmin); }
    }
    if (nvjp) *nvjp = nv;
    if (ngrad) *ngrad = ng;
    if (st) {
        st->attempts_total = att; st->accepts_total = acc; st->rejects_total = rej;
        st->attempts_max = amax; st->attempts_min = amin < 0 ? 0 : amin; st->n_failed = nfail;
    }
    if (nfail) return fail(c, RS_E_MAXITERS, std::to_string(nfail) + " cells hit maxiters");
    return RS_OK;
}

int check_ready(rs_ctx* c, bool need_mech, bool need_wind) {
    if (!c) return RS_E_ARG;
    if (c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!c->have_state) return fail(c, RS_E_STATE, "state not set");
    if (need_mech && !c->have_mech) return fail(c, RS_E_STATE, "mechanism not set");
    if (need_wind && c->nt > 0) {
        if (c->scheme == RS_TRANSPORT_PPM) {
            if (!c->have_ppm_geom || !c->have_ppm_flux)
                return fail(c, RS_E_STATE, "PPM geometry or mass fluxes not set");
        } else if (!c->have_wind) {
            return fail(c, RS_E_STATE, "wind not set");
        }
    }
    return RS_OK;
}

// Loss kind 0: J = sum over level-0 cells of u[species]^2; kind 1: J = <w, u>
// (the reference's linear objective, adjoint_gradient.jl:404-426).
double loss_value(const rs_ctx& c, int kind, int species) {
    double J = 0.0;
    if (kind == 1) {
        for (size_t x = 0; x < c.u.size(); ++x) J += c.wobj[x] * c.u[x];
        return J;
    }
    const double* q = c.u.data() + (size_t)species * c.nc;
    for (size_t x = 0; x < c.ncol; ++x) J += q[x] * q[x];
    return J;
}

int check_loss(rs_ctx* c, int kind, int species) {
    if (kind == 0) return (species < 0 || species >= c->ns) ? fail(c, RS_E_ARG, "bad loss species") : RS_OK;
    if (kind == 1) return c->wobj.empty() ? fail(c, RS_E_STATE, "objective weights not set") : RS_OK;
    return fail(c, RS_E_ARG, "bad loss kind");
}

// One macro step: transport then chemistry, recording the accepted sequences.
int macro(rs_ctx* c, double t0, double t1, double* dtT, double* dtC, const rs_split_opts& o,
          rs_counts* cT, rs_counts* cC, Seq* sT, Seq* sC) {
    int rc = RS_OK;
    if (c->nt > 0) {
        rc = adaptive(c, Op::T, t0, t1, dtT, o.ctrl_T, o.atol_T, o.rtol_T,
                      o.clamp_nonneg & RS_CLAMP_NONNEG,
                      o.maxiters, cT, sT);
        if (rc) return rc;
    }
    if (o.chem_percell) {
        rs_percell_stats st{};
        rc = percell(c, t0, t1, *dtC, nullptr, o.ctrl_C, o.atol_C, o.rtol_C, o.clamp_nonneg,
                     o.maxiters, &st);
        if (cC) { cC->naccept = st.accepts_total; cC->nreject = st.rejects_total; cC->t_end = t1; cC->dt_end = *dtC; }
        return rc;
    }
    return adaptive(c, Op::C, t0, t1, dtC, o.ctrl_C, o.atol_C, o.rtol_C, o.clamp_nonneg,
                    o.maxiters, cC, sC);
}

// Returns the number of non-finite emission-gradient contributions.
long long ros23_vjp_all(rs_ctx* c, const std::vector<double>& u, double dt) {
    std::vector<double> out(c->lam.size());
    long long nbad = 0;
#pragma omp parallel reduction(+ : nbad)
    {
        std::vector<double> uc(c->ns), lc(c->ns), lo(c->ns), eb(c->ns);
#pragma omp for schedule(static)
        for (long long cell = 0; cell < (long long)c->nc; ++cell) {
            CellChem cc(*c, (size_t)cell);
            gather(*c, u.data(), cell, uc.data());
            gather(*c, c->lam.data(), cell, lc.data());
            bool surf = (size_t)cell < c->ncol;
            std::fill(eb.begin(), eb.end(), 0.0);
            ros23_vjp_cell(cc, uc.data(), dt, lc.data(), lo.data(), surf ? eb.data() : nullptr);
            scatter(*c, lo.data(), cell, out.data());
            if (surf)
                for (int s = 0; s < c->ns; ++s) {
                    nbad += !std::isfinite(eb[s]);
                    c->gemis[(size_t)s * c->ncol + cell] += eb[s];
                }
        }
    }
    c->lam.swap(out);
    return nbad;
}

long long count_nonfinite(const std::vector<double>& v) {
    long long n = 0;
    for (double x : v) n += !std::isfinite(x);
    return n;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
// Weights of the PBL mixing from pbl_f and dp (level order, as the GPU).
static void ppm_pbl_weights(rs_ctx* c) {
    if (c->pblf.empty()) return;
    const size_t plane = (size_t)c->nlon * c->nlat;
    c->pblw.assign(c->nc, 0.0);
    c->pblWp.assign(plane, 1.0);
    for (size_t col = 0; col < plane; ++col) {
        double W = 0.0;
        for (int k = 0; k < c->nlev; ++k) {
            const size_t x = (size_t)k * plane + col;
            const double w = c->dp[x] * c->pblf[x];
            c->pblw[x] = w;
            W = W + w;
        }
        c->pblWp[col] = std::max(W, 1.0);
    }
}

extern "C" {

int rs_create(rs_ctx** out, int device) {
    (void)device;
    if (!out) return RS_E_ARG;
    *out = new rs_ctx();
    return RS_OK;
}
int rs_destroy(rs_ctx* c) { delete c; return RS_OK; }
const char* rs_last_error(const rs_ctx* c) { return c ? c->err.c_str() : "null context"; }
const char* rs_backend(void) { return "cpu-oracle"; }

int rs_set_grid(rs_ctx* c, int nlon, int nlat, int nlev, int ns, int nt) {
    if (!c) return RS_E_ARG;
    if (nlon < 1 || nlat < 1 || nlev < 1 || ns < 1 || nt < 0 || nt > ns)
        return fail(c, RS_E_ARG, "bad grid sizes");
    c->nlon = nlon; c->nlat = nlat; c->nlev = nlev; c->ns = ns; c->nt = nt;
    c->ncol = (size_t)nlon * nlat;
    c->nc = c->ncol * nlev;
    c->u.assign((size_t)ns * c->nc, 0.0);
    c->lam.assign((size_t)ns * c->nc, 0.0);
    c->gemis.assign((size_t)ns * c->ncol, 0.0);
    c->jfac.clear(); c->emis.clear();
    c->have_state = c->have_mech = c->have_wind = false;
    c->wobj.clear();
    c->scheme = RS_TRANSPORT_UPWIND3;
    c->have_ppm_geom = c->have_ppm_flux = false;
    c->dPSdt.clear();
    c->msrc.clear();
    c->pblf.clear(); c->pblw.clear(); c->pblWp.clear();
    const size_t nq = nt > 1 ? (size_t)(nt - 1) : 0;
    c->qw.assign(nq * nlev * nlat, 0.0); c->qe = c->qw;
    c->qs.assign(nq * nlev * nlon, 0.0); c->qn = c->qs;
    return RS_OK;
}

int rs_set_wind(rs_ctx* c, const double* cx, const double* wz, const double* wl) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!cx || !wz || !wl) return fail(c, RS_E_ARG, "null wind array");
    c->cx.assign(cx, cx + c->nlat);
    c->wz.assign(wz, wz + c->nlev + 1);
    c->wl.assign(wl, wl + c->nlon);
    for (double x : c->cx) if (!std::isfinite(x)) return fail(c, RS_E_NONFINITE, "cx");
    for (double x : c->wz) if (!std::isfinite(x)) return fail(c, RS_E_NONFINITE, "wz");
    for (double x : c->wl) if (!std::isfinite(x)) return fail(c, RS_E_NONFINITE, "wl");
    c->have_wind = true;
    return RS_OK;
}

static void ppm_update_source(rs_ctx* c);

int rs_set_transport_scheme(rs_ctx* c, int scheme) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (scheme != RS_TRANSPORT_UPWIND3 && scheme != RS_TRANSPORT_PPM)
        return fail(c, RS_E_ARG, "unknown transport scheme");
    if (scheme == RS_TRANSPORT_PPM && c->nt < 1)
        return fail(c, RS_E_ARG, "PPM needs the air mass as transported species 0");
    c->scheme = scheme;
    return RS_OK;
}

static bool all_finite(const std::vector<double>& v) {
    for (double x : v) if (!std::isfinite(x)) return false;
    return true;
}

int rs_set_ppm_geometry(rs_ctx* c, int lon_periodic, double dlon, const double* coslat_e,
                        const double* dS_lat, const double* dphi_lat) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!coslat_e || !dS_lat || !dphi_lat) return fail(c, RS_E_ARG, "null geometry array");
    if (!(dlon > 0.0) || !std::isfinite(dlon)) return fail(c, RS_E_ARG, "dlon must be positive");
    c->lon_periodic = lon_periodic ? 1 : 0;
    c->dlon = dlon;
    c->coslat_e.assign(coslat_e, coslat_e + c->nlat + 1);
    c->dS_lat.assign(dS_lat, dS_lat + c->nlat);
    c->dphi_lat.assign(dphi_lat, dphi_lat + c->nlat);
    if (!all_finite(c->coslat_e) || !all_finite(c->dphi_lat)) return fail(c, RS_E_NONFINITE, "geometry");
    for (double x : c->dS_lat) if (!(x > 0.0) || !std::isfinite(x)) return fail(c, RS_E_ARG, "dS_lat must be positive");
    c->have_ppm_geom = true;
    ppm_update_source(c);
    return RS_OK;
}

int rs_set_mass_fluxes(rs_ctx* c, const double* Mx, const double* My, const double* Mz,
                       const double* dp) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!Mx || !My || !dp) return fail(c, RS_E_ARG, "null flux array");
    const size_t nlon = c->nlon, nlat = c->nlat, nlev = c->nlev;
    c->Mx.assign(Mx, Mx + nlev * nlat * (nlon + 1));
    c->My.assign(My, My + nlev * (nlat + 1) * nlon);
    // Mz NULL: no vertical flux until rs_diagnose_mz supplies it
    if (Mz) c->Mz.assign(Mz, Mz + (nlev + 1) * nlat * nlon);
    else c->Mz.assign((nlev + 1) * nlat * nlon, 0.0);
    c->dp.assign(dp, dp + c->nc);
    if (!all_finite(c->Mx) || !all_finite(c->My) || !all_finite(c->Mz))
        return fail(c, RS_E_NONFINITE, "mass flux");
    for (double x : c->dp) if (!(x > 0.0) || !std::isfinite(x)) return fail(c, RS_E_ARG, "dp must be positive");
    for (size_t x = 0; x < c->ncol; ++x)
        if (c->Mz[x] != 0.0 || c->Mz[nlev * c->ncol + x] != 0.0)
            return fail(c, RS_E_ARG, "Mz must vanish at the surface and the lid");
    c->have_ppm_flux = true;
    ppm_pbl_weights(c);
    ppm_update_source(c);
    return RS_OK;
}

// Air-mass source of the column pressure fixer: (dPSdt + divh_col) dp/dp_col,
// divh_col and dp_col summed over levels in order.
static void ppm_update_source(rs_ctx* c) {
    c->msrc.clear();
    if (c->dPSdt.empty() || !c->have_ppm_geom || !c->have_ppm_flux) return;
    const int nlon = c->nlon, nlat = c->nlat, nlev = c->nlev;
    const size_t plane = (size_t)nlon * nlat;
    c->msrc.assign(c->nc, 0.0);
    for (int j = 0; j < nlat; ++j) {
        const double gx = c->dphi_lat[j] / (c->dlon * c->dS_lat[j]);
        const double rdS = 1.0 / c->dS_lat[j];
        const double ces = c->coslat_e[j], cen = c->coslat_e[j + 1];
        for (int i = 0; i < nlon; ++i) {
            const size_t col = (size_t)j * nlon + i;
            double dcol = 0.0, pcol = 0.0;
            for (int k = 0; k < nlev; ++k) {
                const double* mxr = c->Mx.data() + ((size_t)k * nlat + j) * (nlon + 1);
                const double ms = c->My[((size_t)k * (nlat + 1) + j) * nlon + i];
                const double mn = c->My[((size_t)k * (nlat + 1) + j + 1) * nlon + i];
                const double dx = (mxr[i + 1] - mxr[i]) * gx;
                const double dy = (cen * mn - ces * ms) * rdS;
                dcol += dx + dy;
                pcol += c->dp[(size_t)k * plane + col];
            }
            const double a = (c->dPSdt[col] + dcol) / pcol;
            for (int k = 0; k < nlev; ++k) c->msrc[(size_t)k * plane + col] = a * c->dp[(size_t)k * plane + col];
        }
    }
}

int rs_set_pressure_fixer(rs_ctx* c, const double* dPSdt) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!dPSdt) {
        c->dPSdt.clear();
        c->msrc.clear();
        return RS_OK;
    }
    if (!c->have_ppm_geom || !c->have_ppm_flux)
        return fail(c, RS_E_STATE, "set the PPM geometry and mass fluxes first");
    c->dPSdt.assign(dPSdt, dPSdt + c->ncol);
    if (!all_finite(c->dPSdt)) return fail(c, RS_E_NONFINITE, "dPSdt");
    ppm_update_source(c);
    return RS_OK;
}

// Vertical air-mass flux diagnosed from continuity with the column pressure
// fixer (reseact.esm:262-291, variables divh, divh_col, dp_col, divh_fix, Mz).
// Per column, with divh(k) the horizontal divergence of the fluxes set by
// rs_set_mass_fluxes (the expression of ppm_update_source) and divh_col,
// dp_col their sums in level order:
//   a = (dPSdt + divh_col) / dp_col,   divh_fix(k) = divh(k) - a dp(k)
//     (= divh(k) + (-dPSdt - divh_col) dp(k)/dp_col),
//   Mz[0] = 0,  Mz[k+1] = Mz[k] - (dB[k] dPSdt + divh_fix(k))
//     (the exclusive scan Mz[ke] = -sum_{k<ke} (dB[k] dPSdt + divh_fix(k))).
// Since sum_k divh_fix(k) = -dPSdt, the scan reaches dPSdt (1 - sum_k dB)
// at the lid up to round-off: zero when sum_k dB = 1.  The lid face is then
// set to 0 (the no-flux wall the top face stencil assumes) and the largest
// |scan value at the lid| over the columns goes to lid_resid.  dB (NLEV) and
// dPSdt (NCOL) may be NULL (zeros).
int rs_diagnose_mz(rs_ctx* c, const double* dB, const double* dPSdt, double* Mz_out,
                   double* lid_resid) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!c->have_ppm_geom || !c->have_ppm_flux)
        return fail(c, RS_E_STATE, "set the PPM geometry and mass fluxes first");
    const int nlon = c->nlon, nlat = c->nlat, nlev = c->nlev;
    const size_t plane = (size_t)nlon * nlat;
    if (dB)
        for (int k = 0; k < nlev; ++k)
            if (!std::isfinite(dB[k])) return fail(c, RS_E_NONFINITE, "dB");
    if (dPSdt)
        for (size_t x = 0; x < plane; ++x)
            if (!std::isfinite(dPSdt[x])) return fail(c, RS_E_NONFINITE, "dPSdt");
    double worst = 0.0;
    for (int j = 0; j < nlat; ++j) {
        const double gx = c->dphi_lat[j] / (c->dlon * c->dS_lat[j]);
        const double rdS = 1.0 / c->dS_lat[j];
        const double ces = c->coslat_e[j], cen = c->coslat_e[j + 1];
        for (int i = 0; i < nlon; ++i) {
            const size_t col = (size_t)j * nlon + i;
            auto divh = [&](int k) {
                const double* mxr = c->Mx.data() + ((size_t)k * nlat + j) * (nlon + 1);
                const double ms = c->My[((size_t)k * (nlat + 1) + j) * nlon + i];
                const double mn = c->My[((size_t)k * (nlat + 1) + j + 1) * nlon + i];
                const double dx = (mxr[i + 1] - mxr[i]) * gx;
                const double dy = (cen * mn - ces * ms) * rdS;
                return dx + dy;
            };
            double dcol = 0.0, pcol = 0.0;
            for (int k = 0; k < nlev; ++k) {
                dcol += divh(k);
                pcol += c->dp[(size_t)k * plane + col];
            }
            const double ps = dPSdt ? dPSdt[col] : 0.0;
            const double a = (ps + dcol) / pcol;
            double mz = 0.0;
            c->Mz[col] = 0.0;
            for (int k = 0; k < nlev; ++k) {
                const double fix = divh(k) - a * c->dp[(size_t)k * plane + col];
                const double db = dB ? dB[k] : 0.0;
                const double t = db * ps + fix;
                mz = mz - t;
                if (k + 1 < nlev) c->Mz[(size_t)(k + 1) * plane + col] = mz;
            }
            c->Mz[(size_t)nlev * plane + col] = 0.0;
            // NaN-propagating max of |lid value|
            const double r = std::fabs(mz);
            if (!(r <= worst)) worst = r;
        }
    }
    if (Mz_out) std::copy(c->Mz.begin(), c->Mz.end(), Mz_out);
    if (lid_resid) *lid_resid = worst;
    return RS_OK;
}

int rs_set_pbl_mixing(rs_ctx* c, const double* pbl_f, double tau_pblmix) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!pbl_f) {
        c->pblf.clear(); c->pblw.clear(); c->pblWp.clear();
        return RS_OK;
    }
    if (!c->have_ppm_flux) return fail(c, RS_E_STATE, "set the PPM mass fluxes (dp) first");
    if (!(tau_pblmix > 0.0) || !std::isfinite(tau_pblmix)) return fail(c, RS_E_ARG, "tau_pblmix must be positive");
    c->pblf.assign(pbl_f, pbl_f + c->nc);
    if (!all_finite(c->pblf)) { c->pblf.clear(); return fail(c, RS_E_NONFINITE, "pbl_f"); }
    c->pbl_tau = tau_pblmix;
    ppm_pbl_weights(c);
    return RS_OK;
}

int rs_set_inflow(rs_ctx* c, const double* qbc_w, const double* qbc_e, const double* qbc_s,
                  const double* qbc_n) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    const size_t nq = c->nt > 1 ? (size_t)(c->nt - 1) : 0;
    const size_t nwe = nq * c->nlev * c->nlat, nsn = nq * c->nlev * c->nlon;
    auto put = [&](std::vector<double>& d, const double* src, size_t n) {
        if (src) d.assign(src, src + n); else d.assign(n, 0.0);
    };
    put(c->qw, qbc_w, nwe); put(c->qe, qbc_e, nwe); put(c->qs, qbc_s, nsn); put(c->qn, qbc_n, nsn);
    if (!all_finite(c->qw) || !all_finite(c->qe) || !all_finite(c->qs) || !all_finite(c->qn))
        return fail(c, RS_E_NONFINITE, "inflow mixing ratio");
    return RS_OK;
}

// Mass-action mechanism (rx_traced_integrator.jl:151-253 for the Jacobian
// conventions): net stoichiometry per reaction, species ascending.
static int set_mechanism_common(rs_ctx* c, int nr, const int* re, const double* k,
                                const int* photo,
                                const std::vector<std::vector<std::pair<int, double>>>& prods) {
    c->nr = nr;
    c->ra.resize(nr); c->rb.resize(nr); c->pa.assign(nr, -1); c->pb.assign(nr, -1);
    c->photo.assign(nr, 0); c->kr.resize(nr); c->net.assign(nr, {});
    for (int r = 0; r < nr; ++r) {
        int a = re[2 * r], b = re[2 * r + 1];
        if (a < 0 || a >= c->ns || b < -1 || b >= c->ns)
            return fail(c, RS_E_ARG, "species index out of range in reaction " + std::to_string(r));
        for (auto& q : prods[r]) {
            if (q.first < 0 || q.first >= c->ns)
                return fail(c, RS_E_ARG, "product species out of range in reaction " + std::to_string(r));
            if (!(q.second > 0.0) || !(q.second < 1e6))
                return fail(c, RS_E_ARG, "product coefficient must be positive and finite");
        }
        if (!std::isfinite(k[r])) return fail(c, RS_E_NONFINITE, "rate constant");
        c->ra[r] = a; c->rb[r] = b;
        c->kr[r] = k[r];
        c->photo[r] = photo ? (photo[r] != 0) : 0;
        if (c->photo[r] && b >= 0)
            return fail(c, RS_E_ARG, "photolysis reaction " + std::to_string(r) + " must be first order");
        std::vector<double> nu(c->ns, 0.0);
        nu[a] -= 1.0;
        if (b >= 0) nu[b] -= 1.0;
        for (auto& q : prods[r]) nu[q.first] += q.second;
        for (int s = 0; s < c->ns; ++s)
            if (nu[s] != 0.0) c->net[r].emplace_back(s, nu[s]);
    }
    c->have_mech = true;
    return RS_OK;
}

int rs_set_mechanism(rs_ctx* c, int nr, const int* re, const int* pr, const double* k,
                     const int* photo) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (nr < 0 || (nr > 0 && (!re || !pr || !k))) return fail(c, RS_E_ARG, "bad mechanism arrays");
    std::vector<std::vector<std::pair<int, double>>> prods(nr);
    for (int r = 0; r < nr; ++r) {
        for (int j = 0; j < 2; ++j) {
            const int p = pr[2 * r + j];
            if (p < -1 || p >= c->ns)
                return fail(c, RS_E_ARG, "species index out of range in reaction " + std::to_string(r));
            if (p >= 0) prods[r].emplace_back(p, 1.0);
        }
    }
    int rc = set_mechanism_common(c, nr, re, k, photo, prods);
    if (rc) return rc;
    for (int r = 0; r < nr; ++r) { c->pa[r] = pr[2 * r]; c->pb[r] = pr[2 * r + 1]; }
    return RS_OK;
}

int rs_set_mechanism_stoich(rs_ctx* c, int nr, const int* re, const int* prod_off,
                            const int* prod_species, const double* prod_coef, const double* k,
                            const int* photo) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (nr < 0 || (nr > 0 && (!re || !prod_off || !k))) return fail(c, RS_E_ARG, "bad mechanism arrays");
    if (nr > 0 && (prod_off[0] != 0 || (prod_off[nr] > 0 && (!prod_species || !prod_coef))))
        return fail(c, RS_E_ARG, "bad product offsets");
    std::vector<std::vector<std::pair<int, double>>> prods(nr);
    for (int r = 0; r < nr; ++r) {
        if (prod_off[r + 1] < prod_off[r]) return fail(c, RS_E_ARG, "bad product offsets");
        for (int q = prod_off[r]; q < prod_off[r + 1]; ++q) prods[r].emplace_back(prod_species[q], prod_coef[q]);
    }
    return set_mechanism_common(c, nr, re, k, photo, prods);
}

int rs_set_photolysis(rs_ctx* c, const double* jfac) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!jfac) { c->jfac.clear(); return RS_OK; }
    c->jfac.assign(jfac, jfac + c->nc);
    return RS_OK;
}

int rs_set_emissions(rs_ctx* c, const double* e) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!e) { c->emis.clear(); return RS_OK; }
    c->emis.assign(e, e + (size_t)c->ns * c->ncol);
    return RS_OK;
}

int rs_set_state(rs_ctx* c, const double* u) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!u) return fail(c, RS_E_ARG, "null state");
    c->u.assign(u, u + (size_t)c->ns * c->nc);
    c->have_state = true;
    c->have_cand = false;
    return RS_OK;
}

int rs_get_state(rs_ctx* c, double* u) {
    if (!c || !c->have_state) return fail(c, RS_E_STATE, "state not set");
    if (!u) return fail(c, RS_E_ARG, "null output");
    std::copy(c->u.begin(), c->u.end(), u);
    return RS_OK;
}

int rs_clamp_state(rs_ctx* c) {
    if (!c || !c->have_state) return fail(c, RS_E_STATE, "state not set");
    for (double& x : c->u) x = std::max(x, 0.0);
    return RS_OK;
}

int rs_ssprk43_step(rs_ctx* c, double t, double dt, double atol, double rtol, double* eest) {
    (void)t;
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    double e = ssprk43_attempt(*c, c->u, c->cand, dt, atol, rtol);
    c->have_cand = true;
    if (eest) *eest = e;
    return RS_OK;
}

int rs_ros23_step(rs_ctx* c, double t, double dt, double atol, double rtol, double* eest) {
    (void)t;
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    double e = ros23_attempt(*c, c->u, c->cand, dt, atol, rtol);
    c->have_cand = true;
    if (eest) *eest = e;
    return RS_OK;
}

// The no-allocation attempt of the HIP library on the CPU: the caller's
// buffers are host memory here.  One attempt on the first ncells cells,
// species-major with stride ncells; the context's state is not touched.
int rs_ros23_step_noalloc(rs_ctx* c, const double* u, double* unew, long long ncells, double dt,
                          double atol, double rtol, double* eest) {
    if (!c) return RS_E_ARG;
    if (c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!c->have_mech) return fail(c, RS_E_STATE, "mechanism not set");
    if (!u || !unew) return fail(c, RS_E_ARG, "null buffer");
    if (ncells < 1 || ncells > (long long)c->nc)
        return fail(c, RS_E_ARG, "ncells must be in [1, NC] of the context's grid");
    const int ns = c->ns;
    std::vector<double> sq((size_t)ncells);
#pragma omp parallel
    {
        Ros23Work w(ns);
        std::vector<double> uc(ns), un(ns);
#pragma omp for schedule(static)
        for (long long cell = 0; cell < ncells; ++cell) {
            CellChem cc(*c, (size_t)cell);
            for (int s = 0; s < ns; ++s) uc[s] = u[(size_t)s * ncells + cell];
            sq[cell] = ros23_cell(cc, uc.data(), un.data(), dt, atol, rtol, w);
            for (int s = 0; s < ns; ++s) unew[(size_t)s * ncells + cell] = un[s];
        }
    }
    if (eest) *eest = std::sqrt(fixed_sum(sq) / ((double)ns * (double)ncells));
    return RS_OK;
}

int rs_get_candidate(rs_ctx* c, double* u) {
    if (!c || !c->have_cand) return fail(c, RS_E_STATE, "no candidate");
    std::copy(c->cand.begin(), c->cand.end(), u);
    return RS_OK;
}

int rs_accept_candidate(rs_ctx* c, int clamp) {
    if (!c || !c->have_cand) return fail(c, RS_E_STATE, "no candidate");
    if (clamp & 1) for (double& x : c->cand) x = std::max(x, 0.0);
    c->u.swap(c->cand);
    c->have_cand = false;
    return RS_OK;
}

int rs_adaptive_transport(rs_ctx* c, double t0, double t1, double* dt_io, const rs_pictrl* ctrl,
                          double atol, double rtol, int clamp, long long maxiters, rs_counts* out) {
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    if (!dt_io || !ctrl || !(*dt_io > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    c->lastseq[0].ts.clear();
    return adaptive(c, Op::T, t0, t1, dt_io, *ctrl, atol, rtol, clamp, maxiters, out,
                    &c->lastseq[0], &c->alog);
}

int rs_adaptive_chem(rs_ctx* c, double t0, double t1, double* dt_io, const rs_pictrl* ctrl,
                     double atol, double rtol, int clamp, long long maxiters, rs_counts* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    if (!dt_io || !ctrl || !(*dt_io > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    c->lastseq[1].ts.clear();
    return adaptive(c, Op::C, t0, t1, dt_io, *ctrl, atol, rtol, clamp, maxiters, out,
                    &c->lastseq[1], &c->alog);
}

int rs_get_step_log(rs_ctx* c, int op, double* t_dt, long long cap, long long* n_out) {
    if (!c || (op != 0 && op != 1) || !n_out) return fail(c, RS_E_ARG, "bad step-log arguments");
    const auto& ts = c->lastseq[op].ts;
    *n_out = (long long)ts.size();
    if (t_dt) {
        if (cap < (long long)ts.size()) return fail(c, RS_E_ARG, "step-log buffer too small");
        for (size_t i = 0; i < ts.size(); ++i) { t_dt[2 * i] = ts[i].first; t_dt[2 * i + 1] = ts[i].second; }
    }
    return RS_OK;
}

int rs_chem_percell(rs_ctx* c, double t0, double t1, double dt0, double* dt_cells,
                    const rs_pictrl* ctrl, double atol, double rtol, int clamp,
                    long long maxiters, rs_percell_stats* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    if (!ctrl || !(dt0 > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    return percell(c, t0, t1, dt0, dt_cells, *ctrl, atol, rtol, clamp, maxiters, out);
}

int rs_chem_percell_vjp(rs_ctx* c, double t0, double t1, double dt0, const rs_pictrl* ctrl,
                        double atol, double rtol, int clamp, long long maxiters,
                        rs_percell_stats* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    if (!ctrl || !(dt0 > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    return percell_vjp(c, c->u, t0, t1, dt0, *ctrl, atol, rtol, clamp, maxiters, nullptr, nullptr,
                       nullptr, out);
}

int rs_macro_step(rs_ctx* c, double t0, double t1, double* dtT, double* dtC,
                  const rs_split_opts* o, rs_counts* cT, rs_counts* cC) {
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!o || !dtT || !dtC) return fail(c, RS_E_ARG, "null argument");
    if (cT) *cT = rs_counts{};
    if (cC) *cC = rs_counts{};
    return macro(c, t0, t1, dtT, dtC, *o, cT, cC, nullptr, nullptr);
}

int rs_set_lambda(rs_ctx* c, const double* lam) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    c->lam.assign(lam, lam + (size_t)c->ns * c->nc);
    return RS_OK;
}
int rs_get_lambda(rs_ctx* c, double* lam) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    std::copy(c->lam.begin(), c->lam.end(), lam);
    return RS_OK;
}
int rs_zero_grad(rs_ctx* c) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    std::fill(c->gemis.begin(), c->gemis.end(), 0.0);
    return RS_OK;
}
int rs_get_grad_emis(rs_ctx* c, double* g) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    std::copy(c->gemis.begin(), c->gemis.end(), g);
    return RS_OK;
}

// A performance hint of the device library (keep the PPM stage states of a
// step for its VJP); the oracle always recomputes them, with the same result.
int rs_keep_transport_stages(rs_ctx* c, int keep) {
    (void)keep;
    return c ? RS_OK : RS_E_ARG;
}

int rs_ssprk43_vjp(rs_ctx* c, double t, double dt) {
    (void)t;
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    // PPM: J depends on the state; the step input is the context's state
    if (c->scheme == RS_TRANSPORT_PPM) ssprk43_vjp_ppm(*c, c->u.data(), c->lam, dt);
    else ssprk43_vjp(*c, c->lam, dt);
    return RS_OK;
}

int rs_ros23_vjp(rs_ctx* c, double t, double dt) {
    (void)t;
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    ros23_vjp_all(c, c->u, dt);
    return RS_OK;
}

// The time-loop adjoint; with fixedT/fixedC the forward pass takes those
// accepted sequences (one per macro step) instead of running the controllers.
static int adjoint_impl(rs_ctx* c, int nmacro, const double* tstops, double dtT0, double dtC0,
                        const rs_split_opts* o, int loss_kind, int loss_species, double* J_out,
                        double* grad_u0, double* grad_emis, rs_adjoint_stats* st,
                        const std::vector<Seq>* fixedT, const std::vector<Seq>* fixedC) {
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!o || !tstops || nmacro < 1) return fail(c, RS_E_ARG, "bad adjoint arguments");
    rc = check_loss(c, loss_kind, loss_species);
    if (rc) return rc;
    using clk = std::chrono::steady_clock;
    const bool pcell = o->chem_percell != 0;
    rs_adjoint_stats S{};
    S.bad_op = S.bad_macro = -1;
    S.bad_step = S.bad_nsteps = -1;
    auto tf0 = clk::now();
    std::vector<std::vector<double>> ck(nmacro);
    c->seqT.assign(nmacro, {}); c->seqC.assign(nmacro, {});
    double dtT = dtT0, dtC = dtC0;
    for (int m = 0; m < nmacro; ++m) {
        ck[m] = c->u;
        if (fixedT) {
            // fixed-sequence forward (replay_fixed, adjoint_gradient.jl:647-661)
            std::vector<double> raw;
            c->seqT[m] = (*fixedT)[m];
            c->seqC[m] = (*fixedC)[m];
            for (auto& pq : c->seqT[m].ts) {
                ssprk43_attempt(*c, c->u, raw, pq.second, o->atol_T, o->rtol_T);
                if (o->clamp_nonneg & 1) for (double& x : raw) x = std::max(x, 0.0);
                c->u.swap(raw);
            }
            if (pcell) {
                // per-cell chemistry has no global sequence: each cell runs free
                rs_percell_stats pst{};
                rc = percell(c, tstops[m], tstops[m + 1], dtC0, nullptr, o->ctrl_C, o->atol_C,
                             o->rtol_C, o->clamp_nonneg, o->maxiters, &pst);
                if (rc) return rc;
                c->seqC[m].ts.clear();
                S.naccept_C += pst.accepts_total; S.nreject_C += pst.rejects_total;
            }
            for (auto& pq : c->seqC[m].ts) {
                ros23_attempt(*c, c->u, raw, pq.second, o->atol_C, o->rtol_C);
                if (o->clamp_nonneg & 1) for (double& x : raw) x = std::max(x, 0.0);
                c->u.swap(raw);
            }
            S.naccept_T += (long long)c->seqT[m].ts.size();
            S.naccept_C += (long long)c->seqC[m].ts.size();
            continue;
        }
        rs_counts cT{}, cC{};
        rc = macro(c, tstops[m], tstops[m + 1], &dtT, &dtC, *o, &cT, &cC, &c->seqT[m], &c->seqC[m]);
        if (rc) return rc;
        S.naccept_T += cT.naccept; S.nreject_T += cT.nreject;
        S.naccept_C += cC.naccept; S.nreject_C += cC.nreject;
    }
    std::vector<double> uend = c->u;
    if (J_out) *J_out = loss_value(*c, loss_kind, loss_species);
    S.seconds_forward = std::chrono::duration<double>(clk::now() - tf0).count();
    c->last_opts = *o;
    c->have_record = true;
    c->rec_tstops.assign(tstops, tstops + nmacro + 1);
    c->rec_dtC0 = dtC0;

    auto tb0 = clk::now();
    double trep = 0.0;
    std::fill(c->lam.begin(), c->lam.end(), 0.0);
    std::fill(c->gemis.begin(), c->gemis.end(), 0.0);
    if (loss_kind == 1) {
        c->lam = c->wobj;
    } else {
        const double* q = uend.data() + (size_t)loss_species * c->nc;
        double* L = c->lam.data() + (size_t)loss_species * c->nc;
        for (size_t x = 0; x < c->ncol; ++x) L[x] = 2.0 * q[x];
    }
    // first VJP that made lambda or its gradient non-finite
    // (adjoint_gradient.jl:763-779)
    long long nprev = count_nonfinite(c->lam);
    bool bad = false;
    auto note = [&](int op, int m, int j, const Seq& sq, long long ngrad) {
        const long long nout = count_nonfinite(c->lam);
        if (!bad && (nout > nprev || ngrad > 0)) {
            bad = true;
            S.bad_op = op; S.bad_macro = m; S.bad_step = j; S.bad_nsteps = (long long)sq.ts.size();
            S.bad_t = sq.ts[j].first; S.bad_dt = sq.ts[j].second;
            S.bad_lam_in = nprev; S.bad_lam_out = nout; S.bad_grad = ngrad;
        }
        nprev = nout;
    };
    int worst = 0;
    const int clampf = o->clamp_nonneg & 1;
    for (int m = nmacro - 1; m >= 0; --m) {
        auto tr0 = clk::now();
        // replay_fixed: tape of step inputs, each followed by its clamped output
        std::vector<std::vector<double>> tapeT, tapeC;
        std::vector<double> u = ck[m], raw;
        for (auto& p : c->seqT[m].ts) {
            tapeT.push_back(u);
            ssprk43_attempt(*c, u, raw, p.second, o->atol_T, o->rtol_T);
            if (clampf) for (double& x : raw) x = std::max(x, 0.0);
            u.swap(raw);
        }
        std::vector<double> uT = u;
        for (auto& p : c->seqC[m].ts) {
            tapeC.push_back(u);
            ros23_attempt(*c, u, raw, p.second, o->atol_C, o->rtol_C);
            if (clampf) for (double& x : raw) x = std::max(x, 0.0);
            u.swap(raw);
        }
        // per-cell chemistry: replayed and swept cell by cell (percell_vjp);
        // u gets its end state for the replay check, lambda its adjoint
        long long pc_nv = 0, pc_ng = 0;
        if (pcell) {
            rc = percell_vjp(c, uT, tstops[m], tstops[m + 1], dtC0, o->ctrl_C, o->atol_C,
                             o->rtol_C, o->clamp_nonneg, o->maxiters, &u, &pc_nv, &pc_ng, nullptr);
            if (rc) return rc;
        }
        const std::vector<double>& ref = (m + 1 < nmacro) ? ck[m + 1] : uend;
        double d = 0.0;
        for (size_t x = 0; x < u.size(); ++x) {
            const double e = std::fabs(u[x] - ref[x]) / std::max(std::fabs(ref[x]), 1e-30);
            if (!(e <= d)) d = e;   // a NaN deviation wins
        }
        if (!(d <= S.replay_maxrel)) { S.replay_maxrel = d; worst = m; }
        trep += std::chrono::duration<double>(clk::now() - tr0).count();
        // chemistry backwards (adjoint_gradient.jl:821); the clamp mask is a
        // multiplication (lam .* mask, :736), so a non-finite entry stays
        auto mask = [&](const std::vector<double>& out) {
            if (!clampf) return;
            for (size_t x = 0; x < c->lam.size(); ++x) c->lam[x] = c->lam[x] * ((out[x] > 0.0) ? 1.0 : 0.0);
        };
        for (int jx = (int)tapeC.size() - 1; jx >= 0; --jx) {
            const std::vector<double>& outj = (jx + 1 < (int)tapeC.size()) ? tapeC[jx + 1] : u;
            mask(outj);
            const long long ng = ros23_vjp_all(c, tapeC[jx], c->seqC[m].ts[jx].second);
            note(1, m, jx, c->seqC[m], ng);
            S.nvjp_C += 1;
        }
        if (pcell) {
            // reported as one step spanning the macro step
            Seq whole;
            whole.ts.emplace_back(tstops[m], tstops[m + 1] - tstops[m]);
            note(1, m, 0, whole, pc_ng);
            S.nvjp_C += pc_nv;
        }
        for (int jx = (int)tapeT.size() - 1; jx >= 0; --jx) {
            const std::vector<double>& outj = (jx + 1 < (int)tapeT.size()) ? tapeT[jx + 1] : uT;
            mask(outj);
            if (c->scheme == RS_TRANSPORT_PPM)
                ssprk43_vjp_ppm(*c, tapeT[jx].data(), c->lam, c->seqT[m].ts[jx].second);
            else
                ssprk43_vjp(*c, c->lam, c->seqT[m].ts[jx].second);
            note(0, m, jx, c->seqT[m], 0);
            S.nvjp_T += 1;
        }
    }
    S.seconds_backward = std::chrono::duration<double>(clk::now() - tb0).count();
    S.seconds_replay = trep;
    if (grad_u0) std::copy(c->lam.begin(), c->lam.end(), grad_u0);
    if (grad_emis) std::copy(c->gemis.begin(), c->gemis.end(), grad_emis);
    if (st) *st = S;
    if (!(S.replay_maxrel < 1e-6)) {   // adjoint_gradient.jl:802-808
        char b[200];
        std::snprintf(b, sizeof(b), "fixed-sequence replay of macro step %d lands %.3g relative "
                      "away from the checkpointed end state: the tape is not the forward "
                      "trajectory", worst + 1, S.replay_maxrel);
        return fail(c, RS_E_STATE, b);
    }
    if (bad) {
        char b[320];
        std::snprintf(b, sizeof(b),
                      "non-finite adjoint: first at %s step %lld of %lld in macro step %d "
                      "(t=%.9g dt=%.9g): lambda non-finite %lld -> %lld, gradient non-finite %lld",
                      S.bad_op ? "chemistry" : "transport", S.bad_step + 1, S.bad_nsteps,
                      S.bad_macro + 1, S.bad_t, S.bad_dt, S.bad_lam_in, S.bad_lam_out, S.bad_grad);
        return fail(c, RS_E_NONFINITE, b);
    }
    return RS_OK;
}

int rs_adjoint_run(rs_ctx* c, int nmacro, const double* tstops, double dtT0, double dtC0,
                   const rs_split_opts* o, int loss_kind, int loss_species, double* J_out,
                   double* grad_u0, double* grad_emis, rs_adjoint_stats* st) {
    return adjoint_impl(c, nmacro, tstops, dtT0, dtC0, o, loss_kind, loss_species, J_out, grad_u0,
                        grad_emis, st, nullptr, nullptr);
}

int rs_get_recorded_steps(rs_ctx* c, int macro_step, int op, double* t_dt, long long cap,
                          long long* n_out) {
    if (!c || !n_out || (op != 0 && op != 1)) return fail(c, RS_E_ARG, "bad arguments");
    if (!c->have_record || macro_step < 0 || macro_step >= (int)c->seqT.size())
        return fail(c, RS_E_STATE, "no recorded macro step " + std::to_string(macro_step));
    const auto& ts = (op ? c->seqC : c->seqT)[macro_step].ts;
    *n_out = (long long)ts.size();
    if (t_dt) {
        if (cap < (long long)ts.size()) return fail(c, RS_E_ARG, "buffer too small");
        for (size_t i = 0; i < ts.size(); ++i) { t_dt[2 * i] = ts[i].first; t_dt[2 * i + 1] = ts[i].second; }
    }
    return RS_OK;
}

int rs_oracle_adjoint_on_sequence(rs_ctx* c, int nmacro, const double* tstops,
                                  const long long* nT, const long long* nC, const double* t_dt_T,
                                  const double* t_dt_C, const rs_split_opts* o, int loss_kind,
                                  int loss_species, double* J_out, double* grad_u0,
                                  double* grad_emis, rs_adjoint_stats* st) {
    if (!c || !nT || !nC || !t_dt_T || !t_dt_C || nmacro < 1) return RS_E_ARG;
    std::vector<Seq> fT(nmacro), fC(nmacro);
    size_t oT = 0, oC = 0;
    for (int m = 0; m < nmacro; ++m) {
        for (long long q = 0; q < nT[m]; ++q, ++oT) fT[m].ts.emplace_back(t_dt_T[2 * oT], t_dt_T[2 * oT + 1]);
        for (long long q = 0; q < nC[m]; ++q, ++oC) fC[m].ts.emplace_back(t_dt_C[2 * oC], t_dt_C[2 * oC + 1]);
    }
    return adjoint_impl(c, nmacro, tstops, 0.0, 0.0, o, loss_kind, loss_species, J_out, grad_u0,
                        grad_emis, st, &fT, &fC);
}

int rs_set_objective_weights(rs_ctx* c, const double* w) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!w) { c->wobj.clear(); return RS_OK; }
    c->wobj.assign(w, w + (size_t)c->ns * c->nc);
    return RS_OK;
}

int rs_replay_loss(rs_ctx* c, int loss_kind, int loss_species, double* J_out) {
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!c->have_record) return fail(c, RS_E_STATE, "no recorded sequence");
    rc = check_loss(c, loss_kind, loss_species);
    if (rc) return rc;
    const rs_split_opts& o = c->last_opts;
    std::vector<double> raw;
    for (size_t m = 0; m < c->seqT.size(); ++m) {
        for (auto& p : c->seqT[m].ts) {
            ssprk43_attempt(*c, c->u, raw, p.second, o.atol_T, o.rtol_T);
            if (o.clamp_nonneg & 1) for (double& x : raw) x = std::max(x, 0.0);
            c->u.swap(raw);
        }
        for (auto& p : c->seqC[m].ts) {
            ros23_attempt(*c, c->u, raw, p.second, o.atol_C, o.rtol_C);
            if (o.clamp_nonneg & 1) for (double& x : raw) x = std::max(x, 0.0);
            c->u.swap(raw);
        }
        if (o.chem_percell) {
            // per-cell chemistry records no steps: it runs free, each cell
            // from the run's start dt
            rs_percell_stats pst{};
            rc = percell(c, c->rec_tstops[m], c->rec_tstops[m + 1], c->rec_dtC0, nullptr, o.ctrl_C,
                         o.atol_C, o.rtol_C, o.clamp_nonneg, o.maxiters, &pst);
            if (rc) return rc;
        }
    }
    if (J_out) *J_out = loss_value(*c, loss_kind, loss_species);
    return RS_OK;
}

// ---- ABI extensions (same symbols as the HIP library) -----------------------
int rs_set_lu_ordering(rs_ctx* c, int ordering) {
    if (!c) return RS_E_ARG;
    if (ordering != 0 && ordering != 1) return fail(c, RS_E_ARG, "ordering must be 0 or 1");
    return RS_OK;   // the oracle always eliminates densely in natural order
}

int rs_set_transport_kernel(rs_ctx* c, int mode) {
    if (!c) return RS_E_ARG;
    if (mode < 0 || mode > 3) return fail(c, RS_E_ARG, "mode must be 0, 1, 2 or 3");
    return RS_OK;   // one transport path on the CPU
}

int rs_mechanism_info(rs_ctx* c, rs_mech_info* out) {
    if (!c || !c->have_mech || !out) return fail(c, RS_E_STATE, "mechanism not set");
    rs_mech_info m{};
    m.ns = c->ns;
    m.nr = c->nr;
    m.nnz_J = 0;
    m.nnz_LU = c->ns * c->ns;
    m.ordering = 0;
    const long long n = c->ns;
    m.flops_attempt = (2 * n * n * n) / 3 + 3 * 2 * n * n;   // dense LU + 3 solves
    m.flops_vjp = m.flops_attempt;
    *out = m;
    return RS_OK;
}

int rs_profile(rs_ctx* c, int enable) { (void)enable; return c ? RS_OK : RS_E_ARG; }

int rs_profile_read(rs_ctx* c, rs_kernel_time* out, int max, int* n_out) {
    (void)out; (void)max;
    if (!c) return RS_E_ARG;
    if (n_out) *n_out = 0;
    return RS_OK;
}

// ---- oracle-only extras (declared in oracle/oracle_extra.h) -----------------
// Tangent of one step at the current state: out = (d unew/d u) v + (d unew/d E) Ev.
int rs_oracle_ros23_jvp(rs_ctx* c, double dt, const double* v, const double* Ev, double* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    std::vector<double> res((size_t)c->ns * c->nc);
#pragma omp parallel
    {
        std::vector<double> uc(c->ns), vc(c->ns), oc(c->ns), ec(c->ns);
#pragma omp for schedule(static)
        for (long long cell = 0; cell < (long long)c->nc; ++cell) {
            CellChem cc(*c, (size_t)cell);
            gather(*c, c->u.data(), cell, uc.data());
            gather(*c, v, cell, vc.data());
            bool surf = Ev && (size_t)cell < c->ncol;
            if (surf) for (int s = 0; s < c->ns; ++s) ec[s] = Ev[(size_t)s * c->ncol + cell];
            ros23_jvp_cell(cc, uc.data(), dt, vc.data(), surf ? ec.data() : nullptr, oc.data());
            scatter(*c, oc.data(), cell, res.data());
        }
    }
    std::copy(res.begin(), res.end(), out);
    return RS_OK;
}

int rs_oracle_ssprk43_jvp(rs_ctx* c, double dt, const double* v, double* out) {
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    if (c->scheme == RS_TRANSPORT_PPM) {       // tangent at the context's state
        ssprk43_jvp_ppm(*c, c->u.data(), v, out, dt);
        return RS_OK;
    }
    std::vector<double> vv(v, v + (size_t)c->ns * c->nc), o;
    ssprk43_jvp(*c, vv, o, dt);
    std::copy(o.begin(), o.end(), out);
    return RS_OK;
}

// Transport tendency f = L q (species < NT) and the dense per-cell Jacobian,
// exposed for the operator-level checks.
int rs_oracle_transport_rhs(rs_ctx* c, const double* q, double* f) {
    if (!c) return RS_E_ARG;
    if (c->scheme == RS_TRANSPORT_PPM ? !(c->have_ppm_geom && c->have_ppm_flux) : !c->have_wind)
        return fail(c, RS_E_STATE, "transport not set up");
    transport_rhs_all(*c, q, f);
    std::fill(f + (size_t)c->nt * c->nc, f + (size_t)c->ns * c->nc, 0.0);
    return RS_OK;
}
// PPM: J(u)^T lam and J(u) v at the context's state u (the tendency is nonlinear).
int rs_oracle_transport_rhs_adj(rs_ctx* c, const double* lam, double* g) {
    if (c && c->scheme == RS_TRANSPORT_PPM) {
        if (!(c->have_ppm_geom && c->have_ppm_flux)) return fail(c, RS_E_STATE, "transport not set up");
        std::fill(g, g + (size_t)c->ns * c->nc, 0.0);
        ppm_rhs_vjp_all(*c, c->u.data(), lam, g);
        return RS_OK;
    }
    if (!c || !c->have_wind) return fail(c, RS_E_STATE, "wind not set");
    std::fill(g, g + (size_t)c->ns * c->nc, 0.0);
    for (int s = 0; s < c->nt; ++s)
        transport_rhs_adj_block(*c, lam + (size_t)s * c->nc, g + (size_t)s * c->nc);
    return RS_OK;
}
int rs_oracle_transport_rhs_jvp(rs_ctx* c, const double* v, double* out) {
    if (!c || c->scheme != RS_TRANSPORT_PPM || !(c->have_ppm_geom && c->have_ppm_flux))
        return fail(c, RS_E_STATE, "PPM transport not set up");
    std::fill(out, out + (size_t)c->ns * c->nc, 0.0);
    ppm_rhs_jvp_all(*c, c->u.data(), v, out);
    return RS_OK;
}
int rs_oracle_chem_rhs(rs_ctx* c, const double* u, double* f) {
    if (!c || !c->have_mech) return fail(c, RS_E_STATE, "mechanism not set");
    std::vector<double> uc(c->ns), fc(c->ns);
    for (size_t cell = 0; cell < c->nc; ++cell) {
        CellChem cc(*c, cell);
        gather(*c, u, cell, uc.data());
        cc.rhs(uc.data(), fc.data());
        scatter(*c, fc.data(), cell, f);
    }
    return RS_OK;
}
int rs_oracle_chem_jac(rs_ctx* c, const double* u, size_t cell, double* J) {
    if (!c || !c->have_mech) return fail(c, RS_E_STATE, "mechanism not set");
    std::vector<double> uc(c->ns);
    gather(*c, u, cell, uc.data());
    CellChem cc(*c, cell);
    cc.jac(uc.data(), J);
    return RS_OK;
}
int rs_oracle_attempt_log(rs_ctx* c, double* rec, long long cap, long long* n_out) {
    if (!c || !n_out) return RS_E_ARG;
    *n_out = (long long)c->alog.size();
    if (rec) {
        if (cap < (long long)c->alog.size()) return fail(c, RS_E_ARG, "attempt-log buffer too small");
        for (size_t i = 0; i < c->alog.size(); ++i)
            for (int q = 0; q < 4; ++q) rec[4 * i + q] = c->alog[i][q];
    }
    return RS_OK;
}

int rs_oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

// ---- autotuner and ensembles: the oracle has nothing to tune and runs the
// members one after the other (the CPU baseline of the ensemble workload) ----
struct rs_ensemble {
    std::vector<rs_ctx*> m;
    std::string err;
};

extern "C" {

int rs_autotune(rs_ctx* c, rs_tune_result* out) {
    if (!c) return RS_E_ARG;
    if (out) { *out = rs_tune_result{}; out->transport_tile_levels = 0; out->chem_block = 1; }
    return RS_OK;
}

int rs_ensemble_create(rs_ensemble** out, rs_ctx* const* members, int n) {
    if (!out || !members || n < 1) return RS_E_ARG;
    auto* e = new rs_ensemble();
    for (int d = 0; d < n; ++d) {
        if (!members[d]) { delete e; return RS_E_ARG; }
        e->m.push_back(members[d]);
    }
    *out = e;
    return RS_OK;
}

int rs_ensemble_destroy(rs_ensemble* e) { delete e; return RS_OK; }

const char* rs_ensemble_last_error(const rs_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int rs_ensemble_autotune(rs_ensemble* e, rs_tune_result* out) {
    if (!e) return RS_E_ARG;
    for (size_t d = 0; d < e->m.size(); ++d) rs_autotune(e->m[d], out ? out + d : nullptr);
    return RS_OK;
}

int rs_ensemble_adjoint_run(rs_ensemble* e, int nmacro, const double* tstops, double dtT0,
                            double dtC0, const rs_split_opts* o, int loss_kind, int loss_species,
                            double tape_budget_gib, double* J_out, rs_adjoint_stats* stats,
                            rs_ensemble_stats* es) {
    (void)tape_budget_gib;
    if (!e) return RS_E_ARG;
    rs_ensemble_stats ES{};
    for (size_t d = 0; d < e->m.size(); ++d) {
        rs_adjoint_stats st{};
        double J = 0.0;
        int rc = rs_adjoint_run(e->m[d], nmacro, tstops, dtT0, dtC0, o, loss_kind, loss_species,
                                &J, nullptr, nullptr, &st);
        if (rc) { e->err = "member " + std::to_string(d) + ": " + e->m[d]->err; return rc; }
        if (J_out) J_out[d] = J;
        if (stats) stats[d] = st;
        ES.seconds_forward += st.seconds_forward;
        ES.seconds_backward += st.seconds_backward;
    }
    ES.backward_groups = (int)e->m.size();
    if (es) *es = ES;
    return RS_OK;
}

}  // extern "C"
