// ppm.hip -- monotone PPM transport stages in CWC mixing-ratio form, gfx950.
//
// The reference's own transport (reseact.esm:1536-3100; regional prototype
// prototypes/transport_3d/transport_3d.esm): flux-form PPM with the
// Colella-Woodward (1984) monotone slopes (eq. 1.7-1.8), edge interpolation
// (eq. 1.6: uniform in longitude and latitude, width-weighted in the
// vertical) and parabola limiter (eq. 1.10), in method-of-lines form: the
// face value is the limited edge of the upwind cell.
//   species 0        air mass m:      dm/dt = -div(M)
//   species 1..NT-1  mixing ratio q:  dq/dt = (q div(M) - div(M q_f)) / m
// Lateral open walls read 3 halo cells of prescribed inflow mixing ratio;
// the vertical clamps its stencil at the surface and the lid (Mz = 0 there).
//
// A stage workgroup walks one species up a tile of columns: the horizontal
// faces of each level are reconstructed once each from an LDS copy of the
// tile, the vertical ones once each from the lane's sliding column window.
// The width-only vertical factors are tabulated per column when the fluxes
// are set (rs_ppm_coefs), so a stage does no divisions besides the final one
// by m.  Built with -ffp-contract=off: every statement rounds as the
// oracle's, so stages agree with it bit for bit.

#include <hip/hip_runtime.h>


#include "rs_launch.h"

namespace {

constexpr int PB = 256;                 // lanes per workgroup
constexpr double C13 = 1.0 / 3.0;
constexpr double C23 = 2.0 / 3.0;
constexpr double C16 = 1.0 / 6.0;

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ int wrapi(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// wrapi for -n <= i < 2n, without the integer division
__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// eq. 1.7-1.8, uniform spacing
__device__ __forceinline__ double slope_u(double am, double a0, double ap) {
    const double dl = a0 - am, dr = ap - a0;
    if (dr * dl > 0.0) {
        const double da = 0.5 * (ap - am);
        const double m = dmin(fabs(da), dmin(2.0 * fabs(dl), 2.0 * fabs(dr)));
        return copysign(m, da);
    }
    return 0.0;
}

// eq. 1.6, uniform spacing: edge between cells with means a0, a1 and slopes d0, d1
__device__ __forceinline__ double edge_u(double a0, double a1, double d0, double d1) {
    const double e = a0 + 0.5 * (a1 - a0);
    return e - (d1 - d0) * C16;
}

// eq. 1.7-1.8 with the tabulated width factors (cR, cL) of the cell
__device__ __forceinline__ double slope_w(double cR, double cL, double am, double a0, double ap) {
    const double dl = a0 - am, dr = ap - a0;
    if (dr * dl > 0.0) {
        const double da = cR * dr + cL * dl;
        const double m = dmin(fabs(da), dmin(2.0 * fabs(dl), 2.0 * fabs(dr)));
        return copysign(m, da);
    }
    return 0.0;
}

// eq. 1.6 with the tabulated width factors (cE, cF, cG) of the edge
__device__ __forceinline__ double edge_w(const double* c, double a0, double a1, double d0,
                                         double d1) {
    const double da = a1 - a0;
    const double e0 = a0 + c[0] * da;
    return (e0 - c[1] * d1) + c[2] * d0;
}

// eq. 1.10 on a cell with mean a0 and edges aL, aR
__device__ __forceinline__ void limit(double a0, double& aL, double& aR) {
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

// Uniform face value from v[0..5] = cells f-3 .. f+2 around face f: the
// limited right edge of cell f-1 for M >= 0, the limited left edge of cell f
// otherwise.  Both cases are the same statements on the five cells around
// the upwind cell (w = v[0..4] or v[1..5]), so the window is selected first
// and one reconstruction serves lanes of either flux sign (no divergent
// branches); every value rounds as in the two-branch form of the oracle.
__device__ __forceinline__ double face_u(const double* v, double M) {
    const bool pos = M >= 0.0;
    const double w0 = pos ? v[0] : v[1], w1 = pos ? v[1] : v[2], w2 = pos ? v[2] : v[3];
    const double w3 = pos ? v[3] : v[4], w4 = pos ? v[4] : v[5];
    const double s1 = slope_u(w0, w1, w2);
    const double s2 = slope_u(w1, w2, w3);
    const double s3 = slope_u(w2, w3, w4);
    double aL = edge_u(w1, w2, s1, s2);
    double aR = edge_u(w2, w3, s2, s3);
    limit(w2, aL, aR);
    return pos ? aR : aL;
}

__device__ __forceinline__ double div6(double fw, double fe, double fs, double fn, double fb,
                                       double ft, double gx, double ces, double cen, double rdS) {
    const double dx = (fe - fw) * gx;
    const double dy = (cen * fn - ces * fs) * rdS;
    const double dz = ft - fb;
    return (dx + dy) + dz;
}

// Sum over a workgroup of PB lanes in a fixed order; valid in lane 0.
__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (tid == 0)
        for (int q = 0; q < PB / 64; ++q) t += red[q];
    return t;
}

}  // namespace

// Width factors of the vertical reconstruction per column, from the layer
// widths dp with the stencil clamped to the column: slopes of the cells
// x = -2..NLEV+1 (cR, cL) and edges x = -2..NLEV between cells x and x+1
// (cE, cF, cG), each with the oracle's expressions.
__global__ void __launch_bounds__(PB) rs_ppm_coefs(const double* __restrict__ dp,
                                                   double* __restrict__ vs,
                                                   double* __restrict__ ve, int nlon, int nlat,
                                                   int nlev) {
    const long long ncol = (long long)nlon * nlat;
    const long long t = (long long)blockIdx.x * PB + threadIdx.x;
    const long long nrow = (long long)(nlev + 4);
    if (t >= nrow * ncol) return;
    const long long col = t % ncol;
    const int x = (int)(t / ncol) - 2;
    auto H = [&](int kk) { return dp[(long long)clampi(kk, 0, nlev - 1) * ncol + col]; };
    {
        const double hm = H(x - 1), h0 = H(x), hp = H(x + 1);
        const double w = h0 / (hm + h0 + hp);
        const double cR = w * ((2.0 * hm + h0) / (hp + h0));
        const double cL = w * ((h0 + 2.0 * hp) / (hm + h0));
        const long long r = (long long)(x + 2) * ncol + col;
        vs[r] = cR;
        vs[nrow * ncol + r] = cL;
    }
    if (x <= nlev) {
        const long long nrowe = (long long)(nlev + 3);
        const double hm = H(x - 1), h0 = H(x), h1 = H(x + 1), h2 = H(x + 2);
        const double cA = h0 / (h0 + h1);
        const double g = (2.0 * h1 * h0) / (h0 + h1);
        const double t1 = (hm + h0) / (2.0 * h0 + h1);
        const double t2 = (h2 + h1) / (2.0 * h1 + h0);
        const double cS = 1.0 / (((hm + h0) + h1) + h2);
        const double cE = cA + (g * (t1 - t2)) * cS;
        const double cF = (h0 * t1) * cS;
        const double cG = (h1 * ((h1 + h2) / (h0 + 2.0 * h1))) * cS;
        const long long r = (long long)(x + 2) * ncol + col;
        ve[r] = cE;
        ve[nrowe * ncol + r] = cF;
        ve[2 * nrowe * ncol + r] = cG;
    }
}

constexpr int PTX = 64, PTY = 4;
constexpr int PWX = PTX + 6, PWY = PTY + 6;                 // tile + halo
constexpr int NFX = (PTX + 1) * PTY, NFY = PTX * (PTY + 1);  // faces per tile

// One SSPRK43 stage over the transported block.  KIND 0: out = a + h/2 L a
// (and, with outf, the tendency L a beside it, for the adjoint); KIND 2
// (third stage): the same, and out2 = a4 = 2/3 u + 1/3 out (u = b); KIND 1
// (last): unew = a4 + h/2 L a4 (a4 = c4), the error partial from u = a and
// u3 = b, and (positive domain) the count of negative entries; KIND 3:
// out = L a, the tendency alone.
//
// A workgroup is one species on a tile of PTX longitudes x PTY latitudes over
// a block of A.kblk levels, which it walks upwards.  Per level the tile and
// its 3-cell halo are staged in LDS (double-buffered, so two barriers per
// level) and every longitude and latitude face of the tile is reconstructed
// once, by one lane, into a face-flux array.  In the vertical each lane
// carries its column's stencil values, the slope of the cell above, the
// unlimited edge above and the limited parabola of its cell, so a level adds
// one value, one slope, one edge and one limited parabola: each vertical face
// is reconstructed once too.  The lane then takes its cell's divergence from
// the six face fluxes.
// Grid: x = species (the species of a tile run side by side and share its
// fluxes and width factors in cache), y = tile, z = level block.
// PBL: the PBL mixing term is on (A.pblf set), a separate instantiation so
// that the stage without it is unchanged.
template <int KIND, bool PBL>
__global__ void __launch_bounds__(PB) rs_ppm_stage(RsPPMArgs A) {
    __shared__ double red[PB / 64];
    __shared__ double Q[2][PWY][PWX];
    __shared__ double FX[2][PTY][PTX + 1];
    __shared__ double FY[2][PTY + 1][PTX];
    const RsCtrl* ctl = A.ctrl;
    if (ctl && ctl->done) return;
    const int cur = ctl ? ctl->cur : A.cur_static;
    const double dt = ctl ? ctl->dtc : A.dt_static;
    const double h2 = dt / 2;
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev;
    const long long nc = A.nc, plane = (long long)nlon * nlat;
    const double* __restrict__ pa = A.a_is_state ? (cur ? A.S1 : A.S0) : A.a;
    const double* __restrict__ pb = A.b_is_state ? (cur ? A.S1 : A.S0) : A.b;
    const double* __restrict__ ps = (KIND == 1) ? A.c4 : pa;   // stencil source
    double* __restrict__ po = A.out_is_state ? (cur ? A.S0 : A.S1) : A.out;
    const int tid = threadIdx.x;
    const int tx = tid % PTX, ty = tid / PTX;
    const int ntx = (nlon + PTX - 1) / PTX;
    const int s = blockIdx.x;
    const int i0 = (int)(blockIdx.y % ntx) * PTX, j0 = (int)(blockIdx.y / ntx) * PTY;
    const int k0 = blockIdx.z * A.kblk, k1 = min(nlev, k0 + A.kblk);
    const int i = i0 + tx, j = j0 + ty;
    const bool valid = (i < nlon) && (j < nlat);
    const long long col = (long long)j * nlon + i;
    const long long sb = (long long)s * nc;
    const int sq1 = s - 1;
    const long long nrs = (long long)(nlev + 4) * plane, nre = (long long)(nlev + 3) * plane;
    double gx = 0.0, ces = 0.0, cen = 0.0, rdS = 0.0;
    if (valid) {
        gx = A.gx[j];
        ces = A.ce[j];
        cen = A.ce[j + 1];
        rdS = A.rdS[j];
    }
    // value of the lane's column at level x, clamped to the column
    auto V = [&](int x) { return ps[sb + (long long)clampi(x, 0, nlev - 1) * plane + col]; };
    // slope of cell x (factors at row x + 2) and edge between x and x+1 (row x + 2)
    auto SL = [&](int x, double am, double a0, double ap) {
        const long long r = (long long)(x + 2) * plane + col;
        return slope_w(A.vs[r], A.vs[nrs + r], am, a0, ap);
    };
    auto ED = [&](int x, double a0, double a1, double d0, double d1) {
        const long long r = (long long)(x + 2) * plane + col;
        const double c[3] = {A.ve[r], A.ve[nre + r], A.ve[2 * nre + r]};
        return edge_w(c, a0, a1, d0, d1);
    };
    // vertical walk state at level k: flux of face k (fb), limited parabola
    // of cell k (pL, pR), unlimited edge (k, k+1) (e1), slope of k+1 (d1),
    // values of k+1, k+2 (v1, v2), air-mass flux of face k (mb)
    double fb = 0.0, pL = 0.0, pR = 0.0, e1 = 0.0, d1 = 0.0, v1 = 0.0, v2 = 0.0, mb = 0.0;
    if (valid) {
        mb = A.Mz[(long long)k0 * plane + col];
        if (s > 0) {
            double w[6];
#pragma unroll
            for (int d = 0; d < 6; ++d) w[d] = V(k0 - 3 + d);
            const double sm2 = SL(k0 - 2, w[0], w[1], w[2]);
            const double sm1 = SL(k0 - 1, w[1], w[2], w[3]);
            const double s0 = SL(k0, w[2], w[3], w[4]);
            const double sp1 = SL(k0 + 1, w[3], w[4], w[5]);
            double qL = ED(k0 - 2, w[1], w[2], sm2, sm1);
            double qR = ED(k0 - 1, w[2], w[3], sm1, s0);
            pL = qR;
            pR = ED(k0, w[3], w[4], s0, sp1);
            e1 = pR;
            limit(w[2], qL, qR);     // cell k0 - 1
            limit(w[3], pL, pR);     // cell k0
            fb = mb * (mb >= 0.0 ? qR : pL);
            d1 = sp1;
            v1 = w[4];
            v2 = w[5];
        }
    }
    // PBL mean of this species in this column (PBL mixing on)
    const double pm = (PBL && s > 0 && valid) ? A.pblm[(long long)sq1 * plane + col] : 0.0;
    double sq = 0.0, ng = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int bf = (k - k0) & 1;
        const long long cell = (long long)k * plane + col;
        const long long g = sb + cell;
        double f = 0.0, fa = 0.0, ac = 0.0, ft = 0.0, mt = 0.0;
        if (s > 0) {
            // ---- stage the tile and its halo (corners are never read)
            for (int x = tid; x < PWX * PWY; x += PB) {
                const int r = x / PWX, q = x - r * PWX;
                const int gi = i0 - 3 + q, gj = j0 - 3 + r;
                double v = 0.0;
                const bool inlat = gj >= 0 && gj < nlat;
                const bool inlon = gi >= 0 && gi < nlon;
                if (inlat) {
                    const long long rowb = sb + (long long)k * plane + (long long)gj * nlon;
                    if (inlon) v = ps[rowb + gi];
                    else if (A.lon_periodic) {
                        // faces reach cells <= nlon + 2; columns past them stay 0
                        if (gi <= nlon + 2) v = ps[rowb + (nlon >= 3 ? wrap1(gi, nlon) : wrapi(gi, nlon))];
                    }
                    else if (gi < 0) v = A.qw[((long long)sq1 * nlev + k) * nlat + gj];
                    else v = A.qe[((long long)sq1 * nlev + k) * nlat + gj];
                } else if (inlon) {
                    if (gj < 0) v = A.qs[((long long)sq1 * nlev + k) * nlon + gi];
                    else v = A.qn[((long long)sq1 * nlev + k) * nlon + gi];
                }
                Q[bf][r][q] = v;
            }
            __syncthreads();
            for (int x = tid; x < NFX; x += PB) {
                const int fr = x / (PTX + 1), fx = x - fr * (PTX + 1);
                const int gj = j0 + fr, gi = i0 + fx;
                double F = 0.0;
                if (gj < nlat && gi <= nlon) {
                    const double M = A.Mx[((long long)k * nlat + gj) * (nlon + 1) + gi];
                    F = M * face_u(&Q[bf][fr + 3][fx], M);
                }
                FX[bf][fr][fx] = F;
            }
            for (int x = tid; x < NFY; x += PB) {
                const int fy = x / PTX, fc = x - fy * PTX;
                const int gj = j0 + fy, gi = i0 + fc;
                double F = 0.0;
                if (gj <= nlat && gi < nlon) {
                    const double M = A.My[((long long)k * (nlat + 1) + gj) * nlon + gi];
                    double v[6];
#pragma unroll
                    for (int d = 0; d < 6; ++d) v[d] = Q[bf][fy + d][fc + 3];
                    F = M * face_u(v, M);
                }
                FY[bf][fy][fc] = F;
            }
            if (valid) {
                // ---- face k+1: one new value, slope, edge and parabola
                mt = A.Mz[(long long)(k + 1) * plane + col];
                const double v3 = V(k + 3);
                const double d2 = SL(k + 2, v1, v2, v3);
                const double e2 = ED(k + 1, v1, v2, d1, d2);
                double nL = e1, nR = e2;
                limit(v1, nL, nR);       // cell k + 1
                ft = mt * (mt >= 0.0 ? pR : nL);
                pL = nL;
                pR = nR;
                e1 = e2;
                d1 = d2;
                v1 = v2;
                v2 = v3;
            }
            __syncthreads();
        } else if (valid) {
            mt = A.Mz[(long long)(k + 1) * plane + col];
        }
        if (valid) {
            const double* mxr = A.Mx + ((long long)k * nlat + j) * (nlon + 1);
            const double* myr = A.My + ((long long)k * (nlat + 1) + j) * nlon + i;
            const double divM = div6(mxr[i], mxr[i + 1], myr[0], myr[nlon], mb, mt, gx, ces, cen, rdS);
            if (s == 0) {
                f = A.msrc ? -divM + A.msrc[cell] : -divM;
                fa = f;
                ac = ps[cell];
            } else {
                const double divF = div6(FX[bf][ty][tx], FX[bf][ty][tx + 1], FY[bf][ty][tx],
                                         FY[bf][ty + 1][tx], fb, ft, gx, ces, cen, rdS);
                const double qc = Q[bf][ty + 3][tx + 3];
                const double num = qc * divM - divF;
                fa = num / ps[cell];
                f = fa;
                if (PBL) {
                    // PBL mixing, pbl_f (pbl_mean - q) / tau (reseact.esm:1563-1690)
                    const double t = A.pblf[cell] * (pm - qc);
                    f = fa + t / A.pbl_tau;
                }
                ac = qc;
            }
            const double du = h2 * f;
            const double un = ac + du;
            if (KIND == 3) {
                po[g] = fa;                // the advective tendency (the adjoint's m row)
            } else if (KIND == 1) {
                const double u = pa[g], u3 = pb[g];
                const double a = C13 * u;
                const double b = C23 * u3;
                double ut = a + b;
                const double ud = ut - un;
                ut = 0.5 * ud;
                const double am = dmax(fabs(u), fabs(un));
                double den = am * A.rtol;
                den = den + A.atol;
                const double at = ut / den;
                sq += at * at;
                if (A.posdom && un < 0.0) ng += 1.0;
                po[g] = A.clamp ? dmax(un, 0.0) : un;
                if (A.outf) A.outf[g] = fa;    // the tendency at a4, kept for a step VJP
            } else {
                po[g] = un;
                if (A.outf) A.outf[g] = fa;
                if (KIND == 2) {
                    const double a = C23 * pb[g];
                    const double b = C13 * un;
                    A.out2[g] = a + b;
                }
            }
            fb = ft;
            mb = mt;
        }
    }
    if (KIND == 1) {
        const double t = block_sum(sq, red, tid);
        const long long bid = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        if (tid == 0) A.partial[bid] = t;
        if (A.negpart) {
            const double n = block_sum(ng, red, tid);
            if (tid == 0) A.negpart[bid] = n;
        }
    }
}

// Air-mass source of the column pressure fixer, one lane per column:
// (dPSdt + divh_col) dp / dp_col with divh_col and dp_col summed in level order.
__global__ void __launch_bounds__(PB) rs_ppm_source(const double* __restrict__ Mx,
                                                    const double* __restrict__ My,
                                                    const double* __restrict__ dp,
                                                    const double* __restrict__ gxa,
                                                    const double* __restrict__ ce,
                                                    const double* __restrict__ rdSa,
                                                    const double* __restrict__ dPSdt,
                                                    double* __restrict__ msrc, int nlon, int nlat,
                                                    int nlev) {
    const long long plane = (long long)nlon * nlat;
    const long long col = (long long)blockIdx.x * PB + threadIdx.x;
    if (col >= plane) return;
    const int i = (int)(col % nlon), j = (int)(col / nlon);
    const double gx = gxa[j], rdS = rdSa[j], ces = ce[j], cen = ce[j + 1];
    double dcol = 0.0, pcol = 0.0;
    for (int k = 0; k < nlev; ++k) {
        const double* mxr = Mx + ((long long)k * nlat + j) * (nlon + 1);
        const double ms = My[((long long)k * (nlat + 1) + j) * nlon + i];
        const double mn = My[((long long)k * (nlat + 1) + j + 1) * nlon + i];
        const double dx = (mxr[i + 1] - mxr[i]) * gx;
        const double dy = (cen * mn - ces * ms) * rdS;
        dcol += dx + dy;
        pcol += dp[(long long)k * plane + col];
    }
    const double a = (dPSdt[col] + dcol) / pcol;
    for (int k = 0; k < nlev; ++k) msrc[(long long)k * plane + col] = a * dp[(long long)k * plane + col];
}

// Vertical air-mass flux from continuity with the column pressure fixer
// (reseact.esm:262-291), one lane per column: divh(k) as in rs_ppm_source,
// a = (dPSdt + divh_col) / dp_col, divh_fix(k) = divh(k) - a dp(k), and the
// exclusive scan Mz[k+1] = Mz[k] - (dB[k] dPSdt + divh_fix(k)); the surface
// and lid faces are written 0 and the scan's value at the lid goes to
// lid[col].  dB and dPSdt may be null (zeros).
__global__ void __launch_bounds__(PB) rs_ppm_diag_mz(const double* __restrict__ Mx,
                                                     const double* __restrict__ My,
                                                     const double* __restrict__ dp,
                                                     const double* __restrict__ gxa,
                                                     const double* __restrict__ ce,
                                                     const double* __restrict__ rdSa,
                                                     const double* __restrict__ dB,
                                                     const double* __restrict__ dPSdt,
                                                     double* __restrict__ Mz,
                                                     double* __restrict__ lid, int nlon,
                                                     int nlat, int nlev) {
    const long long plane = (long long)nlon * nlat;
    const long long col = (long long)blockIdx.x * PB + threadIdx.x;
    if (col >= plane) return;
    const int i = (int)(col % nlon), j = (int)(col / nlon);
    const double gx = gxa[j], rdS = rdSa[j], ces = ce[j], cen = ce[j + 1];
    auto divh = [&](int k) {
        const double* mxr = Mx + ((long long)k * nlat + j) * (nlon + 1);
        const double ms = My[((long long)k * (nlat + 1) + j) * nlon + i];
        const double mn = My[((long long)k * (nlat + 1) + j + 1) * nlon + i];
        const double dx = (mxr[i + 1] - mxr[i]) * gx;
        const double dy = (cen * mn - ces * ms) * rdS;
        return dx + dy;
    };
    double dcol = 0.0, pcol = 0.0;
    for (int k = 0; k < nlev; ++k) {
        dcol += divh(k);
        pcol += dp[(long long)k * plane + col];
    }
    const double ps = dPSdt ? dPSdt[col] : 0.0;
    const double a = (ps + dcol) / pcol;
    double mz = 0.0;
    Mz[col] = 0.0;
    for (int k = 0; k < nlev; ++k) {
        const double fix = divh(k) - a * dp[(long long)k * plane + col];
        const double db = dB ? dB[k] : 0.0;
        const double t = db * ps + fix;
        mz = mz - t;
        if (k + 1 < nlev) Mz[(long long)(k + 1) * plane + col] = mz;
    }
    Mz[(long long)nlev * plane + col] = 0.0;
    lid[col] = mz;
}

// ---- adjoint of the tendency ------------------------------------------------
// Derivatives on the branch the forward takes at the linearisation state (the
// limiters switch on products of neighbour differences, so the tendency is
// piecewise smooth; these are its derivatives almost everywhere).
//
// Local gradient of a face value over the five cells w[0..4] = the upwind
// cell's neighbourhood c-2 .. c+2 (the six-cell stencil shifted by one for
// M < 0); `right`: the value is the right edge of w[2] (M >= 0), else its left
// edge.  Width factors (NU): slopes of w[1..3] (sR, sL), edges (w1, w2) and
// (w2, w3) (e1, e2).  Branch-free in the flux sign.
template <bool NU>
__device__ __forceinline__ double face_grad5(const double* w, bool right, const double* sR,
                                             const double* sL, const double* e1, const double* e2,
                                             double* g) {
    double sv[3], sg[3][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double am = w[q], a0 = w[q + 1], ap = w[q + 2];
        const double cR = NU ? sR[q] : 0.5, cL = NU ? sL[q] : 0.5;
        sv[q] = NU ? slope_w(cR, cL, am, a0, ap) : slope_u(am, a0, ap);
        const double dl = a0 - am, dr = ap - a0;
        const bool on = dr * dl > 0.0;
        const double da = NU ? cR * dr + cL * dl : 0.5 * (ap - am);
        const double ax = fabs(da), y = 2.0 * fabs(dl), z = 2.0 * fabs(dr);
        const bool zl = z < y;
        const double in = zl ? z : y;
        const bool lim = in < ax;
        // gradient of the chosen term: da (cR, cL), 2 dr or 2 dl
        double g0 = -cL, g1 = cL - cR, g2 = cR;
        if (lim) { g0 = zl ? 0.0 : -2.0; g1 = 2.0 * (zl ? -1.0 : 1.0); g2 = zl ? 2.0 : 0.0; }
        sg[q][0] = on ? g0 : 0.0;
        sg[q][1] = on ? g1 : 0.0;
        sg[q][2] = on ? g2 : 0.0;
    }
    double aL, aR, eL0, eL1, eL2, eL3, eR0, eR1, eR2, eR3;
    if (NU) {
        aL = edge_w(e1, w[1], w[2], sv[0], sv[1]);
        aR = edge_w(e2, w[2], w[3], sv[1], sv[2]);
        eL0 = 1.0 - e1[0]; eL1 = e1[0]; eL2 = e1[2]; eL3 = -e1[1];
        eR0 = 1.0 - e2[0]; eR1 = e2[0]; eR2 = e2[2]; eR3 = -e2[1];
    } else {
        aL = edge_u(w[1], w[2], sv[0], sv[1]);
        aR = edge_u(w[2], w[3], sv[1], sv[2]);
        eL0 = eR0 = 0.5; eL1 = eR1 = 0.5; eL2 = eR2 = C16; eL3 = eR3 = -C16;
    }
    const double a0 = w[2];
    // limiter branch -> adjoints (bc, bL, bR) of (a0, aL, aR) for the output
    const bool flat = (aR - a0) * (a0 - aL) <= 0.0;
    const double d = aR - aL;
    const double a6 = 6.0 * (a0 - 0.5 * (aL + aR));
    const double dd = d * d;
    const double da6 = d * a6;
    const bool resetL = da6 > dd, resetR = !resetL && (-dd > da6);
    double bc, bL, bR;
    if (flat) { bc = 1.0; bL = 0.0; bR = 0.0; }
    else if (right) { bc = resetR ? 3.0 : 0.0; bL = resetR ? -2.0 : 0.0; bR = resetR ? 0.0 : 1.0; }
    else { bc = resetL ? 3.0 : 0.0; bR = resetL ? -2.0 : 0.0; bL = resetL ? 0.0 : 1.0; }
    // the face value (the forward's limiter)
    double lL = aL, lR = aR;
    limit(a0, lL, lR);
    g[0] = 0.0; g[1] = 0.0; g[2] = bc; g[3] = 0.0; g[4] = 0.0;
    g[1] += bL * eL0;
    g[2] += bL * eL1 + bR * eR0;
    g[3] += bR * eR1;
    const double bs0 = bL * eL2, bs1 = bL * eL3 + bR * eR2, bs2 = bR * eR3;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        g[q] += bs0 * sg[0][q];
        g[q + 1] += bs1 * sg[1][q];
        g[q + 2] += bs2 * sg[2][q];
    }
    return right ? lR : lL;
}

// Local gradient g[d] = d(face value)/d v[d] of one face from its six stencil
// values v[0..5] (cells f-3 .. f+2), for the per-cell gather below: the five
// cells around the upwind cell (v[0..4] for M >= 0, v[1..5] otherwise) go
// through face_grad5.  NU: width-weighted (vertical), with the slope factors
// of stencil cell x at sR[x], sL[x] and the factors of the edge between
// stencil cells x and x+1 at ec[x][0..2].
template <bool NU>
__device__ __forceinline__ void face_grad6(const double* v, double M, const double* sR,
                                           const double* sL, const double (*ec)[3], double* g) {
    const bool right = M >= 0.0;
    double w[5], g5[5], r[3], l[3], e1[3], e2[3];
#pragma unroll
    for (int d = 0; d < 5; ++d) w[d] = right ? v[d] : v[d + 1];
    if (NU) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            r[q] = right ? sR[1 + q] : sR[2 + q];
            l[q] = right ? sL[1 + q] : sL[2 + q];
            e1[q] = right ? ec[1][q] : ec[2][q];
            e2[q] = right ? ec[2][q] : ec[3][q];
        }
    }
    face_grad5<NU>(w, right, r, l, e1, e2, g5);
#pragma unroll
    for (int d = 0; d < 6; ++d) g[d] = right ? (d < 5 ? g5[d] : 0.0) : (d > 0 ? g5[d - 1] : 0.0);
}

// out = [beta b +] x + h2 J(at)^T x, x = alpha a, J the Jacobian of the PPM
// tendency at `at`.  One lane per cell, all transported species of the cell:
//   air mass:  x_m - h2 sum_s x_s f_s / m   (f_s the cell's mixing-ratio
//              tendency at `at`, s in species order)
//   ratio s:   x_s + h2 (mu divM + sum over the faces whose stencil holds the
//              cell of W_f g_f(cell)),  mu = x_s / m,
// with W_f the face's flux times its divergence coefficients in the two cells
// it separates, g_f the face's local gradient (summed over every stencil
// position that maps to the cell: periodic wrap, clamped vertical stencil);
// inflow halo cells are not state and drop out.  A gather: deterministic, no
// atomics.
__global__ void __launch_bounds__(PB) rs_ppm_vjp(RsPPMVjpArgs A) {
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev, nt = A.nt;
    const long long nc = A.nc, plane = (long long)nlon * nlat;
    const long long cell = (long long)blockIdx.x * PB + threadIdx.x;
    if (cell >= nc) return;
    const int i = (int)(cell % nlon), j = (int)((cell / nlon) % nlat), k = (int)(cell / plane);
    const long long col = (long long)j * nlon + i;
    const double* __restrict__ at = A.at;
    const double gx = A.gx[j], ces = A.ce[j], cen = A.ce[j + 1], rdS = A.rdS[j];
    const double* mxr = A.Mx + ((long long)k * nlat + j) * (nlon + 1);
    const double mw = mxr[i], me = mxr[i + 1];
    const double ms = A.My[((long long)k * (nlat + 1) + j) * nlon + i];
    const double mn = A.My[((long long)k * (nlat + 1) + j + 1) * nlon + i];
    const double mb = A.Mz[(long long)k * plane + col];
    const double mt = A.Mz[(long long)(k + 1) * plane + col];
    const double divM = div6(mw, me, ms, mn, mb, mt, gx, ces, cen, rdS);
    const double mval = at[cell];
    const double alpha = A.alpha;
    auto X = [&](const double* p, long long x) { return alpha == 1.0 ? p[x] : alpha * p[x]; };
    double msum = 0.0;
    const long long nrs = (long long)(nlev + 4) * plane, nre = (long long)(nlev + 3) * plane;
    for (int s = 1; s < nt; ++s) {
        const long long sb = (long long)s * nc;
        const int sq1 = s - 1;
        const double* q = at + sb;
        const double* la = A.a + sb;
        // mu of a state cell (x / m)
        auto mu = [&](long long x) { return X(la, x) / at[x]; };
        // ---- longitude: values with wrap or inflow halo; in-domain flag
        auto qx = [&](int ii, bool* st) {
            if (A.lon_periodic) { *st = true; return q[(long long)k * plane + (long long)j * nlon + wrapi(ii, nlon)]; }
            if (ii < 0) { *st = false; return A.qw[((long long)sq1 * nlev + k) * nlat + j]; }
            if (ii >= nlon) { *st = false; return A.qe[((long long)sq1 * nlev + k) * nlat + j]; }
            *st = true;
            return q[(long long)k * plane + (long long)j * nlon + ii];
        };
        auto qy = [&](int jj, bool* st) {
            if (jj < 0) { *st = false; return A.qs[((long long)sq1 * nlev + k) * nlon + i]; }
            if (jj >= nlat) { *st = false; return A.qn[((long long)sq1 * nlev + k) * nlon + i]; }
            *st = true;
            return q[(long long)k * plane + (long long)jj * nlon + i];
        };
        double gsum = mu(cell) * divM;
        double v[6], g[6];
        bool in[6];
        // own faces' values for the cell's tendency (forward expressions)
        double fw_ = 0.0, fe_ = 0.0, fs_ = 0.0, fn_ = 0.0, fb_ = 0.0, ft_ = 0.0;
        // ---- longitude faces: residues i-2..i+3 (all residues if nlon < 6),
        // face index F in [0, nlon] (face nlon also for residue 0)
        {
            const bool allres = A.lon_periodic && nlon < 6;
            const int nloop = allres ? nlon : 6;
            for (int o = 0; o < nloop; ++o) {
                int r;
                if (allres) r = o;
                else if (A.lon_periodic) r = wrapi(i - 2 + o, nlon);
                else r = i - 2 + o;
                for (int rep = 0; rep < 2; ++rep) {
                    int F;
                    if (rep == 0) F = r;
                    else if (A.lon_periodic && r == 0) F = nlon;
                    else break;
                    if (F < 0 || F > nlon) continue;
                    const double M = mxr[F];
                    bool hit = false;
#pragma unroll
                    for (int d = 0; d < 6; ++d) {
                        v[d] = qx(F - 3 + d, &in[d]);
                        const int cc = A.lon_periodic ? wrapi(F - 3 + d, nlon) : F - 3 + d;
                        in[d] = in[d] && cc == i;
                        hit = hit || in[d];
                    }
                    if (F == i) fw_ = M * face_u(v, M);
                    if (F == i + 1) fe_ = M * face_u(v, M);
                    if (!hit || M == 0.0) continue;
                    face_grad6<false>(v, M, nullptr, nullptr, nullptr, g);
                    double gs = 0.0;
#pragma unroll
                    for (int d = 0; d < 6; ++d) if (in[d]) gs += g[d];
                    const long long rowb = (long long)k * plane + (long long)j * nlon;
                    const double mr = F < nlon ? mu(rowb + F) : 0.0;
                    const double ml = F >= 1 ? mu(rowb + F - 1) : 0.0;
                    gsum += (gx * M) * (mr - ml) * gs;
                }
            }
        }
        // ---- latitude faces F in [j-2, j+3] n [0, nlat]
        for (int o = 0; o < 6; ++o) {
            const int F = j - 2 + o;
            if (F < 0 || F > nlat) continue;
            const double M = A.My[((long long)k * (nlat + 1) + F) * nlon + i];
            bool hit = false;
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                v[d] = qy(F - 3 + d, &in[d]);
                in[d] = in[d] && (F - 3 + d == j);
                hit = hit || in[d];
            }
            if (F == j) fs_ = M * face_u(v, M);
            if (F == j + 1) fn_ = M * face_u(v, M);
            if (!hit || M == 0.0) continue;
            face_grad6<false>(v, M, nullptr, nullptr, nullptr, g);
            double gs = 0.0;
#pragma unroll
            for (int d = 0; d < 6; ++d) if (in[d]) gs += g[d];
            const double cf = A.ce[F];
            const double mr = F < nlat ? mu((long long)k * plane + (long long)F * nlon + i) * A.rdS[F] : 0.0;
            const double ml = F >= 1 ? mu((long long)k * plane + (long long)(F - 1) * nlon + i) * A.rdS[F - 1] : 0.0;
            gsum += (cf * M) * (mr - ml) * gs;
        }
        // ---- vertical faces F in [k-2, k+3] n [0, nlev], stencil clamped
        {
            double sR[6], sL[6], ec[6][3];
            for (int o = 0; o < 6; ++o) {
                const int F = k - 2 + o;
                if (F < 0 || F > nlev) continue;
                const double M = A.Mz[(long long)F * plane + col];
                bool hit = false;
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    const int L = F - 3 + d;                  // logical level
                    const int kk = clampi(L, 0, nlev - 1);
                    v[d] = q[(long long)kk * plane + col];
                    in[d] = kk == k;
                    hit = hit || in[d];
                    // slope factors of logical cell L (row L + 2), edge (L, L+1) (row L + 2)
                    const int rs = clampi(L + 2, 0, nlev + 3), re = clampi(L + 2, 0, nlev + 2);
                    sR[d] = A.vs[(long long)rs * plane + col];
                    sL[d] = A.vs[nrs + (long long)rs * plane + col];
                    ec[d][0] = A.ve[(long long)re * plane + col];
                    ec[d][1] = A.ve[nre + (long long)re * plane + col];
                    ec[d][2] = A.ve[2 * nre + (long long)re * plane + col];
                }
                if (F == k || F == k + 1) {
                    // the forward's vertical face value (rs_ppm_stage fz)
                    const int c = M >= 0.0 ? 2 : 3;
                    const double dm = slope_w(sR[c - 1], sL[c - 1], v[c - 2], v[c - 1], v[c]);
                    const double d0 = slope_w(sR[c], sL[c], v[c - 1], v[c], v[c + 1]);
                    const double dp = slope_w(sR[c + 1], sL[c + 1], v[c], v[c + 1], v[c + 2]);
                    double aL = edge_w(ec[c - 1], v[c - 1], v[c], dm, d0);
                    double aR = edge_w(ec[c], v[c], v[c + 1], d0, dp);
                    limit(v[c], aL, aR);
                    const double fv = M * (M >= 0.0 ? aR : aL);
                    if (F == k) fb_ = fv;
                    else ft_ = fv;
                }
                if (!hit || M == 0.0) continue;
                face_grad6<true>(v, M, sR, sL, ec, g);
                double gs = 0.0;
#pragma unroll
                for (int d = 0; d < 6; ++d) if (in[d]) gs += g[d];
                const double mr = F < nlev ? mu((long long)F * plane + col) : 0.0;
                const double ml = F >= 1 ? mu((long long)(F - 1) * plane + col) : 0.0;
                gsum += M * (mr - ml) * gs;
            }
        }
        const double xs = X(la, cell);
        const double divF = div6(fw_, fe_, fs_, fn_, fb_, ft_, gx, ces, cen, rdS);
        const double num = q[cell] * divM - divF;
        const double fsv = num / mval;
        const double t = xs * fsv;
        msum -= t / mval;
        double o = xs + A.h2 * gsum;
        if (A.b) o = A.beta * A.b[sb + cell] + o;
        A.out[sb + cell] = o;
    }
    double o = X(A.a, cell) + A.h2 * msum;
    if (A.b) o = A.beta * A.b[cell] + o;
    A.out[cell] = o;
}

// ---- adjoint of the tendency as three line sweeps --------------------------
// The same J^T x as rs_ppm_vjp, with each face gradient computed once.  In
// the vertical and in latitude a lane walks one grid line of one mixing
// ratio face by face, and adds the face's contribution W_f g_f[d] to the six
// cells of its stencil, held in a register window of "logical" cells f-3 ..
// f+2 that slides by one per face.  A cell leaves the window complete (no
// later face's stencil reaches back to it) and is written once:
// deterministic, no atomics.  Logical cells outside the line are inflow halos
// (dropped) or clamped vertical indices (added to the end cell).  Longitude
// rows (strided across lanes) run on LDS tiles instead (rs_ppm_vjp_xtile).
// Order: z, then y (adding), then x (finishing):
//   G   = mu divM + acc_z;   G += acc_y;   out = [beta b +] x + h2 (G + acc_x)
// and the air-mass row from the tendency F = L(at):
//   out_m = [beta b_m +] x_m - h2 sum_s x_s F_s / m.

// Face divergence per cell, div(M) in the operator's order (species-independent).
__global__ void __launch_bounds__(PB) rs_ppm_divm(RsPPMVjpArgs A, double* __restrict__ divm) {
    const int nlon = A.nlon, nlat = A.nlat;
    const long long plane = (long long)nlon * nlat;
    const long long cell = (long long)blockIdx.x * PB + threadIdx.x;
    if (cell >= A.nc) return;
    const int i = (int)(cell % nlon), j = (int)((cell / nlon) % nlat), k = (int)(cell / plane);
    const long long col = (long long)j * nlon + i;
    const double* mxr = A.Mx + ((long long)k * nlat + j) * (nlon + 1);
    divm[cell] = div6(mxr[i], mxr[i + 1], A.My[((long long)k * (nlat + 1) + j) * nlon + i],
                      A.My[((long long)k * (nlat + 1) + j + 1) * nlon + i], A.Mz[(long long)k * plane + col],
                      A.Mz[(long long)(k + 1) * plane + col], A.gx[j], A.ce[j], A.ce[j + 1], A.rdS[j]);
}

// One line sweep.  DIR 2: vertical columns (lane = species-1, column), writes
// G = mu divM + acc; DIR 1: latitude columns (lane = species-1, level, lon),
// adds to G.  Consecutive lanes are consecutive longitudes: coalesced.
template <int DIR>
__global__ void __launch_bounds__(PB) rs_ppm_vjp_sweep(RsPPMVjpArgs A, const double* __restrict__ divm,
                                                       double* __restrict__ G) {
    static_assert(DIR == 1 || DIR == 2, "latitude or vertical sweep");
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev, nq = A.nt - 1;
    const long long nc = A.nc, plane = (long long)nlon * nlat;
    const long long lane = (long long)blockIdx.x * PB + threadIdx.x;
    const int n = DIR == 1 ? nlat : nlev;                             // cells on the line
    // lines per species; a wave takes 64 consecutive lines of one species and
    // the next waves the same lines of the next species, so the species of a
    // line run side by side and share its fluxes and width factors in cache
    const long long per = DIR == 1 ? (long long)nlev * nlon : plane;
    const long long blk = lane / 64;
    const long long idx = (blk / nq) * 64 + lane % 64;
    if (idx >= per) return;
    const int s = 1 + (int)(blk % nq);
    int k = 0, i = 0;
    long long col = 0;
    if (DIR == 2) {
        col = idx;
    } else {
        k = (int)(idx / nlon);
        i = (int)(idx % nlon);
    }
    const long long sb = (long long)s * nc;
    const double* __restrict__ q = A.at + sb;
    const double* __restrict__ la = A.a + sb;
    const double* __restrict__ m = A.at;
    const int sq1 = s - 1;
    // cell index of position x (0 <= x < n) on the line
    auto cidx = [&](int x) -> long long {
        if (DIR == 2) return (long long)x * plane + col;
        return (long long)k * plane + (long long)x * nlon + i;
    };
    // value of logical position x (inflow halo or clamp)
    auto val = [&](int x) -> double {
        if (DIR == 2) return q[cidx(clampi(x, 0, nlev - 1))];
        if (x < 0) return A.qs[((long long)sq1 * nlev + k) * nlon + i];
        if (x >= nlat) return A.qn[((long long)sq1 * nlev + k) * nlon + i];
        return q[cidx(x)];
    };
    const double alpha = A.alpha;
    auto mu = [&](int x) -> double {          // x in [0, n)
        const long long c = cidx(x);
        const double xv = alpha == 1.0 ? la[c] : alpha * la[c];
        return xv / m[c];
    };
    auto flux = [&](int F) -> double {
        if (DIR == 2) return A.Mz[(long long)F * plane + col];
        return A.My[((long long)k * (nlat + 1) + F) * nlon + i];
    };
    // divergence coefficient of face F in cell F (its lower face) and in
    // cell F-1 (its upper face)
    auto coefR = [&](int F) -> double { return DIR == 2 ? 1.0 : A.ce[F] * A.rdS[F]; };
    auto coefL = [&](int F) -> double { return DIR == 2 ? 1.0 : A.ce[F] * A.rdS[F - 1]; };
    const long long nrs = (long long)(nlev + 4) * plane, nre = (long long)(nlev + 3) * plane;
    // sliding windows: v[d] = value of logical F-3+d, acc[d] = contributions
    // gathered so far for logical F-3+d; vertical width factors (DIR 2) of the
    // slopes of logical cells F-2+d (sw) and of the edges (F-2+d, F-1+d) (ew),
    // rows L + 2 clamped to their tables (rows past the lid are never used)
    double v[6], acc[6], bot = 0.0;
    double swR[5], swL[5], ew[4][3];
#pragma unroll
    for (int d = 0; d < 6; ++d) { v[d] = val(-3 + d); acc[d] = 0.0; }
    auto slope_row = [&](int L, double* r0, double* r1) {
        const long long r = (long long)clampi(L + 2, 0, nlev + 3) * plane + col;
        *r0 = A.vs[r];
        *r1 = A.vs[nrs + r];
    };
    auto edge_row = [&](int L, double* e) {
        const long long r = (long long)clampi(L + 2, 0, nlev + 2) * plane + col;
        e[0] = A.ve[r];
        e[1] = A.ve[nre + r];
        e[2] = A.ve[2 * nre + r];
    };
    if (DIR == 2) {
#pragma unroll
        for (int d = 0; d < 5; ++d) slope_row(-2 + d, &swR[d], &swL[d]);
#pragma unroll
        for (int d = 0; d < 4; ++d) edge_row(-2 + d, ew[d]);
    }
    double mul = 0.0, mur = mu(0);            // mu of cells F-1 and F (0 outside the line)
    // the logical cell leaving the window: write it, fold it (clamped
    // vertical indices below the surface) or drop it (inflow halo)
    auto emit = [&](int L, double a) {
        if (DIR == 2) {
            if (L < 0) { bot += a; return; }
            if (L == 0) a = a + bot;
        } else if (L < 0 || L >= n) {
            return;
        }
        const long long c = cidx(L);
        if (DIR == 2) G[sb - nc + c] = mu(L) * divm[c] + a;
        else G[sb - nc + c] = G[sb - nc + c] + a;
    };
    for (int F = 0; F <= n; ++F) {
        const double M = flux(F);
        if (M != 0.0) {
            const bool right = M >= 0.0;
            double w[5], g[5];
#pragma unroll
            for (int d = 0; d < 5; ++d) w[d] = right ? v[d] : v[d + 1];
            if (DIR == 2) {
                // slopes of logical cells c-1..c+1 and edges (c-1, c), (c, c+1),
                // c = F-1 (M >= 0: window positions 0..2, 0..1) or F (1..3, 1..2)
                double sR[3], sL[3], e1[3], e2[3];
#pragma unroll
                for (int q3 = 0; q3 < 3; ++q3) {
                    sR[q3] = right ? swR[q3] : swR[q3 + 1];
                    sL[q3] = right ? swL[q3] : swL[q3 + 1];
                    e1[q3] = right ? ew[0][q3] : ew[1][q3];
                    e2[q3] = right ? ew[1][q3] : ew[2][q3];
                }
                face_grad5<true>(w, right, sR, sL, e1, e2, g);
            } else {
                face_grad5<false>(w, right, nullptr, nullptr, nullptr, nullptr, g);
            }
            const double W = M * ((F < n ? mur * coefR(F) : 0.0) - (F >= 1 ? mul * coefL(F) : 0.0));
#pragma unroll
            for (int d = 0; d < 5; ++d) {
                const double t = W * g[d];
                if (right) acc[d] += t;
                else acc[d + 1] += t;
            }
        }
        if (F == n) break;
        emit(F - 3, acc[0]);
#pragma unroll
        for (int d = 0; d < 5; ++d) { v[d] = v[d + 1]; acc[d] = acc[d + 1]; }
        v[5] = val(F + 3);
        acc[5] = 0.0;
        mul = mur;
        mur = F + 1 < n ? mu(F + 1) : 0.0;
        if (DIR == 2) {
#pragma unroll
            for (int d = 0; d < 4; ++d) { swR[d] = swR[d + 1]; swL[d] = swL[d + 1]; }
            slope_row(F + 3, &swR[4], &swL[4]);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int q3 = 0; q3 < 3; ++q3) ew[d][q3] = ew[d + 1][q3];
            edge_row(F + 2, ew[3]);
        }
    }
    // the window holds logical n-3 .. n+2, all complete
    if (DIR == 2) {
        // clamped tops: logical n .. n+2 belong to the top cell
        const double top = (acc[3] + acc[4]) + acc[5];
        for (int d = 0; d < 3; ++d) emit(n - 3 + d, d == 2 ? acc[d] + top : acc[d]);
    } else {
        for (int d = 0; d < 6; ++d) emit(n - 3 + d, acc[d]);
    }
}

// Longitude part of J^T x on LDS tiles, finishing the output: a workgroup is
// a tile of PTX longitudes x PTY latitudes at one level of one mixing ratio.
// It stages the tile's values (5-cell halo: wrap or inflow) and mu (3-cell
// halo, 0 outside an open row), computes each longitude face of the logical
// range i0-2 .. i0+PTX+2 once (W_f g_f over its six stencil cells, into LDS),
// then each cell sums the slots of its six faces:
//   out = [beta b +] x + h2 (G + acc_x).
// Periodic rows: logical face FL is the real face FL mod nlon; face 0 (the
// west face of cell 0, flux Mx[0]) also carries face nlon (the east face of
// cell nlon-1, flux Mx[nlon]), whose stencil is the same modulo nlon, so
// every (face, stencil cell) pair is counted once (nlon >= 6).
constexpr int XFW = PTX + 5;                  // logical faces per tile row
__global__ void __launch_bounds__(PB) rs_ppm_vjp_xtile(RsPPMVjpArgs A, double* __restrict__ G) {
    __shared__ double Qx[PTY][PTX + 10];
    __shared__ double MU[PTY][PTX + 6];
    __shared__ double CF[6][PTY][XFW];      // slot-major: lanes read consecutive words
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev;
    const long long nc = A.nc, plane = (long long)nlon * nlat;
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * PTX, j0 = blockIdx.y * PTY;
    const int k = (int)(blockIdx.z % nlev), s = 1 + (int)(blockIdx.z / nlev);
    const int sq1 = s - 1;
    const long long sb = (long long)s * nc;
    const double* __restrict__ q = A.at + sb;
    const double* __restrict__ la = A.a + sb;
    const double alpha = A.alpha;
    const bool periodic = A.lon_periodic;
    // staging: values at columns i0-5 .. i0+PTX+4, mu at i0-3 .. i0+PTX+2
    for (int x = tid; x < PTY * (PTX + 10); x += PB) {
        const int r = x / (PTX + 10), cc = x - r * (PTX + 10);
        const int gj = j0 + r, gi = i0 - 5 + cc;
        double v = 0.0;
        if (gj < nlat) {
            const long long rowb = (long long)k * plane + (long long)gj * nlon;
            if (periodic) { if (gi <= nlon + 4) v = q[rowb + wrap1(gi, nlon)]; }   // faces <= nlon + 2 reach cells <= nlon + 4
            else if (gi < 0) v = A.qw[((long long)sq1 * nlev + k) * nlat + gj];
            else if (gi >= nlon) v = A.qe[((long long)sq1 * nlev + k) * nlat + gj];
            else v = q[rowb + gi];
        }
        Qx[r][cc] = v;
    }
    for (int x = tid; x < PTY * (PTX + 6); x += PB) {
        const int r = x / (PTX + 6), cc = x - r * (PTX + 6);
        const int gj = j0 + r, gi = i0 - 3 + cc;
        double v = 0.0;
        if (gj < nlat && (periodic ? gi <= nlon + 2 : (gi >= 0 && gi < nlon))) {
            const long long c = (long long)k * plane + (long long)gj * nlon + (periodic ? wrap1(gi, nlon) : gi);
            const double xv = alpha == 1.0 ? la[c] : alpha * la[c];
            v = xv / A.at[c];
        }
        MU[r][cc] = v;
    }
    __syncthreads();
    const double* mxb = A.Mx + (long long)k * nlat * (nlon + 1);
    for (int x = tid; x < PTY * XFW; x += PB) {
        const int r = x / XFW, fo = x - r * XFW;
        const int gj = j0 + r, FL = i0 - 2 + fo;         // logical face
        double cf[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (gj < nlat) {
            const double gx = A.gx[gj];
            const double* mxr = mxb + (long long)gj * (nlon + 1);
            // stencil cells FL-3 .. FL+2 sit at Qx[r][fo .. fo+5]; mu of
            // cells FL-1, FL at MU[r][fo], MU[r][fo + 1]
            const double* v = &Qx[r][fo];
            auto add_face = [&](double M, double W) {
                if (M == 0.0 || W == 0.0) return;
                const bool right = M >= 0.0;
                double w[5], g[5];
#pragma unroll
                for (int d = 0; d < 5; ++d) w[d] = right ? v[d] : v[d + 1];
                face_grad5<false>(w, right, nullptr, nullptr, nullptr, nullptr, g);
                // slot d of the six stencil cells: g[d] (M >= 0) or g[d-1]
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    const double gd = right ? (d < 5 ? g[d] : 0.0) : (d > 0 ? g[d - 1] : 0.0);
                    cf[d] += W * gd;
                }
            };
            if (periodic) {
                if (FL <= nlon + 2) {           // cells i < nlon gather faces i-2 .. i+3
                    const int F = wrap1(FL, nlon);
                    const double M = mxr[F];
                    add_face(M, M * ((MU[r][fo + 1] * gx) - (F >= 1 ? MU[r][fo] * gx : 0.0)));
                    if (F == 0) {
                        const double Mn = mxr[nlon];
                        add_face(Mn, Mn * (0.0 - MU[r][fo] * gx));
                    }
                }
            } else if (FL >= 0 && FL <= nlon) {
                const double M = mxr[FL];
                add_face(M, M * ((MU[r][fo + 1] * gx) - (MU[r][fo] * gx)));
            }
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) CF[d][r][fo] = cf[d];
    }
    __syncthreads();
    const int tx = tid % PTX, ty = tid / PTX;
    const int i = i0 + tx, j = j0 + ty;
    if (i >= nlon || j >= nlat) return;
    // faces FL = i-2 .. i+3 are fo = tx .. tx+5; this cell is slot i - FL + 3
    double a = 0.0;
#pragma unroll
    for (int o = 0; o < 6; ++o) a += CF[5 - o][ty][tx + o];
    const long long c = (long long)k * plane + (long long)j * nlon + i;
    const double xv = alpha == 1.0 ? la[c] : alpha * la[c];
    double o = xv + A.h2 * (G[sb - nc + c] + a);
    if (A.b) o = A.beta * A.b[sb + c] + o;
    A.out[sb + c] = o;
}

// out_m = [beta b_m +] x_m - h2 sum_s x_s F_s / m, F the tendency at `at`.
__global__ void __launch_bounds__(PB) rs_ppm_vjp_mrow(RsPPMVjpArgs A, const double* __restrict__ F) {
    const long long cell = (long long)blockIdx.x * PB + threadIdx.x;
    if (cell >= A.nc) return;
    const long long nc = A.nc;
    const double m = A.at[cell];
    const double alpha = A.alpha;
    double msum = 0.0;
    for (int s = 1; s < A.nt; ++s) {
        const long long g = (long long)s * nc + cell;
        const double xs = alpha == 1.0 ? A.a[g] : alpha * A.a[g];
        const double t = xs * F[g];
        msum -= t / m;
    }
    const double xm = alpha == 1.0 ? A.a[cell] : alpha * A.a[cell];
    double o = xm + A.h2 * msum;
    if (A.b) o = A.beta * A.b[cell] + o;
    A.out[cell] = o;
}

// ---- PBL mixing ----------------------------------------------------------
// Weights per column, level order: w = dp pbl_f, Wp = max(sum_k w, 1).
__global__ void __launch_bounds__(PB) rs_ppm_pbl_setup(const double* __restrict__ dp,
                                                       const double* __restrict__ pf,
                                                       double* __restrict__ w, double* __restrict__ Wp,
                                                       int nlev, long long plane) {
    const long long col = (long long)blockIdx.x * PB + threadIdx.x;
    if (col >= plane) return;
    double W = 0.0;
    for (int k = 0; k < nlev; ++k) {
        const long long x = (long long)k * plane + col;
        const double wk = dp[x] * pf[x];
        w[x] = wk;
        W = W + wk;
    }
    Wp[col] = dmax(W, 1.0);
}

// PBL means of a stage's stencil source, per mixing ratio and column:
// pblm = sum_k w_k q_k / Wp (level order).  last: the source is c4.
__global__ void __launch_bounds__(PB) rs_ppm_pbl_mean(RsPPMArgs A, int last) {
    const RsCtrl* ctl = A.ctrl;
    if (ctl && ctl->done) return;
    const int cur = ctl ? ctl->cur : A.cur_static;
    const double* __restrict__ ps = last ? A.c4 : (A.a_is_state ? (cur ? A.S1 : A.S0) : A.a);
    const long long plane = (long long)A.nlon * A.nlat;
    const long long lane = (long long)blockIdx.x * PB + threadIdx.x;
    if (lane >= (long long)(A.nt - 1) * plane) return;
    const long long sq = lane / plane, col = lane % plane;
    const double* q = ps + (sq + 1) * A.nc;
    double S = 0.0;
    for (int k = 0; k < A.nlev; ++k) {
        const long long x = (long long)k * plane + col;
        S = S + A.pblw[x] * q[x];
    }
    A.pblm[lane] = S / A.pblWp[col];
}

// The PBL part of J^T x, added to out (mixing ratios): with x = alpha a,
//   out_j += h2 (w_j L / (Wp tau) - pbl_f_j x_j / tau),  L = sum_k pbl_f_k x_k.
__global__ void __launch_bounds__(PB) rs_ppm_vjp_pbl(RsPPMVjpArgs A) {
    const long long plane = (long long)A.nlon * A.nlat;
    const long long lane = (long long)blockIdx.x * PB + threadIdx.x;
    if (lane >= (long long)(A.nt - 1) * plane) return;
    const long long sb = (lane / plane + 1) * A.nc, col = lane % plane;
    const double alpha = A.alpha;
    double L = 0.0;
    for (int k = 0; k < A.nlev; ++k) {
        const long long x = (long long)k * plane + col;
        const double xv = alpha == 1.0 ? A.a[sb + x] : alpha * A.a[sb + x];
        L = L + A.pblf[x] * xv;
    }
    const double g = L / (A.pblWp[col] * A.pbl_tau);
    for (int k = 0; k < A.nlev; ++k) {
        const long long x = (long long)k * plane + col;
        const double xv = alpha == 1.0 ? A.a[sb + x] : alpha * A.a[sb + x];
        const double t = A.pblw[x] * g - (A.pblf[x] * xv) / A.pbl_tau;
        A.out[sb + x] = A.out[sb + x] + A.h2 * t;
    }
}

hipError_t rs_launch_ppm_pbl_setup(const double* dp, const double* pf, double* w, double* Wp,
                                   int nlon, int nlat, int nlev, hipStream_t st) {
    const long long plane = (long long)nlon * nlat;
    hipLaunchKernelGGL(rs_ppm_pbl_setup, dim3((unsigned)((plane + PB - 1) / PB)), dim3(PB), 0, st,
                       dp, pf, w, Wp, nlev, plane);
    return hipGetLastError();
}

// Level blocks of the column stage: about PKB levels each.
constexpr int PKB = 24;
static int ppm_nkb(int nlev) { return (nlev + PKB - 1) / PKB; }

long long rs_ppm_nblocks(int nlon, int nlat, int nlev, int nt) {
    return (long long)((nlon + PTX - 1) / PTX) * ((nlat + PTY - 1) / PTY) * nt * ppm_nkb(nlev);
}

hipError_t rs_launch_ppm_source(const double* Mx, const double* My, const double* dp,
                                const double* gx, const double* ce, const double* rdS,
                                const double* dPSdt, double* msrc, int nlon, int nlat, int nlev,
                                hipStream_t st) {
    const long long n = (long long)nlon * nlat;
    hipLaunchKernelGGL(rs_ppm_source, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0, st, Mx, My,
                       dp, gx, ce, rdS, dPSdt, msrc, nlon, nlat, nlev);
    return hipGetLastError();
}

hipError_t rs_launch_ppm_diag_mz(const double* Mx, const double* My, const double* dp,
                                 const double* gx, const double* ce, const double* rdS,
                                 const double* dB, const double* dPSdt, double* Mz, double* lid,
                                 int nlon, int nlat, int nlev, hipStream_t st) {
    const long long n = (long long)nlon * nlat;
    hipLaunchKernelGGL(rs_ppm_diag_mz, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0, st, Mx,
                       My, dp, gx, ce, rdS, dB, dPSdt, Mz, lid, nlon, nlat, nlev);
    return hipGetLastError();
}

hipError_t rs_launch_ppm_coefs(const double* dp, double* vs, double* ve, int nlon, int nlat,
                               int nlev, hipStream_t st) {
    const long long n = (long long)(nlev + 4) * nlon * nlat;
    hipLaunchKernelGGL(rs_ppm_coefs, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0, st, dp, vs,
                       ve, nlon, nlat, nlev);
    return hipGetLastError();
}

// kind 0: first and second stages; 2: third stage (also a4); 1: last stage
hipError_t rs_launch_ppm_stage(int kind, const RsPPMArgs* a, hipStream_t st) {
    if (a->nt < 1 || a->nlev < 1) return hipErrorInvalidValue;
    if (a->pblf && a->nt > 1) {
        const long long n = (long long)(a->nt - 1) * a->nlon * a->nlat;
        hipLaunchKernelGGL(rs_ppm_pbl_mean, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0, st, *a,
                           kind == 1 ? 1 : 0);
    }
    RsPPMArgs b = *a;
    const int nkb = ppm_nkb(a->nlev);
    b.kblk = (a->nlev + nkb - 1) / nkb;
    const long long ntile = (long long)((a->nlon + PTX - 1) / PTX) * ((a->nlat + PTY - 1) / PTY);
    const dim3 grid((unsigned)a->nt, (unsigned)ntile, (unsigned)nkb);
    if (b.pblf) {
        if (kind == 1) hipLaunchKernelGGL((rs_ppm_stage<1, true>), grid, dim3(PB), 0, st, b);
        else if (kind == 2) hipLaunchKernelGGL((rs_ppm_stage<2, true>), grid, dim3(PB), 0, st, b);
        else if (kind == 3) hipLaunchKernelGGL((rs_ppm_stage<3, true>), grid, dim3(PB), 0, st, b);
        else hipLaunchKernelGGL((rs_ppm_stage<0, true>), grid, dim3(PB), 0, st, b);
    } else {
        if (kind == 1) hipLaunchKernelGGL((rs_ppm_stage<1, false>), grid, dim3(PB), 0, st, b);
        else if (kind == 2) hipLaunchKernelGGL((rs_ppm_stage<2, false>), grid, dim3(PB), 0, st, b);
        else if (kind == 3) hipLaunchKernelGGL((rs_ppm_stage<3, false>), grid, dim3(PB), 0, st, b);
        else hipLaunchKernelGGL((rs_ppm_stage<0, false>), grid, dim3(PB), 0, st, b);
    }
    return hipGetLastError();
}

static void launch_vjp_pbl(const RsPPMVjpArgs* a, hipStream_t st) {
    if (!a->pblf || a->nt < 2) return;
    const long long n = (long long)(a->nt - 1) * a->nlon * a->nlat;
    hipLaunchKernelGGL(rs_ppm_vjp_pbl, dim3((unsigned)((n + PB - 1) / PB)), dim3(PB), 0, st, *a);
}

hipError_t rs_launch_ppm_vjp(const RsPPMVjpArgs* a, hipStream_t st) {
    hipLaunchKernelGGL(rs_ppm_vjp, dim3((unsigned)((a->nc + PB - 1) / PB)), dim3(PB), 0, st, *a);
    launch_vjp_pbl(a, st);
    return hipGetLastError();
}

// The sweep form of rs_launch_ppm_vjp: F (NT*NC) receives the tendency at
// a->at (computed by the caller with stage kind 3), G (NQ*NC) and divm (NC)
// are scratch.  Periodic rows need nlon >= 6 (the gather kernel otherwise).
hipError_t rs_launch_ppm_vjp_sweeps(const RsPPMVjpArgs* a, const double* F, double* G,
                                    double* divm, hipStream_t st) {
    const long long nq = a->nt - 1;
    if (a->lon_periodic && a->nlon < 6) return hipErrorInvalidValue;
    const long long plane = (long long)a->nlon * a->nlat;
    auto blocks = [](long long n) { return dim3((unsigned)((n + PB - 1) / PB)); };
    if (nq > 0) {
        hipLaunchKernelGGL(rs_ppm_divm, blocks(a->nc), dim3(PB), 0, st, *a, divm);
        // 64-line blocks per species, interleaved over the species
        auto lines = [&](long long per) { return nq * ((per + 63) / 64) * 64; };
        hipLaunchKernelGGL(rs_ppm_vjp_sweep<2>, blocks(lines(plane)), dim3(PB), 0, st, *a,
                           (const double*)divm, G);
        hipLaunchKernelGGL(rs_ppm_vjp_sweep<1>, blocks(lines((long long)a->nlev * a->nlon)), dim3(PB),
                           0, st, *a, (const double*)divm, G);
        const dim3 xg((unsigned)((a->nlon + PTX - 1) / PTX), (unsigned)((a->nlat + PTY - 1) / PTY),
                      (unsigned)(nq * a->nlev));
        hipLaunchKernelGGL(rs_ppm_vjp_xtile, xg, dim3(PB), 0, st, *a, G);
    }
    hipLaunchKernelGGL(rs_ppm_vjp_mrow, blocks(a->nc), dim3(PB), 0, st, *a, F);
    launch_vjp_pbl(a, st);
    return hipGetLastError();
}
