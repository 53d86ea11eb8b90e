// transport.hip -- SSPRK43 flux-form transport stages, their adjoints, the
// device-side PI controller and small utility kernels (FP64, gfx950).
//
// Transport operator (the workloads' 3rd-order upwind scheme): for tracer q,
//   f = -(F_e - F_w) - (G_t - G_b),
//   F_{i+1/2} = cx_j * q_f,   q_f = (-q_mm + 5 q_m + 2 q_p) / 6 (upwind side m),
//   G_{k}     = wz_k * wl_i * q_f  at interior vertical faces, 0 at bottom/top,
// periodic in longitude; vertical stencil indices clamped to the column.
//
// Layout and tiling: species-major, longitude fastest.  A 256-lane workgroup
// (4 waves) owns a 64 (lon) x 16 (level) tile of one (tracer, latitude) slab;
// it stages the tile plus a 2-cell halo in every direction in LDS with
// coalesced loads (periodic wrap in longitude, clamp in the vertical), and each
// lane computes 4 cells from LDS.  v = 0, so latitude rows never interact.
//
// Stages (ssprk43_step_unrolled, rx_traced_integrator.jl:404-437):
//   KIND_STAGE : out = a + dt/2 * L a                    (stages 1-3)
//   KIND_LAST  : u4 = 2/3 u + 1/3 u3 on load;  out = u4 + dt/2 * L u4, plus the
//                scaled error of utilde = (1/3 u + 2/3 u3 - unew)/2 summed per
//                block (deterministic tree) for the global Hairer norm.
// Adjoint stages (transpose of the above, L^T as a gather over the faces each
// cell takes part in):
//   KIND_ADJ_A : out = a + dt/2 L^T a                     (lambda4)
//   KIND_ADJ_B : x = a/3 on load;  out = x + dt/2 L^T x    (lambda2)
//   KIND_ADJ_C : out = a + dt/2 L^T a                     (lambda1)
//   KIND_ADJ_D : out = 2/3 b + a + dt/2 L^T a             (lambda_in; b = lambda4)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "rs_launch.h"

#define TX 64
#define TK_DEFAULT 16
#define LXW (TX + 4)

#define TB 256

#define KIND_STAGE 0
#define KIND_LAST 1
#define KIND_ADJ_A 2
#define KIND_ADJ_B 3
#define KIND_ADJ_C 4
#define KIND_ADJ_D 5

namespace {

constexpr double C16 = 1.0 / 6.0;
constexpr double C13 = 1.0 / 3.0;
constexpr double C23 = 2.0 / 3.0;

__device__ __forceinline__ double face3(double qmm, double qm, double qp) {
    return (5.0 * qm - qmm + 2.0 * qp) * C16;
}

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// Sum over a workgroup of NT lanes (whole waves), deterministic: shuffle tree
// inside each wave, then wave order; the result is valid in lane 0.  `red`
// holds NT/64 doubles.
template <int NT>
__device__ __forceinline__ double block_sum_fixed(double v, double* red, int tid) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (tid == 0)
        for (int q = 0; q < NT / 64; ++q) t += red[q];
    return t;
}

__device__ __forceinline__ int wrapi(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// Forward tendency of cell (kk+2, ii+2) of the LDS tile (level k, lon i).
__device__ __forceinline__ double tend_fwd(const double (*T)[LXW], int kk, int ii, int k, int i,
                                           int nlev, double cxj, const double* __restrict__ wz,
                                           double wli) {
    const int r = kk + 2, c = ii + 2;
    double fe, fw;
    if (cxj >= 0.0) {
        fe = cxj * face3(T[r][c - 1], T[r][c], T[r][c + 1]);
        fw = cxj * face3(T[r][c - 2], T[r][c - 1], T[r][c]);
    } else {
        fe = cxj * face3(T[r][c + 2], T[r][c + 1], T[r][c]);
        fw = cxj * face3(T[r][c + 1], T[r][c], T[r][c - 1]);
    }
    // vertical face kf between levels kf-1 and kf; LDS row of level L is L - k + r
    auto G = [&](int kf) -> double {
        if (kf <= 0 || kf >= nlev) return 0.0;
        const double w = wz[kf] * wli;
        if (w >= 0.0) {
            const int mm = kf - 2 < 0 ? 0 : kf - 2;
            return w * face3(T[mm - k + r][c], T[kf - 1 - k + r][c], T[kf - k + r][c]);
        } else {
            const int mm = kf + 1 > nlev - 1 ? nlev - 1 : kf + 1;
            return w * face3(T[mm - k + r][c], T[kf - k + r][c], T[kf - 1 - k + r][c]);
        }
    };
    const double gt = G(k + 1), gb = G(k);
    return -(fe - fw) - (gt - gb);
}

// (L^T lam) of cell (level k, lon i): every face flux F leaves its lower /
// western cell and enters its upper / eastern one, so Fbar = lam_east - lam_west;
// the cell receives Fbar * rate * (its stencil weight) from each face whose
// 3-point stencil contains it.  Faces: zonal i-2..i+1 (face f between f and
// f+1), vertical k-1..k+2 (face kf between kf-1 and kf); needs nlon >= 4.
__device__ __forceinline__ double tend_adj(const double (*T)[LXW], int kk, int ii, int k, int i,
                                           int nlon, int nlev, double cxj,
                                           const double* __restrict__ wz,
                                           const double* __restrict__ wl) {
    const int r = kk + 2, c = ii + 2;
    double acc = 0.0;
    // zonal: offsets relative to i, valid for nlon >= 4 (all distinct mod nlon)
#pragma unroll
    for (int d = -2; d <= 1; ++d) {
        // face between (i+d) and (i+d+1)
        const double Fb = T[r][c + d + 1] - T[r][c + d];
        int omm, om, op;   // stencil offsets relative to i
        if (cxj >= 0.0) { omm = d - 1; om = d; op = d + 1; }
        else { omm = d + 2; om = d + 1; op = d; }
        double wgt = 0.0;
        if (omm == 0) wgt += -1.0;
        if (om == 0) wgt += 5.0;
        if (op == 0) wgt += 2.0;
        acc += (cxj * Fb * C16) * wgt;
    }
    const double wli = wl[i];
#pragma unroll
    for (int kf = k - 1; kf <= k + 2; ++kf) {
        if (kf <= 0 || kf >= nlev) continue;
        const double w = wz[kf] * wli;
        const double Gb = T[kf - k + r][c] - T[kf - 1 - k + r][c];
        int mm, m, p;
        if (w >= 0.0) { mm = kf - 2 < 0 ? 0 : kf - 2; m = kf - 1; p = kf; }
        else { mm = kf + 1 > nlev - 1 ? nlev - 1 : kf + 1; m = kf; p = kf - 1; }
        double wgt = 0.0;
        if (mm == k) wgt += -1.0;
        if (m == k) wgt += 5.0;
        if (p == k) wgt += 2.0;
        acc += (w * Gb * C16) * wgt;
    }
    (void)nlon;
    return acc;
}

}  // namespace

template <int KIND, int TKV>
__global__ void __launch_bounds__(TB) rs_transport_stage(RsTransArgs A) {
    __shared__ double T[TKV + 4][LXW];
    __shared__ double red[TB / 64];
    const RsCtrl* ctl = A.ctrl;
    if (ctl && ctl->done) return;
    const int cur = ctl ? ctl->cur : A.cur_static;
    const double dt = ctl ? ctl->dtc : A.dt_static;
    const double h2 = dt / 2;
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev;
    const int i0 = blockIdx.x * TX, k0 = blockIdx.y * TKV;
    const int j = blockIdx.z % nlat, s = blockIdx.z / nlat;
    const long long sb = (long long)s * A.nc;
    const long long plane = (long long)nlon * nlat;
    const double* __restrict__ pa = A.a_is_state ? (cur ? A.S1 : A.S0) : A.a;
    const double* __restrict__ pb = A.b;
    double* __restrict__ po = A.out_is_state ? (cur ? A.S0 : A.S1) : A.out;
    const int tid = threadIdx.x;
    // ---- stage the tile + halo
    for (int x = tid; x < (TKV + 4) * LXW; x += TB) {
        const int kk = x / LXW, ii = x - kk * LXW;
        const int gk = k0 - 2 + kk;
        double v = 0.0;
        if (gk >= 0 && gk < nlev) {
            const int gi = wrapi(i0 - 2 + ii, nlon);
            const long long g = sb + gk * plane + (long long)j * nlon + gi;
            if (KIND == KIND_LAST) {
                const double a = C23 * pa[g];
                const double b = C13 * pb[g];
                v = a + b;
            } else if (KIND == KIND_ADJ_B) {
                v = C13 * pa[g];
            } else {
                v = pa[g];
            }
        }
        T[kk][ii] = v;
    }
    __syncthreads();
    const int ii = tid & 63, ky = tid >> 6;
    const int i = i0 + ii;
    const double cxj = A.cx[j];
    double sq = 0.0, ng = 0.0;
#pragma unroll
    for (int rr = 0; rr < TKV / 4; ++rr) {
        const int kk = ky + 4 * rr;
        const int k = k0 + kk;
        if (i >= nlon || k >= nlev) continue;
        const long long g = sb + k * plane + (long long)j * nlon + i;
        const double xc = T[kk + 2][ii + 2];
        if (KIND == KIND_STAGE || KIND == KIND_LAST) {
            const double f = tend_fwd(T, kk, ii, k, i, nlev, cxj, A.wz, A.wl[i]);
            const double du = h2 * f;
            const double un = xc + du;
            if (KIND == KIND_LAST) {
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
            } else {
                po[g] = un;
            }
        } else {
            const double g2 = tend_adj(T, kk, ii, k, i, nlon, nlev, cxj, A.wz, A.wl);
            const double v = xc + h2 * g2;
            if (KIND == KIND_ADJ_D) po[g] = C23 * pb[g] + v;
            else po[g] = v;
        }
    }
    if (KIND == KIND_LAST) {
        const long long bid = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        const double t = block_sum_fixed<TB>(sq, red, tid);
        if (tid == 0) A.partial[bid] = t;
        if (A.negpart) {
            const double nt = block_sum_fixed<TB>(ng, red, tid);
            if (tid == 0) A.negpart[bid] = nt;
        }
    }
}

// ---- fused SSPRK43 attempt / adjoint: all four stages in one launch --------
// A workgroup of NTH lanes (1024, or 512 so that two share a CU and one's
// barrier waits overlap the other's work, on narrower segments) owns one
// (tracer, latitude) slab, or a longitude segment of it, with all levels, in
// LDS.  With seg_halo = 0 the segment is
// the whole periodic row and the zonal stencil wraps inside LDS; otherwise
// FHALO = 8 columns on each side (2 per stage x 4 stages) are recomputed and
// only the seg_in interior columns are written.
//
// LDS layout: column-major, pitch P = (nlev + 4) | 1 doubles per column, with
// two ghost levels below level 0 and above level nlev-1 holding copies of the
// boundary value (the vertical stencil's index clamp becomes a plain read;
// the odd pitch keeps the zonal reads of a wave on distinct banks).
// Lane t owns column t % seg_w and the seg_lev consecutive levels of chunk
// t / seg_w.  It walks its levels upwards with the column values in a sliding
// register window, reuses each vertical face flux for the two cells it
// separates, and keeps its new values in registers until every lane has read
// the old ones.  HBM sees one read of the input and one write of the output
// per tracer-cell (times seg_w / seg_in for the read when segmented).  Every
// face flux and every cell update is the same expression, in the same order,
// as in the stage kernels above, so both paths give the same bits.
#define FT 1024      // lanes per workgroup (the default; 512 is the other instantiation)
#define FL 8
#define FHALO 8

namespace {

// LDS column of local column i = il + d, |d| <= 2: wraps around the row when
// the segment is the whole periodic row, clamps at a halo edge otherwise (those
// columns hold values that never reach a kept interior cell).
__device__ __forceinline__ int fcol(int i, int w, int periodic) {
    if (periodic) {
        i = i < 0 ? i + w : i;
        return i >= w ? i - w : i;
    }
    return i < 0 ? 0 : (i >= w ? w - 1 : i);
}

__host__ __device__ constexpr int fpitch(int nlev) { return (nlev + 4) | 1; }

// Largest dynamic LDS a fused launch asks for (gfx950: 160 KB per workgroup).
#define FUSED_LDS_MAX (160 * 1024)

// One lane's view: pointers at level k0 of its own column and of the four
// zonal neighbours, and at wz[k0]; x[j] is level k0 + j.
struct FCol {
    const double* pc;
    const double* pm2;
    const double* pm1;
    const double* pp1;
    const double* pp2;
    const double* pwz;
    int k0, nlev;
    double wli, cxj;
    // vertical flux at face kf = k0 + jf from a = V(kf-2), b = V(kf-1),
    // c = V(kf), d = V(kf+1) (clamped): zero at the bottom and top faces
    __device__ __forceinline__ double G(int jf, double a, double b, double c, double d) const {
        const int kf = k0 + jf;
        if (kf <= 0 || kf >= nlev) return 0.0;
        const double wk = pwz[jf] * wli;
        if (wk >= 0.0) return wk * face3(a, b, c);
        return wk * face3(d, c, b);
    }
    // fe - fw at level k0 + j, q0 = the cell's own value
    __device__ __forceinline__ double zonal(int j, double q0) const {
        double fe, fw;
        if (cxj >= 0.0) {
            const double qm2 = pm2[j], qm1 = pm1[j], qp1 = pp1[j];
            fe = cxj * face3(qm1, q0, qp1);
            fw = cxj * face3(qm2, qm1, q0);
        } else {
            const double qm1 = pm1[j], qp1 = pp1[j], qp2 = pp2[j];
            fe = cxj * face3(qp2, qp1, q0);
            fw = cxj * face3(qp1, q0, qm1);
        }
        return fe - fw;
    }
};

// out(j, x + dt/2 L x) for the lane's levels k0 + j, j < nk
template <class Out>
__device__ __forceinline__ void fstage_fwd(const FCol& C, int nk, double h2, Out out) {
    double a = C.pc[-2], b = C.pc[-1], c = C.pc[0], d = C.pc[1];
    double gb = C.G(0, a, b, c, d);
#pragma unroll
    for (int j = 0; j < FL; ++j) {
        if (j >= nk) break;
        const double e = C.pc[j + 2];
        const double gt = C.G(j + 1, b, c, d, e);
        const double fz = C.zonal(j, c);
        const double f = -fz - (gt - gb);
        const double du = h2 * f;
        out(j, c + du);
        a = b; b = c; c = d; d = e; gb = gt;
    }
}

// out(j, x + dt/2 L^T x); L^T gathers, for cell k, the faces whose 3-point
// stencil contains it: zonal faces i-2..i+1, vertical faces k-1..k+2
template <class Out>
__device__ __forceinline__ void fstage_adj(const FCol& C, int nk, double h2, Out out) {
    const int nlev = C.nlev;
    double a = C.pc[-2], b = C.pc[-1], c = C.pc[0], d = C.pc[1];
    double w0 = C.pwz[-1] * C.wli, w1 = C.pwz[0] * C.wli, w2 = C.pwz[1] * C.wli;
#pragma unroll
    for (int j = 0; j < FL; ++j) {
        if (j >= nk) break;
        const int k = C.k0 + j;
        const double e = C.pc[j + 2];
        const double w3 = C.pwz[j + 2] * C.wli;
        double q[5];
        q[0] = C.pm2[j]; q[1] = C.pm1[j]; q[2] = c; q[3] = C.pp1[j]; q[4] = C.pp2[j];
        double acc = 0.0;
#pragma unroll
        for (int dd = -2; dd <= 1; ++dd) {
            const double Fb = q[dd + 3] - q[dd + 2];
            int omm, om, op;
            if (C.cxj >= 0.0) { omm = dd - 1; om = dd; op = dd + 1; }
            else { omm = dd + 2; om = dd + 1; op = dd; }
            double wgt = 0.0;
            if (omm == 0) wgt += -1.0;
            if (om == 0) wgt += 5.0;
            if (op == 0) wgt += 2.0;
            acc += (C.cxj * Fb * C16) * wgt;
        }
        // faces kf = k-1 .. k+2 and the values below / above each
        const double lo[4] = {a, b, c, d}, hi[4] = {b, c, d, e}, wvs[4] = {w0, w1, w2, w3};
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
            const int kf = k - 1 + q4;
            if (kf <= 0 || kf >= nlev) continue;
            const double wk = wvs[q4];
            const double Gb = hi[q4] - lo[q4];
            int mm, m, p;
            if (wk >= 0.0) { mm = kf - 2 < 0 ? 0 : kf - 2; m = kf - 1; p = kf; }
            else { mm = kf + 1 > nlev - 1 ? nlev - 1 : kf + 1; m = kf; p = kf - 1; }
            double wgt = 0.0;
            if (mm == k) wgt += -1.0;
            if (m == k) wgt += 5.0;
            if (p == k) wgt += 2.0;
            acc += (wk * Gb * C16) * wgt;
        }
        out(j, c + h2 * acc);
        a = b; b = c; c = d; d = e;
        w0 = w1; w1 = w2; w2 = w3;
    }
}

}  // namespace

// PD: positive-domain solves also count the negative entries of the trial
// state per workgroup (compile-time, so other solves do not carry it).
template <int ADJ, int PD, int NTH>
__global__ void __launch_bounds__(NTH) rs_transport_fused(RsTransArgs P) {
    extern __shared__ double lds[];
    __shared__ double red[NTH / 64];
    const RsCtrl* ctl = P.ctrl;
    if (ctl && ctl->done) return;
    const int cur = ctl ? ctl->cur : P.cur_static;
    const double dt = ctl ? ctl->dtc : P.dt_static;
    const double h2 = dt / 2;
    const int nlon = P.nlon, nlat = P.nlat, nlev = P.nlev, w = P.seg_w, L = P.seg_lev;
    const int pitch = fpitch(nlev);
    const int periodic = P.seg_halo == 0;
    const int seg = blockIdx.x, j = blockIdx.y % nlat, s = blockIdx.y / nlat;
    const int i_start = periodic ? 0 : seg * P.seg_in - FHALO;
    const long long base = (long long)s * P.nc + (long long)j * nlon;
    const long long plane = (long long)nlon * nlat;
    // pa and po may be the same buffer (in-place adjoint on whole rows): every
    // read of pa happens before the first barrier, every write after the last
    const double* pa = P.a_is_state ? (cur ? P.S1 : P.S0) : P.a;
    double* po = P.out_is_state ? (cur ? P.S0 : P.S1) : P.out;
    double* A = lds;                  // w columns of pitch doubles
    double* Wl = lds + w * pitch;     // w
    double* Wz = Wl + w;              // wz[-1 .. nlev+2], zero outside 0..nlev
    const int tid = threadIdx.x;
    // global column of local column x (a segment wraps at most once)
    auto gcol = [&](int x) {
        int gi = i_start + x;
        gi = gi < 0 ? gi + nlon : gi;
        return gi >= nlon ? gi - nlon : gi;
    };
    for (int x = tid; x < w; x += NTH) Wl[x] = P.wl[gcol(x)];
    for (int x = tid; x < nlev + 4; x += NTH) Wz[x] = (x >= 1 && x <= nlev + 1) ? P.wz[x - 1] : 0.0;
    const int il = tid % w, kc = tid / w;
    const int nk = kc < P.seg_chunks ? min(L, nlev - kc * L) : 0;
    const int k0 = nk > 0 ? kc * L : 0;
    const long long g0 = base + gcol(il);
    double* col = A + il * pitch + 2 + k0;   // col[j] = level k0 + j of this column
    // write x[j] for the lane's levels, and the ghost levels at either end
    auto put = [&](auto&& val) {
#pragma unroll
        for (int jj = 0; jj < FL; ++jj) {
            if (jj >= nk) break;
            col[jj] = val(jj);
        }
        if (nk > 0 && k0 == 0) { const double x0 = val(0); col[-1] = x0; col[-2] = x0; }
        if (nk > 0 && k0 + nk == nlev) {
            const double xt = val(nk - 1);
            col[nk] = xt;
            col[nk + 1] = xt;
        }
    };
    double v[FL], y[FL];
#pragma unroll
    for (int jj = 0; jj < FL; ++jj) {
        v[jj] = 0.0;
        y[jj] = 0.0;
        if (jj < nk) {
            const long long ix = g0 + (k0 + jj) * plane;
            v[jj] = pa[ix];
            // the clamp mask of the step output, as the separate mask pass
            // computes it: lambda * (u_out > 0 ? 1 : 0)
            if (ADJ && P.lmask) v[jj] = v[jj] * ((P.lmask[ix] > 0.0) ? 1.0 : 0.0);
        }
    }
    put([&](int jj) { return v[jj]; });
    __syncthreads();
    FCol C;
    C.pc = col;
    C.pm2 = A + fcol(il - 2, w, periodic) * pitch + 2 + k0;
    C.pm1 = A + fcol(il - 1, w, periodic) * pitch + 2 + k0;
    C.pp1 = A + fcol(il + 1, w, periodic) * pitch + 2 + k0;
    C.pp2 = A + fcol(il + 2, w, periodic) * pitch + 2 + k0;
    C.pwz = Wz + 1 + k0;
    C.k0 = k0;
    C.nlev = nlev;
    C.wli = Wl[il];
    C.cxj = P.cx[j];
    // stages 1-3 forward (x <- x + dt/2 L x), or the first three adjoint stages
    auto to_y = [&](int jj, double x) { y[jj] = x; };
    for (int st = 0; st < 3; ++st) {
        if (ADJ) fstage_adj(C, nk, h2, to_y);
        else fstage_fwd(C, nk, h2, to_y);
        __syncthreads();
        if (ADJ && st == 0) {
            // v <- lambda4; the next stage starts from lambda4 / 3
#pragma unroll
            for (int jj = 0; jj < FL; ++jj) v[jj] = y[jj];
            put([&](int jj) { return C13 * y[jj]; });
        } else {
            put([&](int jj) { return y[jj]; });
        }
        __syncthreads();
    }
    if (!ADJ) {
        // u4 = 2/3 u + 1/3 u3 (v = u, y = u3)
        put([&](int jj) {
            const double a = C23 * v[jj];
            const double b = C13 * y[jj];
            return a + b;
        });
        __syncthreads();
    }
    const bool keep = periodic || (il >= FHALO && il < FHALO + P.seg_in &&
                                   seg * P.seg_in + il - FHALO < nlon);
    double sq = 0.0, ng = 0.0;
    if (ADJ) {
        // out = 2/3 lambda4 + (x + dt/2 L^T x)
        unsigned nbad = 0;
        fstage_adj(C, nk, h2, [&](int jj, double z) {
            if (!keep) return;
            const double o = C23 * v[jj] + z;
            po[g0 + (k0 + jj) * plane] = o;
            nbad += !__builtin_isfinite(o);
        });
        if (P.nfcount) {
            // one atomic per wave that wrote a non-finite value
            for (int off = 32; off > 0; off >>= 1) nbad += __shfl_down(nbad, off, 64);
            if ((tid & 63) == 0 && nbad) atomicAdd(P.nfcount, (unsigned long long)nbad);
        }
    } else {
        fstage_fwd(C, nk, h2, [&](int jj, double un) {
            if (!keep) return;
            const double u = v[jj], u3 = y[jj];
            const double a = C13 * u;
            const double b = C23 * u3;
            double ut = a + b;
            const double ud = ut - un;
            ut = 0.5 * ud;
            const double am = dmax(fabs(u), fabs(un));
            double den = am * P.rtol;
            den = den + P.atol;
            const double at = ut / den;
            sq += at * at;
            if (PD && un < 0.0) ng += 1.0;
            po[g0 + (k0 + jj) * plane] = P.clamp ? dmax(un, 0.0) : un;
        });
    }
    if (!ADJ) {
        const long long bid = (long long)blockIdx.y * gridDim.x + blockIdx.x;
        const double t = block_sum_fixed<NTH>(sq, red, tid);
        if (tid == 0) P.partial[bid] = t;
        if (PD) {
            const double nt = block_sum_fixed<NTH>(ng, red, tid);
            if (tid == 0) P.negpart[bid] = nt;
        }
    }
}

// ---- device-side PI controller (rx_traced_integrator.jl:480-505) ----------
// One 1024-lane block: each lane folds its strided share of the block partials
// into four independent accumulators (so the loads overlap instead of
// forming one dependent chain), then a fixed tree sums the lanes.  The order
// depends only on npart, so EEst is reproducible run to run.
#define CTRL_THREADS 1024
// Fixed-order sum of n values by the controller's CTRL_THREADS lanes (four
// independent accumulators per lane, then a tree); valid in lane 0.
__device__ __forceinline__ double ctrl_sum(const double* v, int n, double* sh) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int x = threadIdx.x;
    for (; x + 3 * CTRL_THREADS < n; x += 4 * CTRL_THREADS) {
        s0 += v[x];
        s1 += v[x + CTRL_THREADS];
        s2 += v[x + 2 * CTRL_THREADS];
        s3 += v[x + 3 * CTRL_THREADS];
    }
    for (; x < n; x += CTRL_THREADS) s0 += v[x];
    __syncthreads();
    sh[threadIdx.x] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int w = CTRL_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

// negpart (may be 0): per-block counts of negative entries of the trial state;
// any makes the attempt a rejection with dt <- dtc / 2 (the failfactor rule of
// an out-of-domain trial step, positive-domain mode).
__global__ void __launch_bounds__(CTRL_THREADS) rs_ctrl_update(RsCtrl* c, const double* partial,
                                                               int npart, double* seqlog,
                                                               const double* negpart) {
    __shared__ double sh[CTRL_THREADS];
    if (c->done) return;
    const double sse = ctrl_sum(partial, npart, sh);
    const double nneg = negpart ? ctrl_sum(negpart, npart, sh) : 0.0;
    if (threadIdx.x != 0) return;
    const double raw = sqrt(sse / c->nnorm);
    const double EEst = (raw != raw) ? 1.0e10 : raw;
    const double q11 = pow(dmax(EEst, 1.0e-35), c->beta1);
    const double qden = pow(c->qold, c->beta2);
    double q = q11 / qden;
    const double qg = q / c->gamma;
    q = dmax(c->invqmax, (c->invqmin < qg) ? c->invqmin : qg);
    const int isout = nneg > 0.0;
    const int accept = EEst <= 1.0 && !isout;
    const int insteady = (c->qsmin <= q) && (q <= c->qsmax);
    const double qa = insteady ? 1.0 : q;
    const double dtc = c->dtc;
    c->eest = EEst;
    c->accept = accept;
    if (accept) {
        if (seqlog && c->nlog < c->logcap) {
            seqlog[2 * c->nlog] = c->t;
            seqlog[2 * c->nlog + 1] = dtc;
            c->nlog += 1;
        }
        c->t = c->t + dtc;
        c->dt = dtc / qa;
        c->qold = dmax(EEst, c->qoldinit);
        c->nacc += 1;
        c->cur ^= 1;
    } else if (isout) {
        c->dt = dtc / 2.0;
        c->nrej += 1;
    } else {
        const double q11g = q11 / c->gamma;
        const double qr = (c->invqmin < q11g) ? c->invqmin : q11g;
        c->dt = dtc / qr;
        c->nrej += 1;
    }
    c->iters += 1;
    if (!(c->t < c->tlim) || c->iters >= c->maxiters) {
        c->done = 1;
    } else {
        const double rem = c->t1 - c->t;
        c->dtc = (c->dt < rem) ? c->dt : rem;
    }
}

// ---- utilities ---------------------------------------------------------------
__global__ void rs_fill(double* x, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

__global__ void rs_fill_u64(unsigned long long* x, int n, unsigned long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// lam .*= 1[uout > 0] (a multiplication: a non-finite entry stays non-finite,
// adjoint_gradient.jl:736)
__global__ void rs_mask(double* lam, const double* uout, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        lam[i] = lam[i] * ((uout[i] > 0.0) ? 1.0 : 0.0);
}

// *count += number of non-finite entries of x (a lane with none issues no
// atomic)
__global__ void __launch_bounds__(256) rs_nonfinite(const double* x, long long n,
                                                    unsigned long long* count) {
    unsigned k = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        k += !__builtin_isfinite(x[i]);
    if (k) atomicAdd(count, (unsigned long long)k);
}

// x = max(x, 0)
__global__ void rs_clamp(double* x, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = dmax(x[i], 0.0);
}

// per-block max of |a-b| / max(|b|, 1e-30)
__global__ void __launch_bounds__(256) rs_relmax(const double* a, const double* b, long long n,
                                                 double* part) {
    __shared__ double sh[256];
    double m = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double d = fabs(a[i] - b[i]) / dmax(fabs(b[i]), 1e-30);
        if (!(d <= DBL_MAX)) d = HUGE_VAL;   // a NaN deviation must not vanish in the max
        m = dmax(m, d);
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = dmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// per-block sum of x (fixed order)
__global__ void __launch_bounds__(256) rs_sum(const double* x, long long n, double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        s += x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// J = sum over level-0 cells of q^2 (per block), q = species block of the state
__global__ void __launch_bounds__(256) rs_sumsq(const double* q, long long n, double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        s += q[i] * q[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// per-block sum of w * u (fixed order): the linear objective <w, u>
__global__ void __launch_bounds__(256) rs_dot(const double* w, const double* u, long long n,
                                              double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        s += w[i] * u[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w2 = 128; w2 > 0; w2 >>= 1) {
        if ((int)threadIdx.x < w2) sh[threadIdx.x] += sh[threadIdx.x + w2];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// lam = 2 q on level-0 cells of species s, 0 elsewhere (dJ/du of the loss)
__global__ void rs_loss_seed(double* lam, const double* u, long long ntot, long long sbase,
                             long long ncol) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ntot;
         i += (long long)gridDim.x * blockDim.x) {
        const long long o = i - sbase;
        lam[i] = (o >= 0 && o < ncol) ? 2.0 * u[i] : 0.0;
    }
}

// ---- host-visible launch wrappers ------------------------------------------
// Tile heights the launcher (and the autotuner) can choose from.
template <int TKV>
static hipError_t launch_tk(int kind, const RsTransArgs* a, hipStream_t st) {
    dim3 grid((a->nlon + TX - 1) / TX, (a->nlev + TKV - 1) / TKV, (unsigned)(a->nlat * a->nt));
    dim3 block(TB);
    switch (kind) {
        case KIND_STAGE: hipLaunchKernelGGL((rs_transport_stage<KIND_STAGE, TKV>), grid, block, 0, st, *a); break;
        case KIND_LAST: hipLaunchKernelGGL((rs_transport_stage<KIND_LAST, TKV>), grid, block, 0, st, *a); break;
        case KIND_ADJ_A: hipLaunchKernelGGL((rs_transport_stage<KIND_ADJ_A, TKV>), grid, block, 0, st, *a); break;
        case KIND_ADJ_B: hipLaunchKernelGGL((rs_transport_stage<KIND_ADJ_B, TKV>), grid, block, 0, st, *a); break;
        case KIND_ADJ_C: hipLaunchKernelGGL((rs_transport_stage<KIND_ADJ_C, TKV>), grid, block, 0, st, *a); break;
        case KIND_ADJ_D: hipLaunchKernelGGL((rs_transport_stage<KIND_ADJ_D, TKV>), grid, block, 0, st, *a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" {

hipError_t rs_launch_transport(int kind, int tk, const RsTransArgs* a, hipStream_t st) {
    switch (tk) {
        case 4: return launch_tk<4>(kind, a, st);
        case 8: return launch_tk<8>(kind, a, st);
        case 16: return launch_tk<16>(kind, a, st);
        case 32: return launch_tk<32>(kind, a, st);
        default: return hipErrorInvalidValue;
    }
}

long long rs_transport_nblocks(int nlon, int nlat, int nlev, int nt, int tk) {
    return (long long)((nlon + TX - 1) / TX) * ((nlev + tk - 1) / tk) * (long long)nlat * nt;
}

hipError_t rs_launch_ctrl(RsCtrl* c, const double* partial, int npart, double* seqlog,
                          const double* negpart, hipStream_t st) {
    hipLaunchKernelGGL(rs_ctrl_update, dim3(1), dim3(CTRL_THREADS), 0, st, c, partial, npart, seqlog,
                       negpart);
    return hipGetLastError();
}

static unsigned util_grid(long long n) {
    long long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t rs_launch_fill(double* x, long long n, double v, hipStream_t st) {
    hipLaunchKernelGGL(rs_fill, dim3(util_grid(n)), dim3(256), 0, st, x, n, v);
    return hipGetLastError();
}
hipError_t rs_launch_fill_u64(unsigned long long* x, int n, unsigned long long v, hipStream_t st) {
    hipLaunchKernelGGL(rs_fill_u64, dim3((n + 255) / 256), dim3(256), 0, st, x, n, v);
    return hipGetLastError();
}
hipError_t rs_launch_mask(double* lam, const double* uout, long long n, hipStream_t st) {
    hipLaunchKernelGGL(rs_mask, dim3(util_grid(n)), dim3(256), 0, st, lam, uout, n);
    return hipGetLastError();
}
hipError_t rs_launch_clamp(double* x, long long n, hipStream_t st) {
    hipLaunchKernelGGL(rs_clamp, dim3(util_grid(n)), dim3(256), 0, st, x, n);
    return hipGetLastError();
}
int rs_reduce_blocks(long long n) { return (int)util_grid(n) > 1024 ? 1024 : (int)util_grid(n); }
hipError_t rs_launch_nonfinite(const double* x, long long n, unsigned long long* count,
                               hipStream_t st) {
    hipLaunchKernelGGL(rs_nonfinite, dim3(rs_reduce_blocks(n)), dim3(256), 0, st, x, n, count);
    return hipGetLastError();
}
hipError_t rs_launch_dot(const double* w, const double* u, long long n, double* part,
                         hipStream_t st) {
    hipLaunchKernelGGL(rs_dot, dim3(rs_reduce_blocks(n)), dim3(256), 0, st, w, u, n, part);
    return hipGetLastError();
}
hipError_t rs_launch_relmax(const double* a, const double* b, long long n, double* part,
                            hipStream_t st) {
    hipLaunchKernelGGL(rs_relmax, dim3(rs_reduce_blocks(n)), dim3(256), 0, st, a, b, n, part);
    return hipGetLastError();
}
hipError_t rs_launch_sum(const double* x, long long n, double* part, hipStream_t st) {
    hipLaunchKernelGGL(rs_sum, dim3(rs_reduce_blocks(n)), dim3(256), 0, st, x, n, part);
    return hipGetLastError();
}
hipError_t rs_launch_sumsq(const double* q, long long n, double* part, hipStream_t st) {
    hipLaunchKernelGGL(rs_sumsq, dim3(rs_reduce_blocks(n)), dim3(256), 0, st, q, n, part);
    return hipGetLastError();
}
hipError_t rs_launch_loss_seed(double* lam, const double* u, long long ntot, long long sbase,
                               long long ncol, hipStream_t st) {
    hipLaunchKernelGGL(rs_loss_seed, dim3(util_grid(ntot)), dim3(256), 0, st, lam, u, ntot, sbase,
                       ncol);
    return hipGetLastError();
}

}  // extern "C"

// ---- fused launch wrappers ---------------------------------------------------
extern "C" {

// Segment plan for the fused kernels on an nlon x nlev slab: plan[0..5] =
// seg_w, seg_in, seg_halo, nseg, seg_lev, seg_chunks.  Returns 0 if the fused
// kernels cannot take this grid (nlon < 4, or too many levels for a useful
// segment: at most FL levels per lane and nth (1024 or 512) lanes per
// workgroup, and the tile within lds_bytes of LDS per workgroup).
int rs_transport_fused_plan(int nlon, int nlev, long long lds_bytes, int nth, int* plan) {
    if (nth != 1024 && nth != 512) return 0;
    const long long lds_doubles = (std::min<long long>(lds_bytes, FUSED_LDS_MAX) - 256) / 8;
    auto fits = [&](long long w, int* lev, int* chunks) {
        if (w < 4 || w > nth) return false;
        if (w * fpitch(nlev) + w + nlev + 4 > lds_doubles) return false;
        const int per_col = (int)(nth / w);
        const int L = (nlev + per_col - 1) / per_col;
        if (L > FL) return false;
        *lev = L;
        *chunks = (nlev + L - 1) / L;
        return true;
    };
    if (nlon < 4 || nlev < 1) return 0;
    int L = 0, ch = 0;
    if (fits(nlon, &L, &ch)) {
        plan[0] = nlon; plan[1] = nlon; plan[2] = 0; plan[3] = 1; plan[4] = L; plan[5] = ch;
        return 1;
    }
    long long wmax = std::min<long long>(nth, nlon + 2 * FHALO);
    while (wmax >= 2 * FHALO + 16 && !fits(wmax, &L, &ch)) --wmax;
    if (wmax < 2 * FHALO + 16) return 0;
    const int nseg = (int)((nlon + (wmax - 2 * FHALO) - 1) / (wmax - 2 * FHALO));
    const int seg_in = (nlon + nseg - 1) / nseg;
    // every local column maps into [-nlon, 2 nlon) (one wrap at most)
    if (seg_in + FHALO > nlon || FHALO > nlon) return 0;
    if (!fits(seg_in + 2 * FHALO, &L, &ch)) return 0;
    plan[0] = seg_in + 2 * FHALO; plan[1] = seg_in; plan[2] = FHALO; plan[3] = nseg;
    plan[4] = L; plan[5] = ch;
    return 1;
}

long long rs_transport_fused_nblocks(int nlat, int nt, int nseg) {
    return (long long)nseg * nlat * nt;
}

hipError_t rs_launch_transport_fused(int adj, const RsTransArgs* a, hipStream_t st) {
    const int nth = a->fnth;
    if ((nth != 1024 && nth != 512) || a->seg_w < 4 || a->seg_w > nth || a->nseg < 1 ||
        a->seg_lev < 1 || a->seg_lev > FL || (long long)a->seg_w * a->seg_chunks > nth ||
        (long long)a->seg_lev * a->seg_chunks < a->nlev)
        return hipErrorInvalidValue;
    const size_t lds = ((size_t)a->seg_w * fpitch(a->nlev) + a->seg_w + a->nlev + 4) * sizeof(double);
    if (lds > FUSED_LDS_MAX - 256) return hipErrorInvalidValue;
    dim3 grid((unsigned)a->nseg, (unsigned)(a->nlat * a->nt));
    if (nth == 512) {
        if (adj) hipLaunchKernelGGL((rs_transport_fused<1, 0, 512>), grid, dim3(512), lds, st, *a);
        else if (a->negpart) hipLaunchKernelGGL((rs_transport_fused<0, 1, 512>), grid, dim3(512), lds, st, *a);
        else hipLaunchKernelGGL((rs_transport_fused<0, 0, 512>), grid, dim3(512), lds, st, *a);
    } else {
        if (adj) hipLaunchKernelGGL((rs_transport_fused<1, 0, FT>), grid, dim3(FT), lds, st, *a);
        else if (a->negpart) hipLaunchKernelGGL((rs_transport_fused<0, 1, FT>), grid, dim3(FT), lds, st, *a);
        else hipLaunchKernelGGL((rs_transport_fused<0, 0, FT>), grid, dim3(FT), lds, st, *a);
    }
    return hipGetLastError();
}

}  // extern "C"
