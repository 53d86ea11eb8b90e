/* chem_kernels.hip -- per-cell Rosenbrock23 kernels, compiled at run time
 * together with the mechanism code emitted by mechgen (rs_rhs, rs_jac_W,
 * rs_lu, rs_solve, rs_solveT, rs_jtv, rs_hess and the RS_NS / RS_NNZ sizes).
 *
 * Execution model: one lane per grid cell.  The species-major state makes the
 * per-species loads of consecutive lanes contiguous (coalesced); all of a
 * cell's working set -- u, f0, f1, k1..k3, the W = I - h*gamma*J factors --
 * lives in registers, because every index in the emitted code is a literal.
 */

/* One ROS23 attempt on one cell (ros23_step, rx_traced_integrator.jl:279-349,
 * mass matrix I, dT = 0).  Returns the cell's sum of squared scaled errors.
 * E (may be 0) points at this cell's emission of species 0, stride es.
 * The stage combinations the reference forms at the end -- c = 2(k1 - f0),
 * a = c32 (k2 - f1), k1 - 2 k2 -- are formed as soon as their inputs exist,
 * with the same operations, so that f0, f1 and k2 die early: six state
 * vectors plus the factors are live at the peak instead of ten. */
RS_FN double rs_ros23_attempt(const double* __restrict__ K, double jf,
                              const double* __restrict__ E, long long es,
                              const double* __restrict__ u, double* __restrict__ unew,
                              double dt, double atol, double rtol) {
    double W[RS_NNZ];
    double f[RS_NS], k1[RS_NS], k2[RS_NS], cc[RS_NS], aa[RS_NS];
    rs_rhs(K, jf, E, es, u, f);                                    /* f0, :282 */
    const double hg = dt * RS_ROS23_D;                             /* :291 */
    rs_jac_W(K, jf, u, hg, W);                                     /* :287-296 */
    rs_lu(W);
    rs_solve(W, f, k1);                                      /* :298 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { double c = k1[i] - f[i]; cc[i] = 2.0 * c; }   /* :323-324 */
    const double dto2 = dt / 2;
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double du1 = dto2 * k1[i]; aa[i] = u[i] + du1; }
    rs_rhs(K, jf, E, es, aa, f);                                   /* f1, :303 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) aa[i] = f[i] - k1[i];          /* :306-309 */
    rs_solve(W, aa, k2);                                     /* :310 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) k2[i] = k2[i] + k1[i];         /* :311-313 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { double a = k2[i] - f[i]; aa[i] = RS_ROS23_C32 * a; }  /* :321-322 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double du2 = dt * k2[i]; unew[i] = u[i] + du2; }
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double w2 = 2.0 * k2[i]; k1[i] = k1[i] - w2; }  /* :333-334 */
    rs_rhs(K, jf, E, es, unew, f);                                 /* f2, :318 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double d3 = f[i] - aa[i]; f[i] = d3 - cc[i]; }  /* :325-326 */
    rs_solve(W, f, k2);                                      /* k3, :328 */
    const double dto6 = dt / 6;
    double sse = 0.0;
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {                              /* :335-346 */
        double ut = k1[i] + k2[i];
        ut = dto6 * ut;
        const double am = rs_max(RS_FABS(u[i]), RS_FABS(unew[i]));
        double den = am * rtol;
        den = den + atol;
        const double at = ut / den;
        sse += at * at;
    }
    return sse;
}

/* Hand-derived adjoint of the part of one ROS23 step that reaches unew
 * (f2, k3 and EEst are dead, rx_traced_integrator.jl:540-547):
 *   f0 = f(u); k1 = W\f0; u1 = u + h/2 k1; f1 = f(u1); k2 = W\(f1-k1) + k1;
 *   unew = u + h k2,   W = I - h*gamma*J(u).
 * Given lam = dL/dunew, returns dL/du in lam and adds dL/dE to G[i*es] (G
 * may be 0; E enters f additively on the cells that have it).  Reverse:
 *   z2 = W^-T (h lam)             = f1bar;  k1bar = h lam - z2
 *   u1bar = J(u1)^T z2;  ubar = lam + u1bar;  k1bar += h/2 u1bar
 *   z1 = W^-T k1bar               = f0bar;  ubar += J(u)^T z1
 *   Jbar = h gamma (z2 (k2-k1)^T + z1 k1^T)  ->  ubar += d/du [z^T J(u) v]
 *   Ebar = z1 + z2.
 * The stages are recomputed from u, not stored. */
RS_FN void rs_ros23_vjp(const double* __restrict__ K, double jf, const double* __restrict__ E,
                        long long es, const double* __restrict__ u, double dt,
                        double* __restrict__ lam, double* __restrict__ G) {
    double W[RS_NNZ];
    double f[RS_NS], k1[RS_NS], k2[RS_NS], tmp[RS_NS], z[RS_NS], ub[RS_NS];
    rs_rhs(K, jf, E, es, u, f);
    const double hg = dt * RS_ROS23_D;
    rs_jac_W(K, jf, u, hg, W);
    rs_lu(W);
    rs_solve(W, f, k1);
    const double dto2 = dt / 2;
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double du1 = dto2 * k1[i]; tmp[i] = u[i] + du1; }
    rs_rhs(K, jf, E, es, tmp, f);                                  /* f1 at u1 = tmp */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) f[i] = f[i] - k1[i];
    rs_solve(W, f, k2);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) k2[i] = k2[i] + k1[i];
    /* reverse */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { f[i] = dt * lam[i]; ub[i] = lam[i]; }     /* k2bar */
    rs_solveT(W, f, z);                                      /* z2 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) f[i] = f[i] - z[i];            /* k1bar */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) k2[i] = k2[i] - k1[i];         /* v = k2 - k1 */
    rs_hess(K, jf, z, k2, hg, ub);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { k2[i] = z[i]; lam[i] = 0.0; }   /* keep z2 for Ebar */
    rs_jtv(K, jf, tmp, z, lam);                                    /* u1bar = J(u1)^T z2 */
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { ub[i] = ub[i] + lam[i]; f[i] = f[i] + dto2 * lam[i]; }
    rs_solveT(W, f, z);                                      /* z1 */
    rs_jtv(K, jf, u, z, ub);
    rs_hess(K, jf, z, k1, hg, ub);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) lam[i] = ub[i];
    if (G) {
#pragma unroll
        for (int i = 0; i < RS_NS; ++i) G[i * es] += z[i] + k2[i];
    }
}

/* PI controller update of one attempt (rx_traced_integrator.jl:484-504).
 * Returns 1 on accept.  t, dt, qold, nacc, nrej updated in place. */
RS_FN int rs_pi_update(double beta1, double beta2, double invqmin, double invqmax,
                       double gamma, double qsmin, double qsmax, double qoldinit,
                       double dtc, double eest_raw, double* t, double* dt, double* qold,
                       long long* nacc, long long* nrej) {
    const double EEst = (eest_raw != eest_raw) ? 1.0e10 : eest_raw;
    const double q11 = RS_POW(rs_max(EEst, 1.0e-35), beta1);
    const double qden = RS_POW(*qold, beta2);
    double q = q11 / qden;
    q = rs_max(invqmax, (invqmin < q / gamma) ? invqmin : q / gamma);
    const int accept = EEst <= 1.0;
    const int insteady = (qsmin <= q) && (q <= qsmax);
    const double qa = insteady ? 1.0 : q;
    if (accept) {
        *t = *t + dtc;
        *dt = dtc / qa;
        *qold = rs_max(EEst, qoldinit);
        *nacc += 1;
    } else {
        const double q11g = q11 / gamma;
        const double qr = (invqmin < q11g) ? invqmin : q11g;
        *dt = dtc / qr;
        *nrej += 1;
    }
    return accept;
}

#ifndef RS_HOST_BUILD

#define RS_CHEM_BLOCK 256   /* largest block the launcher uses (64, 128 or 256) */

__device__ inline double rs_block_sum(double v, double* sh) {
    /* deterministic: fixed shuffle tree inside the wave, then wave order */
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    return s;
}

extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK)
rs_chem_attempt(RsChemArgs a) {
    __shared__ double sh[RS_CHEM_BLOCK / 64];
    const RsCtrl* c = a.ctrl;
    if (c && c->done) return;
    const int cur = c ? c->cur : a.cur_static;
    const double dt = c ? c->dtc : a.dt_static;
    const double* __restrict__ src = cur ? a.S1 : a.S0;
    double* __restrict__ dst = cur ? a.S0 : a.S1;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sse = 0.0;
    if (cell < a.nc) {
        double u[RS_NS], un[RS_NS];
#pragma unroll
        for (int s = 0; s < RS_NS; ++s) u[s] = src[(long long)s * a.nc + cell];
        const double jf = a.jfac ? a.jfac[cell] : 1.0;
        const double* Ec = (a.E && cell < a.ncol) ? a.E + cell : nullptr;
        sse = rs_ros23_attempt(a.K, jf, Ec, a.ncol, u, un, dt, a.atol, a.rtol);
#pragma unroll
        for (int s = 0; s < RS_NS; ++s)
            dst[(long long)s * a.nc + cell] = a.clamp ? rs_max(un[s], 0.0) : un[s];
    }
    const double tot = rs_block_sum(sse, sh);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = tot;
}

/* Persistent per-cell adaptive solve with wave-aggregated work stealing: a
 * lane that finishes its cell takes the next index from the queue head; all
 * lanes run the same attempt code, each on its own cell and its own dt. */
extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK)
rs_chem_percell(RsPercellArgs a) {
    const int lane = threadIdx.x & 63;
    long long cell = -1;     /* -1: needs work, -2: queue drained */
    double u[RS_NS];
    const double* Ec = nullptr;
    double t = 0.0, dt = 0.0, qold = 0.0, jf = 1.0;
    long long iters = 0, nacc = 0, nrej = 0;
    long long s_att = 0, s_acc = 0, s_rej = 0, s_max = 0, s_min = -1, s_fail = 0;
    const double tlim = a.t1 - 1.0e-9;
    for (;;) {
        const bool need = (cell == -1);
        const unsigned long long m = __ballot(need);
        if (m) {
            const int leader = __builtin_ctzll(m);
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.head, (unsigned long long)__popcll(m));
            base = __shfl(base, leader, 64);
            if (need) {
                const unsigned long long rank = __popcll(m & ((1ull << lane) - 1ull));
                const unsigned long long idx = base + rank;
                if (idx < (unsigned long long)a.nc) {
                    cell = (long long)idx;
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) u[s] = a.u[(long long)s * a.nc + cell];
                    jf = a.jfac ? a.jfac[cell] : 1.0;
                    Ec = (a.E && cell < a.ncol) ? a.E + cell : nullptr;
                    t = a.t0;
                    dt = a.dtcell ? a.dtcell[cell] : a.dt0;
                    qold = a.qoldinit;
                    iters = nacc = nrej = 0;
                } else {
                    cell = -2;
                }
            }
        }
        if (!__any(cell >= 0)) break;
        if (cell >= 0) {
            if ((t < tlim) && (iters < a.maxiters)) {
                const double dtc = (dt < a.t1 - t) ? dt : a.t1 - t;
                double un[RS_NS];
                const double sse = rs_ros23_attempt(a.K, jf, Ec, a.ncol, u, un, dtc, a.atol, a.rtol);
                const double eest = RS_SQRT(sse / (double)RS_NS);
                if (rs_pi_update(a.beta1, a.beta2, a.invqmin, a.invqmax, a.gamma, a.qsmin,
                                 a.qsmax, a.qoldinit, dtc, eest, &t, &dt, &qold, &nacc, &nrej)) {
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) u[s] = a.clamp ? rs_max(un[s], 0.0) : un[s];
                }
                iters += 1;
            }
            if (!((t < tlim) && (iters < a.maxiters))) {
#pragma unroll
                for (int s = 0; s < RS_NS; ++s) a.u[(long long)s * a.nc + cell] = u[s];
                if (a.dtcell) a.dtcell[cell] = dt;
                s_att += iters; s_acc += nacc; s_rej += nrej;
                s_max = iters > s_max ? iters : s_max;
                s_min = (s_min < 0 || iters < s_min) ? iters : s_min;
                if (t < tlim) s_fail += 1;
                cell = -1;
            }
        }
    }
    /* wave reduction of the statistics, one atomic per wave */
    for (int off = 32; off > 0; off >>= 1) {
        s_att += __shfl_down(s_att, off, 64);
        s_acc += __shfl_down(s_acc, off, 64);
        s_rej += __shfl_down(s_rej, off, 64);
        s_fail += __shfl_down(s_fail, off, 64);
        const long long om = __shfl_down(s_max, off, 64);
        s_max = om > s_max ? om : s_max;
        const long long on = __shfl_down(s_min, off, 64);
        s_min = (s_min < 0) ? on : ((on >= 0 && on < s_min) ? on : s_min);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long*)&a.stats[0], (unsigned long long)s_att);
        atomicAdd((unsigned long long*)&a.stats[1], (unsigned long long)s_acc);
        atomicAdd((unsigned long long*)&a.stats[2], (unsigned long long)s_rej);
        atomicMax((unsigned long long*)&a.stats[3], (unsigned long long)s_max);
        if (s_min >= 0) atomicMin((unsigned long long*)&a.stats[4], (unsigned long long)s_min);
        atomicAdd((unsigned long long*)&a.stats[5], (unsigned long long)s_fail);
    }
}

/* One ROS23 step VJP for every cell, with the clamp mask of that step
 * (lam *= 1[uout > 0], adjoint_gradient.jl:736) applied first. */
extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK)
rs_chem_vjp(RsChemVjpArgs a) {
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= a.nc) return;
    double u[RS_NS], lam[RS_NS];
#pragma unroll
    for (int s = 0; s < RS_NS; ++s) {
        u[s] = a.u[(long long)s * a.nc + cell];
        double l = a.lam[(long long)s * a.nc + cell];
        if (a.uout && !(a.uout[(long long)s * a.nc + cell] > 0.0)) l = 0.0;
        lam[s] = l;
    }
    const double jf = a.jfac ? a.jfac[cell] : 1.0;
    const bool surf = cell < a.ncol;
    const double* Ec = (surf && a.E) ? a.E + cell : nullptr;
    double* Gc = (surf && a.gE) ? a.gE + cell : nullptr;
    rs_ros23_vjp(a.K, jf, Ec, a.ncol, u, a.dt, lam, Gc);
#pragma unroll
    for (int s = 0; s < RS_NS; ++s) a.lam[(long long)s * a.nc + cell] = lam[s];
}

#endif /* RS_HOST_BUILD */
