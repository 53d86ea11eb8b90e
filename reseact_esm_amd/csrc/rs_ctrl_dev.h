/* rs_ctrl_dev.h -- device-side PI controller shared by the transport kernels
 * (built by hipcc) and the runtime-compiled chemistry kernels (hipRTC).
 *
 * rs_ctrl_arrive is called by every block of an attempt kernel with its
 * partial sum of squared scaled errors.  The last block to arrive sums all
 * partials in a fixed order (one wave, stride 64, then a fixed shuffle tree:
 * the same order whatever the block size) and applies the PI update of
 * rx_traced_integrator.jl:480-505 to the controller block, so an attempt is
 * one kernel launch and EEst is deterministic. */
#ifndef RS_CTRL_DEV_H
#define RS_CTRL_DEV_H

__device__ inline double rs_cmax(double a, double b) { return (a < b) ? b : a; }

/* PI update for the attempt whose total squared error is sse (one thread). */
__device__ inline void rs_ctrl_pi(RsCtrl* c, double sse, double* seqlog) {
    const double raw = sqrt(sse / c->nnorm);
    const double EEst = (raw != raw) ? 1.0e10 : raw;                 /* :484 */
    const double q11 = pow(rs_cmax(EEst, 1.0e-35), c->beta1);        /* :486 */
    const double qden = pow(c->qold, c->beta2);                      /* :487 */
    double q = q11 / qden;
    const double qg = q / c->gamma;
    q = rs_cmax(c->invqmax, (c->invqmin < qg) ? c->invqmin : qg);    /* :489 */
    const int accept = EEst <= 1.0;
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
        c->qold = rs_cmax(EEst, c->qoldinit);
        c->nacc += 1;
        c->cur ^= 1;
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

/* part0: this block's partial (meaningful in thread 0).  flag: a __shared__ int. */
__device__ inline void rs_ctrl_arrive(RsCtrl* c, double* partial, unsigned nblocks, unsigned bid,
                                      double part0, double* seqlog, int* flag) {
    if (threadIdx.x == 0) {
        partial[bid] = part0;
        __threadfence();
        const unsigned prev = atomicAdd(&c->arrive, 1u);
        *flag = (prev == nblocks - 1u) ? 1 : 0;
    }
    __syncthreads();
    if (!*flag) return;
    __threadfence();
    if (threadIdx.x < 64) {
        const volatile double* vp = partial;
        double s = 0.0;
        for (unsigned x = threadIdx.x; x < nblocks; x += 64u) s += vp[x];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if (threadIdx.x == 0) {
            c->arrive = 0u;
            rs_ctrl_pi(c, s, seqlog);
        }
    }
}

#endif
