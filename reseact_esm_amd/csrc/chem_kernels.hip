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
 * E (may be 0) points at this cell's NS emissions (contiguous: the device
 * keeps the surface emissions cell-major, [ncol][NS]).
 *
 * Storage: three long-lived vectors sit in a slab sv with stride ST between
 * species (the lane's LDS slab on the device, ST = 64; a local array on the
 * host, ST = 1), reused as values die:
 *   slot A: 2(k1 - f0)  ->  + c32 (k2 - f1)       (the two terms of k3's rhs)
 *   slot B: k1          ->  k1 - 2 k2             (the error's first term)
 *   slot C: u on entry  ->  atol + rtol max(|u|, |unew|) (unless KEEP_U)
 * W (the LU factors, stride WS) is a register array for small mechanisms and
 * the lane's LDS slice for large ones; three short-lived vectors stay in
 * registers.  unew is
 * handed to out(i, unew_i) as soon as it exists.  The exact build
 * (RS_EXACT_DIV) keeps the reference's evaluation order (f2 - a) - c for k3's
 * rhs; the fast build adds a + c first. */
/* Phase boundary for the compiler only (no instruction): values of the slab
 * are read again from LDS after it instead of being held in registers across
 * the phases, which is what the slab is for. */
#ifdef RS_HOST_BUILD
#define RS_PHASE()
#else
#define RS_PHASE() __asm__ volatile("" ::: "memory")
#endif

/* Scheduling fences between the attempt's phases (no instruction), selected
 * by the bits of RS_AFENCE_MASK: 1 after f0, 2 after the LU, 4 after k1, 8
 * around f1, 16 after k2, 32 around f2, 64 after k3.  A fence keeps the
 * compiler from interleaving the phases on either side.  Measured on the
 * 20-species mechanism (one attempt over 616 k cells): no fence 0.107 ms
 * (196 B/lane of scratch), all fences 0.117 ms (no scratch, but the phases
 * no longer overlap), the fence after the LU alone 0.103 ms. */
#ifndef RS_AFENCE_MASK
#define RS_AFENCE_MASK 2
#endif
#ifdef RS_HOST_BUILD
#define RS_AFENCE(bit)
#else
#define RS_AFENCE(bit) do { if (RS_AFENCE_MASK & (bit)) __builtin_amdgcn_sched_barrier(0); } while (0)
#endif
#define RS_SA(i) sv[(i) * ST]
#define RS_SB(i) sv[(RS_NS + (i)) * ST]
#define RS_SC(i) sv[(2 * RS_NS + (i)) * ST]
template <int ST, int WS, bool KEEP_U, class Out>
RS_FN double rs_ros23_attempt(const double* __restrict__ K, double jf,
                              const double* __restrict__ E,
                              double* __restrict__ sv, double* __restrict__ W, double dt,
                              double atol, double rtol, Out out, bool unew_only = false) {
    /* two short-lived register vectors; the solves work in place */
    double f[RS_NS], y[RS_NS];
    rs_rhs<ST>(K, jf, E, &RS_SC(0), f);                        /* f0, :282 */
    RS_AFENCE(1);
    const double hg = dt * RS_ROS23_D;                             /* :291 */
    rs_jac_W<ST, WS>(K, jf, &RS_SC(0), hg, W);                     /* :287-296 */
    rs_lu<WS>(W);
    RS_AFENCE(2);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) RS_SA(i) = f[i];               /* f0, until k1 exists */
    rs_solve_ip<WS>(W, f);                                         /* k1, :298 */
    RS_AFENCE(4);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {                              /* :323-324 */
        const double c = f[i] - RS_SA(i);
        RS_SA(i) = 2.0 * c;
        RS_SB(i) = f[i];
    }
    const double dto2 = dt / 2;
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double du1 = dto2 * f[i]; y[i] = RS_SC(i) + du1; }
    RS_AFENCE(8);
    rs_rhs(K, jf, E, y, f);                                    /* f1, :303 */
    RS_AFENCE(8);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) y[i] = f[i] - RS_SB(i);        /* :306-309 */
    rs_solve_ip<WS>(W, y);                                         /* :310 */
    RS_AFENCE(16);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) y[i] = y[i] + RS_SB(i);        /* k2, :311-313 */
#ifdef RS_EXACT_DIV
    double c3[RS_NS];
#endif
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {                              /* :321-322 */
        const double a = y[i] - f[i];
#ifdef RS_EXACT_DIV
        c3[i] = RS_ROS23_C32 * a;
#else
        const double ca = RS_ROS23_C32 * a;
        RS_SA(i) = ca + RS_SA(i);
#endif
    }
    /* error denominators: slot C (u is dead by then), or registers when u
       must survive the attempt (KEEP_U) */
    double den[KEEP_U ? RS_NS : 1];
#define RS_DEN(i) den[KEEP_U ? (i) : 0]
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {
        const double u = RS_SC(i);
        const double du2 = dt * y[i];
        const double un = u + du2;
        out(i, un);
        const double am = rs_max(RS_FABS(u), RS_FABS(un));        /* :335-346 */
        double d = am * rtol;
        d = d + atol;
        if (KEEP_U) RS_DEN(i) = d;
        else RS_SC(i) = d;
        const double w2 = 2.0 * y[i];                              /* :333-334 */
        RS_SB(i) = RS_SB(i) - w2;
        y[i] = un;
    }
    /* A replay on a known step sequence needs unew alone: it stops here,
       before f2, k3 and the error estimate.  A uniform run-time flag rather
       than a second kernel, so that the replay executes the forward's own
       instructions up to this point and reproduces its unew bit for bit
       (separately compiled code may contract multiply-adds differently). */
    if (unew_only) return 0.0;
    RS_AFENCE(32);
    rs_rhs(K, jf, E, y, f);                                    /* f2, :318 */
    RS_AFENCE(32);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {                              /* :325-326 */
#ifdef RS_EXACT_DIV
        const double d3 = f[i] - c3[i];
        f[i] = d3 - RS_SA(i);
#else
        f[i] = f[i] - RS_SA(i);
#endif
    }
    rs_solve_ip<WS>(W, f);                                         /* k3, :328 */
    RS_AFENCE(64);
    const double dto6 = dt / 6;
    double sse = 0.0;
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {
        double ut = RS_SB(i) + f[i];
        ut = dto6 * ut;
        const double at = RS_DIV(ut, (KEEP_U ? RS_DEN(i) : RS_SC(i)));
        sse += at * at;
    }
    return sse;
}
#undef RS_DEN
#undef RS_SA
#undef RS_SB
#undef RS_SC

/* Hand-derived adjoint of the part of one ROS23 step that reaches unew
 * (f2, k3 and EEst are dead, rx_traced_integrator.jl:540-547):
 *   f0 = f(u); k1 = W\f0; u1 = u + h/2 k1; f1 = f(u1); k2 = W\(f1-k1) + k1;
 *   unew = u + h k2,   W = I - h*gamma*J(u).
 * Given lam = dL/dunew, returns dL/du in lam and adds dL/dE to G[i*gs] (G
 * may be 0; E, this cell's emissions as for the attempt, enters f additively
 * on the cells that have it); returns the
 * number of non-finite values among those dL/dE contributions.  Reverse:
 *   z2 = W^-T (h lam)             = f1bar;  k1bar = h lam - z2
 *   u1bar = J(u1)^T z2;  ubar = lam + u1bar;  k1bar += h/2 u1bar
 *   z1 = W^-T k1bar               = f0bar;  ubar += J(u)^T z1
 *   Jbar = h gamma (z2 (k2-k1)^T + z1 k1^T)  ->  ubar += d/du [z^T J(u) v]
 *   Ebar = z1 + z2.
 * The stages are recomputed from u, not stored.  Slab sv (stride ST, as for
 * the attempt): slot A holds u on entry, slot B gets u1, slot C gets k1.
 * lam is read through lam_in(i) when the reverse sweep starts and written
 * through lam_out(i, v), so it is not held across the forward recompute. */
#define RS_SA(i) sv[(i) * ST]
#define RS_SB(i) sv[(RS_NS + (i)) * ST]
#define RS_SC(i) sv[(2 * RS_NS + (i)) * ST]
template <int ST, int WS, class LIn, class LOut>
RS_FN int rs_ros23_vjp(const double* __restrict__ K, double jf, const double* __restrict__ E,
                        long long gs, double* __restrict__ sv, double* __restrict__ W, double dt,
                        LIn lam_in, LOut lam_out, double* __restrict__ G) {
    double f[RS_NS], x[RS_NS], y[RS_NS];
    rs_rhs<ST>(K, jf, E, &RS_SA(0), f);
    const double hg = dt * RS_ROS23_D;
    RS_PHASE();
    rs_jac_W<ST, WS>(K, jf, &RS_SA(0), hg, W);
    rs_lu<WS>(W);
    rs_solve<WS>(W, f, x);                                         /* k1 */
    RS_PHASE();
    const double dto2 = dt / 2;
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) {
        RS_SC(i) = x[i];
        const double du1 = dto2 * x[i];
        RS_SB(i) = RS_SA(i) + du1;                                 /* u1 */
    }
    RS_PHASE();
    rs_rhs<ST>(K, jf, E, &RS_SB(0), f);                        /* f1 */
    RS_PHASE();
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) f[i] = f[i] - x[i];
    rs_solve<WS>(W, f, y);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) y[i] = y[i] + x[i];            /* k2 */
    /* reverse */
    double ub[RS_NS];
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { const double l = lam_in(i); f[i] = dt * l; ub[i] = l; }  /* k2bar */
    rs_solveT<WS>(W, f, x);                                        /* z2 */
    RS_PHASE();
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) { f[i] = f[i] - x[i]; y[i] = y[i] - RS_SC(i); }  /* k1bar; v = k2 - k1 */
    rs_hess(K, jf, x, y, hg, ub);
    RS_PHASE();
    /* ubar += J(u1)^T z2, k1bar += h/2 J(u1)^T z2 */
    rs_jtv2<ST>(K, jf, &RS_SB(0), x, ub, f, dto2);
    RS_PHASE();
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) y[i] = x[i];                   /* keep z2 for Ebar */
    rs_solveT<WS>(W, f, x);                                        /* z1 */
    RS_PHASE();
    rs_jtv<ST>(K, jf, &RS_SA(0), x, ub);
    RS_PHASE();
    rs_hess<ST>(K, jf, x, &RS_SC(0), hg, ub);
#pragma unroll
    for (int i = 0; i < RS_NS; ++i) lam_out(i, ub[i]);
    int nbad = 0;
    if (G) {
#pragma unroll
        for (int i = 0; i < RS_NS; ++i) {
            const double g = x[i] + y[i];
            nbad += !RS_ISFINITE(g);
            G[i * gs] += g;
        }
    }
    return nbad;
}
#undef RS_SA
#undef RS_SB
#undef RS_SC

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

/* Sum over the workgroup (whole waves), deterministic: a fixed shuffle tree
 * inside each wave, then wave order; valid in lane 0. */
__device__ inline double rs_block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    return s;
}

/* Rate constants: a kernel argument (scalar loads from the argument segment,
 * uniform across the wave) for mechanisms up to RS_K_INLINE_MAX reactions,
 * otherwise a pointer. */
#if RS_NR > 0 && RS_NR <= RS_K_INLINE_MAX
typedef struct { double k[RS_NR]; } RsKTab;
#define RS_KPARAM , RsKTab kt
#define RS_KPTR(a) (kt.k)
#else
#define RS_KPARAM
#define RS_KPTR(a) ((a).K)
#endif

/* Storage of a lane's cell.  RS_WDECL(W): the LU factors, a register array
 * (stride RS_W_ST = 1).  RS_SLAB(sv): three state vectors with stride
 * RS_SLAB_ST, in LDS ([3][NS][64] per wave, sized by the launcher) when four
 * waves per CU fit, otherwise a local array. */
#define RS_W_ST 1
#define RS_WDECL(name) double name##_arr[RS_NNZ]; double* name = name##_arr
#if RS_SLOW_DOUBLES(RS_NS) > 0
#define RS_SLAB_ST 64
#define RS_SLAB(name) extern __shared__ double rs_slab_mem[]; \
    double* name = rs_slab_mem + (threadIdx.x >> 6) * (3 * RS_NS * 64) + (threadIdx.x & 63)
#else
#define RS_SLAB_ST 1
#define RS_SLAB(name) double name##_arr[3 * RS_NS]; double* name = name##_arr
#endif

/* Kernel groups of a split compile: with RS_JIT_ONLY=n only group n is
 * emitted (1-3 the three attempts, 4 the per-cell solve, 5 its adjoint, 6 and
 * 7 the two VJPs), so the groups compile in parallel; without it, all. */
#ifndef RS_JIT_ONLY
#define RS_JIT_PART(n) 1
#else
#define RS_JIT_PART(n) (RS_JIT_ONLY == (n))
#endif

/* One attempt over all cells, in three entry points: the rate constants as a
 * kernel argument (scalar loads; RS_KPARAM) or through the pointer in the
 * argument block (vector loads, fewer SGPRs held), the autotuner keeping the
 * faster on the context's grid; and, for positive-domain solves only, the
 * pointer form that also counts negative entries of the trial state per
 * workgroup (a compile-time switch: the count costs the other two 6%).  The
 * body is spelled out in each kernel (a macro): shared through a device
 * function, the compiler spilled 100-140 B/lane more. */
#define RS_CHEM_ATTEMPT_KERNEL(NAME, KPARAM, KEXPR, POSDOM)                                 \
    extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK) NAME(RsChemArgs a KPARAM) {  \
        __shared__ double sh[RS_CHEM_BLOCK / 64];                                           \
        RS_SLAB(sv);                                                                        \
        RS_WDECL(W);                                                                        \
        const RsCtrl* c = a.ctrl;                                                           \
        if (c && c->done) return;                                                           \
        const int cur = c ? c->cur : a.cur_static;                                          \
        const double dt = c ? c->dtc : a.dt_static;                                         \
        const double* __restrict__ src = cur ? a.S1 : a.S0;                                 \
        double* __restrict__ dst = cur ? a.S0 : a.S1;                                       \
        const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;            \
        double sse = 0.0, neg = 0.0;                                                        \
        if (cell < a.nc) {                                                                  \
            _Pragma("unroll") for (int s = 0; s < RS_NS; ++s)                               \
                sv[(2 * RS_NS + s) * RS_SLAB_ST] = src[(long long)s * a.nc + cell];         \
            const double jf = a.jfac ? a.jfac[cell] : 1.0;                                  \
            const double* Ec = (a.E && cell < a.ncol) ? a.E + cell * RS_NS : nullptr;       \
            const int clamp = a.clamp;                                                      \
            sse = rs_ros23_attempt<RS_SLAB_ST, RS_W_ST, false>(                             \
                KEXPR, jf, Ec, sv, W, dt, a.atol, a.rtol, [&](int s, double un) {           \
                    if (POSDOM && un < 0.0) neg += 1.0;                                     \
                    dst[(long long)s * a.nc + cell] = clamp ? rs_max(un, 0.0) : un;         \
                }, a.unew_only != 0);                                                       \
        }                                                                                   \
        const double tot = rs_block_sum(sse, sh);                                           \
        if (threadIdx.x == 0) a.partial[blockIdx.x] = tot;                                  \
        if (POSDOM) {                                                                       \
            __syncthreads();                                                                \
            const double nt = rs_block_sum(neg, sh);                                        \
            if (threadIdx.x == 0) a.negpart[blockIdx.x] = nt;                               \
        }                                                                                   \
    }

#if RS_JIT_PART(1)
RS_CHEM_ATTEMPT_KERNEL(rs_chem_attempt, RS_KPARAM, RS_KPTR(a), 0)
#endif
#if RS_JIT_PART(2)
RS_CHEM_ATTEMPT_KERNEL(rs_chem_attempt_kp, , a.K, 0)
#endif
#if RS_JIT_PART(3)
RS_CHEM_ATTEMPT_KERNEL(rs_chem_attempt_pd, , a.K, 1)
#endif

/* Persistent per-cell adaptive solve with wave-aggregated work stealing: a
 * lane that finishes its cell takes the next index from the queue head; all
 * lanes run the same attempt code, each on its own cell and its own dt. */
/* RS_PERCELL_WAVES=n caps the registers so n waves fit per SIMD (more spill,
 * more latency hiding); default: the compiler's choice. */
#ifdef RS_PERCELL_WAVES
#define RS_PERCELL_ATTR __attribute__((amdgpu_waves_per_eu(RS_PERCELL_WAVES, RS_PERCELL_WAVES)))
#else
#define RS_PERCELL_ATTR
#endif
#if RS_JIT_PART(4)
extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK) RS_PERCELL_ATTR
rs_chem_percell(RsPercellArgs a RS_KPARAM) {
    const int lane = threadIdx.x & 63;
    long long cell = -1;     /* -1: needs work, -2: queue drained */
    RS_SLAB(sv);
    RS_WDECL(W);
#define RS_U(s) sv[(2 * RS_NS + (s)) * RS_SLAB_ST]
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
                    for (int s = 0; s < RS_NS; ++s) RS_U(s) = a.u[(long long)s * a.nc + cell];
                    jf = a.jfac ? a.jfac[cell] : 1.0;
                    Ec = (a.E && cell < a.ncol) ? a.E + cell * RS_NS : nullptr;
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
#define RS_UN(s) un[s]
                const double sse = rs_ros23_attempt<RS_SLAB_ST, RS_W_ST, true>(
                    RS_KPTR(a), jf, Ec, sv, W, dtc, a.atol, a.rtol,
                    [&](int s, double v) { RS_UN(s) = v; });
                const double eest = RS_SQRT(sse / (double)RS_NS);
                bool isout = false;
                if (a.posdom) {
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) isout = isout || RS_UN(s) < 0.0;
                }
                if (isout) {
                    /* out-of-domain trial state: retry at half the step */
                    dt = dtc / 2.0;
                    nrej += 1;
                } else if (rs_pi_update(a.beta1, a.beta2, a.invqmin, a.invqmax, a.gamma, a.qsmin,
                                        a.qsmax, a.qoldinit, dtc, eest, &t, &dt, &qold, &nacc, &nrej)) {
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) RS_U(s) = a.clamp ? rs_max(RS_UN(s), 0.0) : RS_UN(s);
                }
#undef RS_UN
                iters += 1;
            }
            if (!((t < tlim) && (iters < a.maxiters))) {
#pragma unroll
                for (int s = 0; s < RS_NS; ++s) a.u[(long long)s * a.nc + cell] = RS_U(s);
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

#undef RS_U
#endif

/* Adjoint of the per-cell solve (the reference's backward sweep,
 * adjoint_gradient.jl:731-834, over each cell's own accepted sequence).
 * Persistent with the same wave-aggregated queue as rs_chem_percell; a lane
 * takes a cell, integrates it again from u with the same attempt code (so
 * the same trajectory), recording in its own slice of a.ws the state before
 * every accepted step, the step, and the end state; then sweeps the record
 * backwards: lam .*= 1[u_{j+1} > 0] when clamping (the accepted state is
 * max(u, 0)), lam <- VJP of step j at u_j (rs_ros23_vjp; lam lives in a.lam,
 * read and written per step).  Lanes of a wave move through the two phases
 * independently.  A cell that would need more than a.nmax recorded states
 * is counted in stats[6] and left without its sweep (the host fails the
 * call). */
#if RS_JIT_PART(5)
extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK) RS_PERCELL_ATTR
rs_chem_percell_adj(RsPercellAdjArgs a RS_KPARAM) {
    const int lane = threadIdx.x & 63;
    const long long NL = a.nlanes;
    double* __restrict__ ws = a.ws + ((long long)blockIdx.x * blockDim.x + threadIdx.x);
    long long cell = -1;     /* -1: needs work, -2: queue drained */
    int mode = 0;            /* 1: forward with record, 2: backward sweep */
    RS_SLAB(sv);
    RS_WDECL(W);
#define RS_U(s) sv[(2 * RS_NS + (s)) * RS_SLAB_ST]
#define RS_REC(j, s) ws[((j) * (RS_NS + 1) + (s)) * NL]
    const double* Ec = nullptr;
    double* Gc = nullptr;
    double t = 0.0, dt = 0.0, qold = 0.0, jf = 1.0;
    long long iters = 0, nacc = 0, nrej = 0, nst = 0, q = -1;
    unsigned long long s_att = 0, s_acc = 0, s_rej = 0, s_max = 0, s_fail = 0, s_vjp = 0,
                       s_ovf = 0, s_nl = 0, s_ng = 0;
    bool ovf = false;
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
                    for (int s = 0; s < RS_NS; ++s) {
                        const double v = a.u[(long long)s * a.nc + cell];
                        RS_U(s) = v;
                        RS_REC(0, s) = v;
                    }
                    jf = a.jfac ? a.jfac[cell] : 1.0;
                    Ec = (a.E && cell < a.ncol) ? a.E + cell * RS_NS : nullptr;
                    Gc = (a.gE && cell < a.ncol) ? a.gE + cell : nullptr;
                    t = a.t0;
                    dt = a.dt0;
                    qold = a.qoldinit;
                    iters = nacc = nrej = nst = 0;
                    ovf = false;
                    mode = 1;
                } else {
                    cell = -2;
                }
            }
        }
        if (!__any(cell >= 0)) break;
        if (cell >= 0 && mode == 1) {
            if ((t < tlim) && (iters < a.maxiters)) {
                const double dtc = (dt < a.t1 - t) ? dt : a.t1 - t;
                double un[RS_NS];
                const double sse = rs_ros23_attempt<RS_SLAB_ST, RS_W_ST, true>(
                    RS_KPTR(a), jf, Ec, sv, W, dtc, a.atol, a.rtol,
                    [&](int s, double v) { un[s] = v; });
                const double eest = RS_SQRT(sse / (double)RS_NS);
                bool isout = false;
                if (a.posdom) {
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) isout = isout || un[s] < 0.0;
                }
                if (isout) {
                    dt = dtc / 2.0;
                    nrej += 1;
                } else if (rs_pi_update(a.beta1, a.beta2, a.invqmin, a.invqmax, a.gamma, a.qsmin,
                                        a.qsmax, a.qoldinit, dtc, eest, &t, &dt, &qold, &nacc, &nrej)) {
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) RS_U(s) = a.clamp ? rs_max(un[s], 0.0) : un[s];
                    if (nst + 1 < a.nmax) {
                        RS_REC(nst, RS_NS) = dtc;
                        nst += 1;
#pragma unroll
                        for (int s = 0; s < RS_NS; ++s) RS_REC(nst, s) = RS_U(s);
                    } else {
                        ovf = true;
                    }
                }
                iters += 1;
            }
            if (ovf || !((t < tlim) && (iters < a.maxiters))) {
                if (a.uend) {
#pragma unroll
                    for (int s = 0; s < RS_NS; ++s) a.uend[(long long)s * a.nc + cell] = RS_U(s);
                }
                s_att += iters; s_acc += nacc; s_rej += nrej;
                s_max = (unsigned long long)iters > s_max ? (unsigned long long)iters : s_max;
                if (ovf) {
                    s_ovf += 1;
                    cell = -1;
                    mode = 0;
                } else {
                    if (t < tlim) s_fail += 1;
                    mode = 2;
                    q = nst - 1;
                }
            }
        } else if (cell >= 0 && mode == 2) {
            if (q >= 0) {
#pragma unroll
                for (int s = 0; s < RS_NS; ++s) sv[s * RS_SLAB_ST] = RS_REC(q, s);   /* slot A: u_q */
                const double dq = RS_REC(q, RS_NS);
                const long long qn = q + 1;
                const long long cb = cell;
                s_ng += (unsigned long long)rs_ros23_vjp<RS_SLAB_ST, RS_W_ST>(
                    RS_KPTR(a), jf, Ec, a.ncol, sv, W, dq,
                    [&](int s) {
                        double l = a.lam[(long long)s * a.nc + cb];
                        if (a.clamp) l = l * ((RS_REC(qn, s) > 0.0) ? 1.0 : 0.0);
                        return l;
                    },
                    [&](int s, double v) { a.lam[(long long)s * a.nc + cb] = v; }, Gc);
                s_vjp += 1;
                q -= 1;
            }
            if (q < 0) {
#pragma unroll
                for (int s = 0; s < RS_NS; ++s) s_nl += !RS_ISFINITE(a.lam[(long long)s * a.nc + cell]);
                cell = -1;
                mode = 0;
            }
        }
    }
#undef RS_REC
    /* wave reduction of the statistics, one atomic per counter and wave */
    for (int off = 32; off > 0; off >>= 1) {
        s_att += __shfl_down(s_att, off, 64);
        s_acc += __shfl_down(s_acc, off, 64);
        s_rej += __shfl_down(s_rej, off, 64);
        s_fail += __shfl_down(s_fail, off, 64);
        s_vjp += __shfl_down(s_vjp, off, 64);
        s_ovf += __shfl_down(s_ovf, off, 64);
        s_nl += __shfl_down(s_nl, off, 64);
        s_ng += __shfl_down(s_ng, off, 64);
        const unsigned long long om = __shfl_down(s_max, off, 64);
        s_max = om > s_max ? om : s_max;
    }
    if (lane == 0) {
        atomicAdd(&a.stats[0], s_att);
        atomicAdd(&a.stats[1], s_acc);
        atomicAdd(&a.stats[2], s_rej);
        atomicMax(&a.stats[3], s_max);
        atomicAdd(&a.stats[4], s_fail);
        atomicAdd(&a.stats[5], s_vjp);
        atomicAdd(&a.stats[6], s_ovf);
        atomicAdd(&a.stats[7], s_nl);
        atomicAdd(&a.stats[8], s_ng);
        if (a.nfc && s_ng) atomicAdd(&a.nfc[1], s_ng);
    }
}

#undef RS_U
#endif

/* One ROS23 step VJP for every cell, with the clamp mask of that step
 * (lam .*= 1[uout > 0], adjoint_gradient.jl:736; a multiplication, so a
 * non-finite entry stays non-finite) applied first.  With a.nfc the step's
 * non-finite outputs are counted (adjoint_gradient.jl:763-779): a lane with
 * none issues no atomic. */
/* MB = 1 (rs_chem_vjp_mb, species counts up to 32): the clamp mask comes
 * from a.mask_in, the sign bits of the step output written by the VJP of the
 * step after (one 32-bit word per cell instead of NS doubles); MB = 0
 * (rs_chem_vjp): from a.uout, or none.  Either writes the sign bits of its
 * own input u to a.mask_out when that is given.  The factor is 1.0 or 0.0
 * from the same predicate u > 0, so both give the same lambda. */
#define RS_CHEM_VJP_KERNEL(NAME, MB)                                                       \
    extern "C" __global__ void __launch_bounds__(RS_CHEM_BLOCK) NAME(RsChemVjpArgs a RS_KPARAM) { \
        const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;          \
        if (cell >= a.nc) return;                                                         \
        RS_SLAB(sv);                                                                      \
        RS_WDECL(W);                                                                      \
        _Pragma("unroll") for (int s = 0; s < RS_NS; ++s)                                 \
            sv[s * RS_SLAB_ST] = a.u[(long long)s * a.nc + cell];                         \
        if (RS_NS <= 32 && a.mask_out) {                                                  \
            unsigned mo = 0;                                                              \
            _Pragma("unroll") for (int s = 0; s < RS_NS; ++s)                             \
                mo |= (sv[s * RS_SLAB_ST] > 0.0 ? 1u : 0u) << (s & 31);                   \
            a.mask_out[cell] = mo;                                                        \
        }                                                                                 \
        const double jf = a.jfac ? a.jfac[cell] : 1.0;                                    \
        const bool surf = cell < a.ncol;                                                  \
        const double* Ec = (surf && a.E) ? a.E + cell * RS_NS : nullptr;                  \
        double* Gc = (surf && a.gE) ? a.gE + cell : nullptr;                              \
        unsigned nlam = 0;                                                                \
        const int ngrad = rs_ros23_vjp<RS_SLAB_ST, RS_W_ST>(                              \
            RS_KPTR(a), jf, Ec, a.ncol, sv, W, a.dt,                                      \
            [&](int s) {                                                                  \
                double l = a.lam[(long long)s * a.nc + cell];                             \
                if (MB) {                                                                 \
                    const unsigned bit = (a.mask_in[cell] >> (s & 31)) & 1u;              \
                    l = l * (bit ? 1.0 : 0.0);                                            \
                } else if (a.uout) {                                                      \
                    l = l * ((a.uout[(long long)s * a.nc + cell] > 0.0) ? 1.0 : 0.0);     \
                }                                                                         \
                return l;                                                                 \
            },                                                                            \
            [&](int s, double v) {                                                        \
                nlam += !RS_ISFINITE(v);                                                  \
                a.lam[(long long)s * a.nc + cell] = v;                                    \
            }, Gc);                                                                       \
        if (a.nfc) {                                                                      \
            if (nlam) atomicAdd(&a.nfc[0], (unsigned long long)nlam);                     \
            if (ngrad) atomicAdd(&a.nfc[1], (unsigned long long)ngrad);                   \
        }                                                                                 \
    }

#if RS_JIT_PART(6)
RS_CHEM_VJP_KERNEL(rs_chem_vjp, 0)
#endif
#if RS_JIT_PART(7) && RS_NS <= 32   /* one mask bit per species */
RS_CHEM_VJP_KERNEL(rs_chem_vjp_mb, 1)
#endif

#endif /* RS_HOST_BUILD */
