/* rs_launch.h -- argument blocks shared by the host code, the statically built
 * kernels and the runtime-compiled (per-mechanism) chemistry kernels.  Plain C
 * layout: the same text is compiled by hipcc and embedded into the hipRTC
 * source, so both sides agree on every offset. */
#ifndef RS_LAUNCH_H
#define RS_LAUNCH_H

/* Device-resident state of one PI-controlled adaptive solve with a single
 * global error norm (rx_traced_integrator.jl:455-507). */
typedef struct RsCtrl {
    double t, t1, tlim, dt, dtc, qold, eest, nnorm;
    double beta1, beta2, invqmin, invqmax, gamma, qsmin, qsmax, qoldinit;
    long long nacc, nrej, iters, maxiters;
    int done, accept, cur, nlog, logcap, pad;
} RsCtrl;

/* LDS slab of the chemistry kernels, in doubles per lane (0 = none): three
 * state vectors per lane, laid out per wave as [3][ns][64] so that a wave's
 * accesses to one species are 64 consecutive doubles; used while four waves
 * (one per SIMD) still fit in the 160 KiB of LDS of a CU. */
#define RS_SLOW_DOUBLES(ns) ((3 * (ns) * 8 * 256 <= 160 * 1024) ? 3 * (ns) : 0)

/* Mechanisms with at most this many reactions get their rate constants as a
 * kernel argument (read with scalar loads) instead of through a pointer. */
#define RS_K_INLINE_MAX 448

/* One ROS23 attempt over all cells.  With ctrl != 0 the step size, the
 * current state buffer and the done flag are read from it (adaptive mode). */
typedef struct RsChemArgs {
    double* S0;
    double* S1;
    const RsCtrl* ctrl;
    const double* K;        /* rate constants, nr          */
    const double* jfac;     /* photolysis factor, nc or 0  */
    const double* E;        /* surface emissions, ns*ncol or 0 */
    double* partial;        /* one sse per block           */
    long long nc, ncol;
    double dt_static, atol, rtol;
    int cur_static, clamp;
    double* negpart;        /* positive domain: negative entries per block, or 0 */
    int posdom;
    int unew_only;          /* 1: stop after unew (the fixed-sequence replay; static dt) */
} RsChemArgs;

/* Per-cell adaptive chemistry, persistent work-stealing launch. */
typedef struct RsPercellArgs {
    double* u;              /* state, in place             */
    const double* K;
    const double* jfac;
    const double* E;
    double* dtcell;         /* per-cell dt in/out, or 0    */
    unsigned long long* head;   /* work-queue head (zeroed before launch) */
    long long* stats;       /* attempts, accepts, rejects, max, min, failed */
    long long nc, ncol, maxiters;
    double t0, t1, dt0, atol, rtol;
    double beta1, beta2, invqmin, invqmax, gamma, qsmin, qsmax, qoldinit;
    int clamp, posdom;      /* posdom: reject a trial state with a negative entry */
} RsPercellArgs;

/* Adjoint of the per-cell adaptive chemistry, persistent work-stealing
 * launch: each lane integrates its cell again from u recording the accepted
 * steps in its own slice of ws (lane stride nlanes; per recorded state j:
 * the NS values at [j (NS+1) + s] and the step that leaves it at
 * [j (NS+1) + NS]), then sweeps them backwards. */
typedef struct RsPercellAdjArgs {
    const double* u;        /* chemistry input state       */
    double* uend;           /* end state of every cell, or 0 */
    double* lam;            /* in/out                      */
    double* gE;             /* emission gradient, ns*ncol, or 0 */
    const double* K;
    const double* jfac;
    const double* E;
    double* ws;             /* nmax * (ns + 1) * nlanes doubles */
    unsigned long long* head;   /* work-queue head (zeroed before launch) */
    unsigned long long* stats;  /* attempts, accepts, rejects, max attempts, failed cells,
                                   step VJPs, record overflows, non-finite lam outputs,
                                   non-finite emission-gradient contributions (zeroed) */
    unsigned long long* nfc;    /* or 0: [1] += non-finite emission-gradient contributions */
    long long nc, ncol, maxiters, nmax, nlanes;
    double t0, t1, dt0, atol, rtol;
    double beta1, beta2, invqmin, invqmax, gamma, qsmin, qsmax, qoldinit;
    int clamp, posdom;
} RsPercellAdjArgs;

/* One ROS23 step VJP over all cells (stages recomputed from u). */
typedef struct RsChemVjpArgs {
    const double* u;        /* step input state            */
    const double* uout;     /* step output (clamp mask), or 0 */
    double* lam;            /* in/out                      */
    double* gE;             /* emission gradient, ns*ncol, or 0 */
    const double* K;
    const double* jfac;
    const double* E;
    long long nc, ncol;
    double dt;
    unsigned long long* nfc;    /* or 0: [0] += non-finite lam outputs, [1] += non-finite
                                   emission-gradient contributions of this step */
    /* The clamp mask as sign bits, one word per cell (bit s: species s > 0;
     * species counts up to 32): rs_chem_vjp_mb reads mask_in instead of uout;
     * mask_out, if given, gets the bits of u, which are the mask of the step
     * before in the backward sweep. */
    const unsigned* mask_in;
    unsigned* mask_out;
} RsChemVjpArgs;

/* One transport stage over all transported species (see transport.hip). */
typedef struct RsTransArgs {
    double* S0;
    double* S1;
    const double* a;        /* stage input                 */
    const double* b;        /* second input (last stage / adjoint) */
    double* out;
    const double* cx;       /* nlat   */
    const double* wz;       /* nlev+1 */
    const double* wl;       /* nlon   */
    const RsCtrl* ctrl;
    double* partial;
    long long nc;
    double dt_static, atol, rtol;
    int nlon, nlat, nlev, nt;
    int cur_static, a_is_state, out_is_state, clamp;
    /* fused kernels: longitude segment of seg_w columns (seg_in interior plus
     * seg_halo on each side; seg_halo = 0 means the whole periodic row); each
     * lane owns one column and seg_lev consecutive levels, seg_chunks lanes
     * per column */
    int seg_w, seg_in, seg_halo, nseg, seg_lev, seg_chunks;
    double* negpart;        /* positive domain: negative entries per block, or 0 */
    int posdom;
    int fnth;               /* fused kernels: lanes per workgroup (1024 or 512) */
    /* fused adjoint only: lambda is multiplied by 1[lmask > 0] as it is read
     * (the clamp mask of the step output), and *nfcount gets the number of
     * non-finite values written; 0 = neither */
    const double* lmask;
    unsigned long long* nfcount;
} RsTransArgs;

/* Monotone PPM transport stage (CWC mixing-ratio form; species 0 = air
 * mass).  Stage input a (or the state buffer); the third stage also writes
 * a4 from u (b); the last stage reads a4 (c4), u (a) and u3 (b). */
typedef struct RsPPMArgs {
    double* S0;
    double* S1;
    const double* a;        /* stage input                               */
    const double* b;        /* u (third stage) or u3 (last stage)        */
    const double* c4;       /* last stage: a4 = 2/3 u + 1/3 u3           */
    double* out;
    double* out2;           /* third stage: a4                            */
    const RsCtrl* ctrl;
    double* partial;
    double* negpart;
    const double* Mx;       /* nlev x nlat x (nlon+1) */
    const double* My;       /* nlev x (nlat+1) x nlon */
    const double* Mz;       /* (nlev+1) x nlat x nlon */
    const double* ce;       /* nlat+1: meridional face lengths */
    const double* gx;       /* nlat: dphi_lat / (dlon dS_lat) */
    const double* rdS;      /* nlat: 1 / dS_lat */
    const double* vs;       /* vertical slope factors, 2 x (nlev+4) x ncol */
    const double* ve;       /* vertical edge factors, 3 x (nlev+3) x ncol */
    const double* qw;       /* inflow halos of species 1..nt-1 */
    const double* qe;
    const double* qs;
    const double* qn;
    const double* msrc;     /* air-mass source of the pressure fixer, or 0 */
    double* outf;           /* stages 1-3: also the advective tendency at a (adjoint), or 0 */
    const double* pblf;     /* PBL mixing: layer fraction in the PBL (nc), or 0 (off) */
    const double* pblw;     /* dp * pbl_f (nc) */
    const double* pblWp;    /* max(column sum of pblw, 1) (ncol) */
    double* pblm;           /* PBL means of the stencil source ((nt-1) x ncol) */
    double pbl_tau;
    long long nc;
    double dt_static, atol, rtol;
    int nlon, nlat, nlev, nt;
    int cur_static, a_is_state, out_is_state, clamp;
    int lon_periodic, posdom, b_is_state;
    int kblk;               /* levels per workgroup (set by the launcher) */
} RsPPMArgs;

/* Adjoint of the PPM tendency at the state `at` (species 0 = air mass):
 * out = [beta b +] x + h2 J(at)^T x with x = alpha a, over the transported
 * block; a and out must differ (each cell gathers from its neighbours). */
typedef struct RsPPMVjpArgs {
    const double* at;
    const double* a;
    const double* b;        /* or 0 */
    double* out;
    const double* Mx;
    const double* My;
    const double* Mz;
    const double* ce;
    const double* gx;
    const double* rdS;
    const double* vs;
    const double* ve;
    const double* qw;
    const double* qe;
    const double* qs;
    const double* qn;
    const double* pblf;     /* PBL mixing (as RsPPMArgs), or 0 */
    const double* pblw;
    const double* pblWp;
    long long nc;
    double h2, alpha, beta, pbl_tau;
    int nlon, nlat, nlev, nt, lon_periodic, pad;
} RsPPMVjpArgs;

#endif
