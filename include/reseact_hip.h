/*
 * reseact_hip.h -- C ABI of the operator-split step library (transport SSPRK43 +
 * per-cell Rosenbrock23 chemistry + discrete adjoint), FP64 throughout.
 *
 * Two shared libraries export exactly this ABI:
 *   libreseact_hip.so   hand-written HIP kernels for gfx950 (the product);
 *   libreseact_cpu.so   the CPU restatement used as the parity oracle and the
 *                       CPU baseline (built from oracle/, never shipped).
 *
 * Conventions
 *   * State vectors are SPECIES-MAJOR, longitude fastest: element
 *     u[s*NC + c] with c = i + NLON*(j + NLAT*k), i = lon, j = lat, k = level
 *     (k = 0 is the surface).  Same contract as the reference's
 *     tools/reactant_handoff/rx_traced_integrator.jl:39-43.
 *   * Every call returns an int status: RS_OK (0) or one of RS_E_*; the
 *     message of the last failure is rs_last_error(ctx).
 *   * A context owns all device memory; the caller owns host buffers.  Calls on
 *     one context are not thread-safe; use one context per host thread.
 *   * The transport operator acts on species 0..NT-1; species NT..NS-1 are
 *     left bit-for-bit untouched by it.  Error norms are always taken over the
 *     full NS*NC state, as the reference's operators both see the full state
 *     (tools/reactant_handoff/op_split.jl:107-109).
 */
#ifndef RESEACT_HIP_H
#define RESEACT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define RS_API __attribute__((visibility("default")))
#else
#define RS_API
#endif

#define RS_OK           0
#define RS_E_ARG        1   /* bad argument / call order                       */
#define RS_E_HIP        2   /* a HIP runtime call failed                       */
#define RS_E_NONFINITE  3   /* a non-finite value where none is allowed        */
#define RS_E_MAXITERS   4   /* adaptive loop hit maxiters (adjoint_gradient.jl:580) */
#define RS_E_STATE      5   /* missing setup (grid / mechanism / state)        */
#define RS_E_JIT        6   /* runtime compilation of the mechanism kernels    */

typedef struct rs_ctx rs_ctx;

/* PI step-size controller.  Reference: PICtrl, rx_traced_integrator.jl:86-95;
 * defaults pictrl_ros23 / pictrl_ssprk43 at :98-99. */
typedef struct rs_pictrl {
    double beta1, beta2;
    double qmin, qmax, gamma;
    double qsteady_min, qsteady_max;
    double qoldinit;
} rs_pictrl;

/* Outcome of one adaptive solve (adaptive_solve returns (u, t, dt, nacc, nrej),
 * rx_traced_integrator.jl:455-507). */
typedef struct rs_counts {
    long long naccept, nreject;
    double t_end;     /* time reached                                          */
    double dt_end;    /* proposed dt carried into the next solve                */
} rs_counts;

/* Clamp flags of the adaptive solves (the `clamp_nonneg` argument and field):
 * RS_CLAMP_NONNEG (1): max(u, 0) on every accepted inner step
 *   (rx_traced_integrator.jl:466-470);
 * RS_POSITIVE_DOMAIN (2): a trial state with a negative entry is rejected and
 *   the step retried at half its size, values untouched (the out-of-domain
 *   rejection the reference wants for the Rosenbrock23 trial sub-steps,
 *   HELPERS.md:396-447, as an isoutofdomain rule with failfactor 2; per cell
 *   in rs_chem_percell, any entry of the operator's species in the
 *   global-norm solves).  In rs_split_opts it applies to the chemistry
 *   operator only: the transport's linear flux-form scheme undershoots at a
 *   sharp front by an amount proportional to the step, so halving never
 *   clears it there.  The recorded step sequences and the adjoint see only
 *   accepted steps. */
#define RS_CLAMP_NONNEG 1
#define RS_POSITIVE_DOMAIN 2

/* Options of one Lie-Trotter macro step (adjoint_gradient.jl:609-619 and the
 * tolerances/controllers of run_reseact_reactant.jl:144,310-355). */
typedef struct rs_split_opts {
    rs_pictrl ctrl_T, ctrl_C;
    double atol_T, rtol_T, atol_C, rtol_C;
    int clamp_nonneg;        /* flags: RS_CLAMP_NONNEG | RS_POSITIVE_DOMAIN    */
    int chem_percell;        /* 0: one global error norm (reference), 1: per cell */
    long long maxiters;
} rs_split_opts;

/* Per-cell (work-stealing) chemistry statistics. */
typedef struct rs_percell_stats {
    long long attempts_total, accepts_total, rejects_total;
    long long attempts_max, attempts_min;   /* over cells                       */
    long long n_failed;                     /* cells that hit maxiters          */
} rs_percell_stats;

/* Statistics of a forward + backward (adjoint) run. */
typedef struct rs_adjoint_stats {
    long long naccept_T, naccept_C, nreject_T, nreject_C;
    long long nvjp_T, nvjp_C;
    double seconds_forward, seconds_backward, seconds_replay;
    double replay_maxrel;    /* replayed end state vs checkpoint (adjoint_gradient.jl:802-808) */
    /* The first VJP of the backward sweep that made lambda or its gradient
     * non-finite (adjoint_gradient.jl:763-779: more non-finite entries in
     * lambda after it than before, or any in its gradient contribution):
     * bad_op -1 none, 0 transport, 1 chemistry; bad_macro the macro step
     * (0-based); bad_step the step (0-based, forward order) of bad_nsteps of
     * that operator in it, with its t and dt; non-finite counts of lambda
     * before and after it and of its gradient contribution. */
    int bad_op, bad_macro;
    long long bad_step, bad_nsteps;
    double bad_t, bad_dt;
    long long bad_lam_in, bad_lam_out, bad_grad;
} rs_adjoint_stats;

/* ---- lifecycle ------------------------------------------------------------ */
RS_API int         rs_create(rs_ctx** out, int device);
RS_API int         rs_destroy(rs_ctx* ctx);
RS_API const char* rs_last_error(const rs_ctx* ctx);
RS_API const char* rs_backend(void);                 /* "hip-gfx950" or "cpu-oracle" */

/* ---- setup ----------------------------------------------------------------
 * rs_set_grid: NLON x NLAT x NLEV cells, NS species of which the first NT are
 * transported.  Replaces the model build `(f, u0, p, tspan, var_map)` of
 * prototypes/reseact_3d_chem/split_common.jl:530-533 with explicit sizes.     */
RS_API int rs_set_grid(rs_ctx* ctx, int nlon, int nlat, int nlev, int ns, int nt);

/* Flux-form advection coefficients (units 1/s), periodic in longitude:
 *   cx[j]      zonal rate u_j / (a cos(phi_j) dlambda) of latitude row j;
 *   wz[k]      vertical face profile w(z_k)/dz at faces k = 0..NLEV
 *              (wz[0] = wz[NLEV] = 0: closed bottom and top);
 *   wl[i]      longitude factor; vertical rate at face (i,j,k) = wz[k]*wl[i].
 * Replaces the transport RHS f_T of reseact.esm:1564-3100 (the synthetic
 * 3rd-order upwind operator of the workloads). */
RS_API int rs_set_wind(rs_ctx* ctx, const double* cx, const double* wz, const double* wl);

/* ---- monotone PPM transport in CWC mixing-ratio form ------------------------
 * The reference's own transport (reseact.esm:1536-3100; the regional
 * prototype prototypes/transport_3d/transport_3d.esm): flux-form PPM with the
 * CW84 monotone slopes and parabola limiter, open lateral walls with
 * prescribed inflow mixing ratios, and the hybrid sigma-pressure mass-form
 * vertical with no flux through the surface and the lid.  The discretisation
 * rules are imported by the reference from EarthSciDiscretizations
 * (reseact.esm:61-110), which is not in its tree; the operator here restates
 * Colella & Woodward (1984) eqs. 1.6-1.10 in method-of-lines form (the face
 * value is the limited edge value of the upwind cell).
 *
 * With scheme RS_TRANSPORT_PPM the transported block is
 *   species 0        air mass m (Pa):  dm/dt = -div(M)
 *   species 1..NT-1  mixing ratios q:  dq/dt = (-div(M q_f) + q div(M)) / m
 * where div(F) = (Fx[i+1]-Fx[i]) dphi_lat[j] / (dlon dS_lat[j])
 *              + (coslat_e[j+1] Fy[j+1] - coslat_e[j] Fy[j]) / dS_lat[j]
 *              + (Fz[k+1] - Fz[k])   (k = 0 at the surface, +ve upward).
 * A uniform q stays uniform to the last bit under any air-mass flux (the
 * reference's CWC gate, transport_3d.esm:269).  rs_set_wind is not needed. */
#define RS_TRANSPORT_UPWIND3 0   /* the synthetic workloads' operator (default) */
#define RS_TRANSPORT_PPM     1
RS_API int rs_set_transport_scheme(rs_ctx* ctx, int scheme);

/* Grid metrics on the unit sphere: lon_periodic (1) or open west/east walls
 * (0); dlon (rad); coslat_e[NLAT+1] meridional face lengths (0 at a pole);
 * dS_lat[NLAT] cell area weights sin(phie[j+1]) - sin(phie[j]); dphi_lat[NLAT]
 * zonal face lengths (rad).  reseact.esm:146-177 (phie, coslat_e, dS_lat,
 * dphi_lat). */
RS_API int rs_set_ppm_geometry(rs_ctx* ctx, int lon_periodic, double dlon,
                               const double* coslat_e, const double* dS_lat,
                               const double* dphi_lat);

/* Face air-mass fluxes (Pa/s) and layer thicknesses (Pa), lon fastest:
 *   Mx[(k*NLAT + j)*(NLON+1) + i], i = 0..NLON (west face of cell i; with a
 *      periodic longitude Mx[..NLON] must equal Mx[..0]);
 *   My[(k*(NLAT+1) + j)*NLON + i], j = 0..NLAT (south face of row j);
 *   Mz[(k*NLAT + j)*NLON + i],     k = 0..NLEV (lower face of level k, +ve
 *      upward; zero at k = 0 and k = NLEV: no flux through surface and lid;
 *      NULL: zero everywhere until rs_diagnose_mz diagnoses it);
 *   dp[(k*NLAT + j)*NLON + i] > 0, the widths of the vertical parabolae
 *      (reseact.esm:178-197 Mx, My; :225-234 dp; :282-291 Mz). */
RS_API int rs_set_mass_fluxes(rs_ctx* ctx, const double* Mx, const double* My,
                              const double* Mz, const double* dp);

/* Inflow mixing ratios of the NT-1 mixing-ratio species in the 3-cell halos
 * of the open walls (qbc_w_X/qbc_e_X/qbc_s_X/qbc_n_X of reseact.esm:726-1121,
 * one set per species X): qbc_w[(s*NLEV + k)*NLAT + j], the same for qbc_e;
 * qbc_s[(s*NLEV + k)*NLON + i], the same for qbc_n; s = 0 is species 1.
 * NULL leaves that wall's halo at 0.  A periodic longitude ignores w/e. */
RS_API int rs_set_inflow(rs_ctx* ctx, const double* qbc_w, const double* qbc_e,
                         const double* qbc_s, const double* qbc_n);

/* Column pressure fixer of the air-mass equation (reseact.esm:5 caveat 4 and
 * the divh_fix variable, :262-281): with dPSdt[NLAT*NLON] (Pa/s, surface
 * pressure tendency) the air mass follows
 *   dm/dt = -div(M) + (dPSdt + divh_col) dp(k) / dp_col,
 * divh_col = sum_k of the horizontal divergence, dp_col = sum_k dp (sums in
 * level order), i.e. the continuity equation reads the corrected horizontal
 * divergence while the mixing ratios keep the raw div(M).  With Mz diagnosed
 * from the same corrected divergence, dm/dt = dB dPSdt up to round-off.
 * NULL switches it off.  Recomputed whenever the fluxes or the geometry are
 * set again; call after both. */
RS_API int rs_set_pressure_fixer(rs_ctx* ctx, const double* dPSdt);

/* Vertical air-mass flux diagnosed from continuity, replacing Mz = -OMEGA
 * (reseact.esm:262-291: divh, divh_col, dp_col, divh_fix, Mz), from the
 * horizontal fluxes and dp of rs_set_mass_fluxes and the geometry:
 *   divh_fix(k) = divh(k) + (-dPSdt - divh_col) dp(k) / dp_col,
 *   Mz[k] = -sum_{k' < k} (dB[k'] dPSdt + divh_fix(k'))   (exclusive scan),
 * sums in level order.  The scan reaches dPSdt (1 - sum dB) at the lid up to
 * round-off; the lid face is set to 0 and the largest |scan at the lid| over
 * the columns is returned in *lid_resid (may be NULL).  dB[NLEV] (the
 * hybrid B differences) and dPSdt[NLAT*NLON] (Pa/s, the pressure fixer's)
 * may be NULL (zeros).  The result becomes the context's Mz; Mz_out (may be
 * NULL) receives a copy in the layout of rs_set_mass_fluxes. */
RS_API int rs_diagnose_mz(rs_ctx* ctx, const double* dB, const double* dPSdt, double* Mz_out,
                          double* lid_resid);

/* PBL mixing of the mixing ratios (reseact.esm:1563-1690, the tracer
 * tendency's GEOS-Chem-style full PBL mixing; weights :7536 pbl_f, :7698
 * pbl_wt, :7747 pbl_sum_X, :8455 pbl_mean_X; tau_pblmix :1170): each mixing
 * ratio q gains
 *   pbl_f (pbl_mean - q) / tau_pblmix,
 *   pbl_mean = sum_k (dp pbl_f)_k q_k / max(sum_k (dp pbl_f)_k, 1)
 * per column (sums in level order).  pbl_f: NC layer fractions inside the PBL
 * (cell order), computed by the caller from its forcing (reseact.esm:7455
 * p_pbl_top, :7536); tau_pblmix > 0 in seconds.  The weights use the dp of
 * rs_set_mass_fluxes and follow it when the fluxes are set again.  NULL
 * switches it off.  Conserves sum_k dp q per column exactly in exact
 * arithmetic. */
RS_API int rs_set_pbl_mixing(rs_ctx* ctx, const double* pbl_f, double tau_pblmix);

/* Mass-action mechanism, NR reactions: reactants[2r], reactants[2r+1]
 * (second = -1 for unimolecular; equal entries = 2A), products[2r],
 * products[2r+1] (-1 = none), rate constants k[r] (1/s or 1/(conc s)),
 * photo[r] != 0 marks a photolysis reaction whose rate is k[r]*jfac[cell].
 * Replaces the chemistry RHS f_C and its block Jacobian
 * (rx_traced_integrator.jl:151-253, prototypes/reseact_3d_chem/block_jac.jl:57-136). */
RS_API int rs_set_mechanism(rs_ctx* ctx, int nr, const int* reactants, const int* products,
                     const double* k, const int* photo);
/* The same with general products, for mechanisms ingested from reaction
 * systems (fractional and multiple products, catalysts): reactants as above;
 * the products of reaction r are prod_species[q] with coefficient
 * prod_coef[q] (> 0) for q in [prod_off[r], prod_off[r+1]), prod_off[0] = 0.
 * Rate laws are mass action, k[r] times the product of the reactants (no
 * combinatoric factor for 2A).  Replaces the ReactionSystem the reference
 * builds from an .esm `reaction_systems` entry (SuperFast,
 * reseact.esm:12714-12717; prototypes/g2_depsink/g2_nodep.esm:450-1860). */
RS_API int rs_set_mechanism_stoich(rs_ctx* ctx, int nr, const int* reactants, const int* prod_off,
                                   const int* prod_species, const double* prod_coef,
                                   const double* k, const int* photo);
/* Per-cell photolysis factor (NC values, e.g. max(0, cos SZA)); NULL = 1. */
RS_API int rs_set_photolysis(rs_ctx* ctx, const double* jfac);
/* Surface emission source E[s*NLAT*NLON + j*NLON + i] (conc/s) added to the
 * chemistry tendency of level-0 cells (reseact.esm:12132); NULL = none. */
RS_API int rs_set_emissions(rs_ctx* ctx, const double* emis);

/* ---- state ---------------------------------------------------------------- */
RS_API int rs_set_state(rs_ctx* ctx, const double* u);   /* NS*NC, host -> device   */
RS_API int rs_get_state(rs_ctx* ctx, double* u);         /* NS*NC, device -> host   */
/* max(u, 0) on the whole state, in place: the macro-boundary clamp of the
 * Lie-Trotter driver (op_split.jl:151). */
RS_API int rs_clamp_state(rs_ctx* ctx);

/* ---- single attempts ------------------------------------------------------
 * One SSPRK43 / ROS23 attempt from the current state with step dt; the raw
 * result becomes the "candidate" and *eest its Hairer RMS error.
 * Reference: ssprk43_step_unrolled rx_traced_integrator.jl:404-437,
 *            ros23_step :279-349 (jac = :sym, an exact Jacobian).           */
RS_API int rs_ssprk43_step(rs_ctx* ctx, double t, double dt, double atol, double rtol, double* eest);
RS_API int rs_ros23_step(rs_ctx* ctx, double t, double dt, double atol, double rtol, double* eest);
RS_API int rs_get_candidate(rs_ctx* ctx, double* u);
RS_API int rs_accept_candidate(rs_ctx* ctx, int clamp_nonneg);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: one ROS23 attempt, as rs_ros23_step, on the first ncells
 * cells of the context's grid (with their photolysis factors and emissions),
 * from the caller's DEVICE buffer u_dev into the caller's DEVICE buffer
 * unew_dev (both NS x ncells doubles, species-major with stride ncells),
 * *eest its error norm.  The context's state is not touched.  Needs the grid
 * and the mechanism (set up beforehand); allocates nothing: the
 * per-workgroup error partials live in a fixed device array of the library,
 * RS_NOALLOC_PARTIALS doubles, one copy per GPU kept until the process ends,
 * so ncells is limited to RS_NOALLOC_PARTIALS times the chemistry workgroup
 * size (64 by default: 8,388,608 cells); larger is RS_E_ARG.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (other threads,
 * other contexts) race; each call synchronises its context's stream before
 * it returns. */
#define RS_NOALLOC_PARTIALS 131072
RS_API int rs_ros23_step_noalloc(rs_ctx* ctx, const double* u_dev, double* unew_dev,
                                 long long ncells, double dt, double atol, double rtol,
                                 double* eest);

/* ---- adaptive solves ------------------------------------------------------
 * The PI-controlled accept/reject loop over [t0, t1] with a single global
 * error norm.  *dt_io: initial dt in, proposed next dt out.
 * Reference: adaptive_solve rx_traced_integrator.jl:455-507 and its host twin
 * host_adaptive! tools/adjoint_gradient.jl:538-582.                         */
RS_API int rs_adaptive_transport(rs_ctx* ctx, double t0, double t1, double* dt_io,
                          const rs_pictrl* ctrl, double atol, double rtol,
                          int clamp_nonneg, long long maxiters, rs_counts* out);
RS_API int rs_adaptive_chem(rs_ctx* ctx, double t0, double t1, double* dt_io,
                     const rs_pictrl* ctrl, double atol, double rtol,
                     int clamp_nonneg, long long maxiters, rs_counts* out);

/* Accepted steps of the last rs_adaptive_transport (op 0) or rs_adaptive_chem
 * (op 1) call on this context: *n_out pairs (t, dt), t the time the step
 * started from, written to t_dt[2i], t_dt[2i+1] when t_dt is not NULL (cap =
 * pairs it can hold).  The sequence the reference's adaptive_solve takes
 * (rx_traced_integrator.jl:480-505) and its adjoint replays
 * (tools/adjoint_gradient.jl:647-661 replay_fixed). */
RS_API int rs_get_step_log(rs_ctx* ctx, int op, double* t_dt, long long cap, long long* n_out);

/* Per-cell adaptive chemistry: every cell integrates [t0, t1] as its own ODE
 * (the reference loop applied with NC = 1), all cells in one persistent
 * work-stealing launch.  dt_cells (NC values, may be NULL = dt0 everywhere)
 * carries each cell's proposed dt in and out. */
RS_API int rs_chem_percell(rs_ctx* ctx, double t0, double t1, double dt0, double* dt_cells,
                    const rs_pictrl* ctrl, double atol, double rtol,
                    int clamp_nonneg, long long maxiters, rs_percell_stats* out);

/* VJP of the per-cell adaptive chemistry over [t0, t1] from the current state
 * (unchanged), every cell starting at dt0: the reference's backward sweep
 * (tools/adjoint_gradient.jl:731-834) over each cell's own accepted step
 * sequence (adaptive_solve with NC = 1, rx_traced_integrator.jl:455-507).
 * Each cell is integrated again with its accepted steps recorded, then swept
 * last step first: lambda .*= 1[u_{j+1} > 0] when clamping (the clamp of the
 * accepted state, adjoint_gradient.jl:736), lambda <- VJP of ros23_step j at
 * u_j.  The lambda of rs_set_lambda is dJ/du(t1) on entry and dJ/du(t0) on
 * return; rs_get_grad_emis accumulates dJ/dE (rs_zero_grad clears it).
 * Device memory: (NS + 1) doubles per resident lane and recorded state, the
 * longest cell's accepted steps (found by a forward pass on scratch). */
RS_API int rs_chem_percell_vjp(rs_ctx* ctx, double t0, double t1, double dt0,
                               const rs_pictrl* ctrl, double atol, double rtol,
                               int clamp_nonneg, long long maxiters, rs_percell_stats* out);

/* One Lie-Trotter macro step: transport over [t0,t1] then chemistry over the
 * same interval from the transported state.
 * Reference: macro_step tools/adjoint_gradient.jl:609-619;
 *            lie_trotter_solve tools/reactant_handoff/op_split.jl:100-164. */
RS_API int rs_macro_step(rs_ctx* ctx, double t0, double t1, double* dtT_io, double* dtC_io,
                  const rs_split_opts* opts, rs_counts* cT, rs_counts* cC);

/* ---- discrete adjoint of one step (fixed dt) ------------------------------
 * lambda <- (d unew / d u)^T lambda at the current state; for ROS23 also
 * grad_emis += (d unew / d E)^T lambda.  The Rosenbrock stages are recomputed
 * from the state, not stored.  With the PPM transport (nonlinear: limiters,
 * division by the air mass) the SSPRK43 stage states are recomputed from the
 * state too, and the derivative is that of the limiter branches the forward
 * takes there (the derivative almost everywhere; DIFFERENTIABILITY_PLAN.md:
 * 709-736 for the caveat at exactly uniform fields).
 * Reference: ssprk43_step_vjp / ros23_step_vjp rx_traced_integrator.jl:610-623. */
RS_API int rs_set_lambda(rs_ctx* ctx, const double* lam);      /* NS*NC               */
RS_API int rs_get_lambda(rs_ctx* ctx, double* lam);
RS_API int rs_zero_grad(rs_ctx* ctx);
RS_API int rs_get_grad_emis(rs_ctx* ctx, double* g);           /* NS*NLAT*NLON        */
RS_API int rs_ssprk43_vjp(rs_ctx* ctx, double t, double dt);
RS_API int rs_ros23_vjp(rs_ctx* ctx, double t, double dt);

/* PPM transport: keep != 0 makes rs_ssprk43_step also keep its stage states
 * and the tendencies at them (7 transported blocks of device memory, already
 * held as scratch by the step VJP), so that a following rs_ssprk43_vjp of
 * the same step -- same state, same dt, with nothing but lambda/gradient
 * accessors or reads in between -- skips recomputing them: the primal values
 * a pullback holds on to (the reference's Enzyme ssprk43_step_vjp takes them
 * from its forward trace, rx_traced_integrator.jl:618-623).  The step costs
 * the extra writes; the VJP then does its four adjoint applications only.
 * Results are bit for bit those of the recomputing VJP.  Default 0. */
RS_API int rs_keep_transport_stages(rs_ctx* ctx, int keep);

/* ---- time-loop adjoint ----------------------------------------------------
 * Forward pass over macro stops tstops[0..nmacro] with a checkpoint of the
 * state and the accepted (t, dt) sequences at every macro step, then the
 * backward sweep: per macro step (last first) a fixed-sequence replay from
 * its checkpoint, chemistry VJPs (last step first), then transport VJPs, with
 * the clamp mask lambda .*= 1[u_raw > 0] before every step.
 * Loss RS_LOSS_SURFACE_SQ (0): J = sum over level-0 cells of
 * u[loss_species]^2.  RS_LOSS_WEIGHTS (1): J = <w, u(t_end)> with the weights
 * of rs_set_objective_weights (loss_species unused).
 * Outputs: *J_out, grad_u0 (NS*NC), grad_emis (NS*NLAT*NLON); either
 * gradient pointer may be NULL.
 * Errors, after the outputs and stats are written: RS_E_STATE when the
 * fixed-sequence replay of a macro step lands >= 1e-6 (relative, or
 * non-finite) from the next checkpoint (adjoint_gradient.jl:802-808);
 * RS_E_NONFINITE when a VJP made lambda or its gradient non-finite, with the
 * first such VJP in stats->bad_* and in rs_last_error.
 * With opts->chem_percell = 1 the chemistry of every macro step is the
 * per-cell solve of rs_chem_percell (each cell from dtC0); its backward is
 * that of rs_chem_percell_vjp, which replays each cell from the transported
 * state (the replay check compares the cells' end states), and it is
 * reported as one chemistry step spanning the macro step (bad_step 0 of 1,
 * t = the macro start, dt = its length); stats->nvjp_C counts the per-cell
 * step VJPs; rs_get_recorded_steps(op 1) returns no steps for it.
 * Reference: forward_pass / replay_fixed / backward_stage! / backward_sweep,
 * tools/adjoint_gradient.jl:647-834.                                        */
#define RS_LOSS_SURFACE_SQ 0
#define RS_LOSS_WEIGHTS    1
RS_API int rs_adjoint_run(rs_ctx* ctx, int nmacro, const double* tstops,
                   double dtT0, double dtC0, const rs_split_opts* opts,
                   int loss_kind, int loss_species,
                   double* J_out, double* grad_u0, double* grad_emis,
                   rs_adjoint_stats* stats);

/* Accepted steps of operator op (0 transport, 1 chemistry) in macro step
 * `macro_step` (0-based) of the last rs_adjoint_run: *n_out pairs (t, dt)
 * into t_dt when not NULL (cap pairs).  The recorded sequence the backward
 * sweep replays (Ckpt seqT / seqC, tools/adjoint_gradient.jl:621-629). */
RS_API int rs_get_recorded_steps(rs_ctx* ctx, int macro_step, int op, double* t_dt,
                                 long long cap, long long* n_out);

/* Weights w (NS*NC, species-major) of the linear objective J = <w, u(t_end)>
 * (loss kind RS_LOSS_WEIGHTS); NULL clears them.  Replaces the reference's
 * objective as a weight vector, objective_weights(var_map, spec),
 * tools/adjoint_gradient.jl:404-426 (the mean of a species over all cells,
 * the surface or one level: w = 1/count on the matched states). */
RS_API int rs_set_objective_weights(rs_ctx* ctx, const double* w);

/* Frozen-sequence forward map used by the finite-difference checks of the
 * time-loop adjoint (adjoint_gradient.jl:1183-1244 "fdtape"): re-runs the
 * accepted step sequences recorded by the last rs_adjoint_run from the
 * current state with the current emissions, and returns the loss. */
RS_API int rs_replay_loss(rs_ctx* ctx, int loss_kind, int loss_species, double* J_out);


/* ---- extensions with no counterpart in the reference ----------------------
 * Elimination order of the per-cell sparse LU, set before rs_set_mechanism:
 * 0 = natural species order, the elimination order of the reference's
 *     blocksolve (rounding differs only by FMA contraction and reciprocal
 *     multiplication);
 * 1 = Markowitz fill-reducing order (default; fewer flops). */
RS_API int rs_set_lu_ordering(rs_ctx* ctx, int ordering);

/* Transport kernels for each SSPRK43 attempt and its VJP:
 * 1 = one fused launch per attempt (all four stages in LDS; default, used
 *     where the grid allows: nlon >= 4 and a longitude segment with an
 *     8-column halo fits in LDS at this nlev), with 1024-lane workgroups
 *     where a whole periodic row fits one, else 512-lane workgroups
 *     (narrower segments, two workgroups per CU);
 * 2 = fused, 1024-lane workgroups;  3 = fused, 512-lane workgroups;
 * 0 = one launch per stage.  All give the same bits.  rs_autotune times
 * the per-stage tiles and both fused sizes and keeps the fastest. */
RS_API int rs_set_transport_kernel(rs_ctx* ctx, int mode);

/* Static work of the compiled mechanism kernels (roofline accounting). */
typedef struct rs_mech_info {
    int ns, nr, nnz_J, nnz_LU, ordering;
    long long flops_attempt;     /* FP64 flops of one ROS23 attempt on one cell  */
    long long flops_vjp;         /* FP64 flops of one ROS23 step VJP on one cell */
    int regs_attempt, regs_percell, regs_vjp;   /* VGPRs per lane (0 on CPU)   */
    long long scratch_attempt, scratch_percell, scratch_vjp;  /* spill bytes/lane */
} rs_mech_info;
RS_API int rs_mechanism_info(rs_ctx* ctx, rs_mech_info* out);

/* Per-kernel timing with HIP events recorded on the library's stream around
 * every launch while enabled (rs_profile(ctx, 1)); rs_profile_read returns
 * launches and total milliseconds per kernel kind. */
typedef struct rs_kernel_time {
    char name[32];
    long long launches;
    double total_ms;
} rs_kernel_time;
RS_API int rs_profile(rs_ctx* ctx, int enable);

/* Autotuner: times one SSPRK43 attempt for every transport tile height
 * (4, 8, 16, 32 levels per 64-longitude workgroup) and for the fused
 * kernel, one ROS23 attempt for every chemistry workgroup size (64, 128,
 * 256) and, at the fastest size, with the rate constants passed as a kernel
 * argument or through a pointer, on this context's own grid and state, and
 * keeps the fastest of each.
 * The state is not changed. */
typedef struct rs_tune_result {
    int transport_tile_levels;
    int chem_block;
    double transport_ms[4];      /* per candidate, 0 if not timed */
    double chem_ms[3];
    int transport_fused;         /* 1 if the fused kernel was kept */
    int transport_fused_threads; /* its workgroup size (1024 or 512), 0 if not kept */
    double transport_fused_ms;   /* 0 if the grid does not allow it */
    int chem_kpointer;           /* 1: attempt reads rate constants through a pointer */
    int pad2;
    double chem_kp_ms;           /* attempt time of the pointer variant at chem_block */
} rs_tune_result;
RS_API int rs_autotune(rs_ctx* ctx, rs_tune_result* out);

/* Ensembles: N independent contexts (regional domains of different sizes,
 * each fully set up) advanced together.  Forward: every adaptive phase of
 * every member runs from one instantiated HIP graph (one branch per member
 * stream) launched until all controllers are done.  Backward: members are
 * grouped so their replay tapes fit `tape_budget_gib`, and each group's whole
 * backward sweep (replays + VJPs over all split steps) is one graph launch.
 * The groups' tapes are slices of one device pool, sized once for the
 * largest group and kept by the ensemble.
 * Members may use either transport scheme (upwind or PPM).
 * Each member's gradients stay in its context (rs_get_lambda = dJ/du0,
 * rs_get_grad_emis).  No counterpart in the reference, which runs one domain
 * per process. */
typedef struct rs_ensemble rs_ensemble;
typedef struct rs_ensemble_stats {
    double seconds_forward, seconds_backward;
    long long graph_launches_forward, graph_launches_backward;
    long long graph_nodes_backward;
    int backward_groups;
    int pad;
    double seconds_backward_setup;   /* tape pool, capture and instantiation */
    double seconds_backward_run;     /* graph launches until done */
} rs_ensemble_stats;
RS_API int rs_ensemble_create(rs_ensemble** out, rs_ctx* const* members, int n);
RS_API int rs_ensemble_destroy(rs_ensemble* e);
RS_API const char* rs_ensemble_last_error(const rs_ensemble* e);
RS_API int rs_ensemble_autotune(rs_ensemble* e, rs_tune_result* out);
RS_API int rs_ensemble_adjoint_run(rs_ensemble* e, int nmacro, const double* tstops,
                                   double dtT0, double dtC0, const rs_split_opts* opts,
                                   int loss_kind, int loss_species, double tape_budget_gib,
                                   double* J_out, rs_adjoint_stats* stats,
                                   rs_ensemble_stats* es);
RS_API int rs_profile_read(rs_ctx* ctx, rs_kernel_time* out, int max, int* n_out);

#ifdef __cplusplus
}
#endif
#endif /* RESEACT_HIP_H */
