/* oracle_extra.h -- oracle-only entry points of libreseact_cpu.so (test
 * infrastructure: tangents and operator-level probes used by tests/ as the
 * checker).  The HIP library does not export these. */
#ifndef RESEACT_ORACLE_EXTRA_H
#define RESEACT_ORACLE_EXTRA_H
#include <stddef.h>
#include "../include/reseact_hip.h"
#ifdef __cplusplus
extern "C" {
#endif
int rs_oracle_ros23_jvp(rs_ctx* c, double dt, const double* v, const double* Ev, double* out);
int rs_oracle_ssprk43_jvp(rs_ctx* c, double dt, const double* v, double* out);
int rs_oracle_transport_rhs(rs_ctx* c, const double* q, double* f);
int rs_oracle_transport_rhs_adj(rs_ctx* c, const double* lam, double* g);
/* PPM only: J(u) v at the context's state u */
int rs_oracle_transport_rhs_jvp(rs_ctx* c, const double* v, double* out);
int rs_oracle_chem_rhs(rs_ctx* c, const double* u, double* f);
int rs_oracle_chem_jac(rs_ctx* c, const double* u, size_t cell, double* J);
/* Every attempt of the last rs_adaptive_transport / rs_adaptive_chem call:
 * rec[4i..4i+3] = (t, dtc, EEst, outcome 1 accept / 0 reject / -1 out of
 * domain); *n_out attempts; rec may be NULL to query the count. */
int rs_oracle_attempt_log(rs_ctx* c, double* rec, long long cap, long long* n_out);
/* The time-loop adjoint (rs_adjoint_run) with the forward pass taking the
 * given accepted sequences instead of running the controllers: macro step m
 * has nT[m] transport and nC[m] chemistry steps, (t, dt) pairs consecutive in
 * t_dt_T / t_dt_C.  For comparing with a run that took those sequences. */
int rs_oracle_adjoint_on_sequence(rs_ctx* c, int nmacro, const double* tstops,
                                  const long long* nT, const long long* nC, const double* t_dt_T,
                                  const double* t_dt_C, const rs_split_opts* o, int loss_kind,
                                  int loss_species, double* J_out, double* grad_u0,
                                  double* grad_emis, rs_adjoint_stats* st);
int rs_oracle_threads(void);
#ifdef __cplusplus
}
#endif
#endif
