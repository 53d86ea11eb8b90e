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
int rs_oracle_threads(void);
#ifdef __cplusplus
}
#endif
#endif
