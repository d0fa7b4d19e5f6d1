/*
 * senga_oracle.h -- CPU restatement of the SENGA+ RK-substage right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: tests/, the smoke
 * check and bench.py's cpu_baseline leg call it as the checker.  The product
 * library (include/senga_gpu.h) never links or calls it.
 *
 * It follows the reference Fortran statement by statement (each function
 * cites its reference source lines) in the same floating-point evaluation order, so
 * that built with gcc -O2 and no FMA it reproduces the reference binary's
 * results; it is pinned against the reference build (oracle/_ref) and the
 * committed golden fixtures by tests/test_oracle_*.py.
 */
#ifndef SENGA_ORACLE_H
#define SENGA_ORACLE_H

#include "senga_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

orc_t *orc_create(const sg_config *cfg, int *err);
void orc_destroy(orc_t *o);
/* run(nx,ny,nz,5+N), Fortran order; also sets the stage register and runs
 * the initial bisection (TEMPIN). */
int orc_set_state(orc_t *o, const double *run, double etime);
void orc_set_timestep(orc_t *o, double dt);
/* one RHS evaluation at substage irk; rhs (5+N)*n interior, Fortran order.
 * The stage register is restored afterwards; T, p, interval index and the
 * chemical rate keep the values of this evaluation. */
int orc_rhs(orc_t *o, int irk, double *rhs);
int orc_substage(orc_t *o, int irk);
int orc_end_step(orc_t *o);
/* Adaptive time step (ADAPTT, src/adaptt.f) after each step when on; errold
 * and errldr are the controller history (the dump's ERROLD, ERRLDR). */
int orc_set_adaptive(orc_t *o, int on, double errold, double errldr);
int orc_get_timestep(orc_t *o, double *tstep, double *errold, double *errldr);
int orc_step(orc_t *o, int nsteps);
int orc_get(orc_t *o, double *run, double *temp, double *pres, double *rate);
/* stage register (DRHS..YRHS interior) */
int orc_get_stage(orc_t *o, double *stage);
int orc_derivatives(orc_t *o, const double *f, double *out9);
/* interval indices of species 1-15 from the last TEMPIN/TEMPER: the first
 * ITNDEX word (interior, packed as src/com_senga2.h:88-98) */
int orc_get_itndex(orc_t *o, int *itndex);
int orc_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
