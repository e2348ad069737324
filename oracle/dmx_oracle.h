/*
 * dmx_oracle.h -- CPU FP64 restatement of the reference MHD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg);
 * the product (libdmx.so) never links, loads or calls it.
 *
 * It restates marcjoos-cea/dumses-hybrid (MPI=0, NDIM=3, Cartesian) array by
 * array, in the reference's operation order, so that it reproduces the
 * reference's float64 results bit for bit (checked against the compiled
 * reference by tests/test_oracle.py).  Arrays use the reference's index
 * ranges; uin is exchanged in the reference layout (see include/dmx.h).
 */
#ifndef DMX_ORACLE_H
#define DMX_ORACLE_H

#include "../include/dmx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmxo dmxo;

dmxo*  dmxo_create(const dmx_params* p);
void   dmxo_destroy(dmxo* s);
void   dmxo_load(dmxo* s, const double* uin);    /* reference-layout uin */
void   dmxo_store(const dmxo* s, double* uin);
double dmxo_compute_dt(dmxo* s);                 /* physics.f90:18-132   */
void   dmxo_godunov(dmxo* s, double time, double dt); /* scheme.f90:104  */
void   dmxo_fargo(dmxo* s, double dt);           /* fargo.f90:17-124     */
void   dmxo_boundary(dmxo* s, double time, double dt); /* boundary.f90:18 */
void   dmxo_dissipation(dmxo* s, double time, double dt); /* physics.f90:272 */
/* nsteps of the reference loop (dumses.f90:80-120); returns the new time */
double dmxo_advance(dmxo* s, int nsteps, double time, double* dts);
void   dmxo_history(dmxo* s, double out[9]);     /* history.f90:62-118   */
/* shear = 1: stresses of src/problem/mri/history.f90 (-<BxBy>, <rho dvx dvy>) */
void   dmxo_history_variant(dmxo* s, int shear, double out[9]);

/* Function-level entry points of the Riemann solvers.
 * 1D: q = [rho, p, u_n, B_n, u_t1, B_t1, u_t2, B_t2] left and right;
 *     flux[8]; star[7] = rho, u, v, w, B_t1, B_t2, p_T of the sampled state
 *     (zeros for the solvers other than HLLD).
 * 2D: q{LL,RL,LR,RR} = [rho, p, u, v, A, B, C, E] per corner; returns emf*dt */
void   dmxo_riemann1d(int solver, const double* ql, const double* qr,
                      double gamma, double ciso, int iso,
                      double* flux, double* star);
double dmxo_riemann2d(int solver, const double* qLL, const double* qRL,
                      const double* qLR, const double* qRR, double dt,
                      double gamma, double ciso, int iso);

#ifdef __cplusplus
}
#endif
#endif
