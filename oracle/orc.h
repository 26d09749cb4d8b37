/*
 * orc.h -- ORACLE (test infrastructure only). FP64 CPU restatement of the
 * reference's Scheutjens-Fleer hot path. Loaded by tests/, smoke() (as the
 * checker) and bench.py's cpu_baseline leg only; never by the product path.
 */
#ifndef ORC_H
#define ORC_H

#include "../include/sfx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_lat {
  int grad, geom, type, order;
  int n[3];
  int MX, MY, MZ;
  long M, jx, jy;
  int code1[3], code2[3]; /* boundary choice per gradient (SFX_BOUND_*) */
  int b1[3], b2[3];       /* halo source coordinate per gradient (0-based) */
  double l1, l1chi, ladj, offset, dimensions;
  int adjust_outer;
  double l0_2d, l1x, l1y;
  double *L, *lm, *lp, *lmc, *lpc; /* per-x-layer tables of curved lattices */
} orc_lat;

orc_lat* orc_lat_new(const sfx_lattice_desc* d);
void orc_lat_free(orc_lat* L);
long orc_lat_M(const orc_lat* L);
int orc_lat_is_interior(const orc_lat* L, long i);
void orc_setb(const orc_lat* L, double* P);
void orc_zero_halos(const orc_lat* L, double* P);
void orc_propagate(const orc_lat* L, double* Gin, const double* g, double* Gout);
void orc_subtract_boundaries(const orc_lat* L, double* P);
void orc_multiply_sites(const orc_lat* L, double* P);
double orc_lngn(const orc_lat* L, const double* G);
void orc_side(const orc_lat* L, double* P, double* out);
/* second order (3-D simple cubic), six direction components per site */
void orc_setb6(const orc_lat* L, double* P);
int orc_ncomp2(const orc_lat* L);
double orc_wcomp2(const orc_lat* L, int k);
void orc_propagate2(const orc_lat* L, double* Gin, const double* g, double* Gout, double S, int backward);
double orc_lngn2(const orc_lat* L, const double* G);
void orc_connect2(const orc_lat* L, const double* A, const double* B, double* out, double* OUT);
void orc_correct2(const orc_lat* L, double* in, double* IN, const double* G);
/* force ensemble and normalisations */
void orc_propagate_force(const orc_lat* L, double* Gin, const double* g, double* Gout, double f);
void orc_norm_phi_free(const orc_lat* L, double* phi, double* PHI, double C);
void orc_norm_phi_restr(const orc_lat* L, double* phi, double* PHI, const double* Gi, double C);
/* tests only: 1 makes the 2-D cylinder's backward sweep split its outward
 * radial step as the forward sweep does (the reference does not) */
void orc_set_cyl2d_split_outward(int on);

/* ---- SF system ---- */
typedef struct orc_sys orc_sys;
orc_sys* orc_sys_new(const sfx_lattice_desc* ld, const sfx_system_desc* sd);
void orc_sys_free(orc_sys* s);
long orc_sys_nvar(const orc_sys* s);
long orc_sys_M(const orc_sys* s);
int orc_sys_n_itvar(const orc_sys* s);
/* UpdateSWF + ComputePhi + GetResiduals; returns 0 or a negative code. */
int orc_sys_residual(orc_sys* s, const double* x, double* f);
int orc_sys_compute_phi(orc_sys* s, const double* x);
void orc_sys_get_phi(const orc_sys* s, int state, double* out);
void orc_sys_get_swf(const orc_sys* s, int state, double* out);
double orc_sys_lngn(const orc_sys* s, int mol);
double orc_sys_phibulk(const orc_sys* s, int mol);
double orc_sys_free_energy(orc_sys* s);
double orc_sys_chemical_potential(const orc_sys* s, int mol);
double orc_sys_grand_potential(orc_sys* s);
/* Homogeneous initial guess (x = 0 everywhere). */
int orc_sys_solve_diis(orc_sys* s, double* x, double tol, int max_iter, int m, double eta,
                       int* iterations, double* final_norm);

/* ---- MPS.cpp literal restatement (3-D SC periodic AB melt) ---- */
typedef struct orc_mps orc_mps;
orc_mps* orc_mps_new(int mx, int my, int mz, int n1, int n2, double chi, double n_chains,
                     double phib_a, double phib_b);
void orc_mps_free(orc_mps* p);
long orc_mps_M(const orc_mps* p);
/* ComputeG at u = x (x: 2*M, halos zero); g out: 2*M. */
void orc_mps_compute_g(orc_mps* p, const double* x, double* g);
void orc_mps_get_phi(const orc_mps* p, double* phi /* 2*M */);
double orc_mps_gn(const orc_mps* p);
double orc_mps_helmholtz(orc_mps* p, const double* x);
int orc_mps_scf(orc_mps* p, double* x, double tol, int max_iter, int m, double eta,
                int* iterations, double* final_error);

/* small dense symmetric solve used by DIIS: X = A^{-1} 1 (MPS.cpp:200-226) */
void orc_diis_coeffs(const double* A, int n, double* X);

/* SFNewton's pseudo-Hessian Newton with the Scheutjens line search
 * (orc_newton.c). The factorisation routines work on the reference's float
 * n x n L D U array (element (r, c) at h[r + n c]). */
typedef void (*orc_resid_fn)(void* ctx, const double* x, double* g);
typedef struct orc_newton_opts {
  double tolerance;
  int iterationlimit;
  double deltamax, deltamin;
  int linesearchlimit;
  double reset_hessian_criterion;
  double small_alpha;
  int max_n_small_alpha;
} orc_newton_opts;
typedef struct orc_newton_stats {
  int iterations, functioncalls, hessian_resets, reversed_directions, time_dependent, converged;
  double accuracy, residual_norm;
} orc_newton_stats;
void orc_newton_defaults(orc_newton_opts* o);
int orc_newton(orc_resid_fn f, void* ctx, double* x, int n, const orc_newton_opts* o, orc_newton_stats* st);
int orc_newton_sys(orc_sys* s, double* x, int n, const orc_newton_opts* o, orc_newton_stats* st);
void orc_ldu_multiply(const float* h, int n, const double* w, double alpha, double* v);
int orc_ldu_sign(const float* h, int n);
void orc_ldu_update(float* h, int n, double* w, double* v, double alpha);
void orc_ldu_solve(const float* h, int n, const double* g, double* p);

#ifdef __cplusplus
}
#endif
#endif
