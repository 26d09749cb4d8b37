/* orc_newton.c -- the reference's default SCF iteration, restated in C for
 * tests: SFNewton's Newton method with a pseudo-Hessian (a float L D U
 * factorisation updated by one rank-one term per step), the Scheutjens line
 * search and trust region, and the reset rules SF_Solve applies around each
 * direction. Test infrastructure only: the library never links this file.
 *
 * Follows (Scheutjens branch; DIIS, L-BFGS, CG and TN off; pseudohessian on):
 *   sfnewton.cpp:844-1143    iterate
 *   sfnewton.cpp:1240-1256   resethessian, :1523-1530 startderivatives
 *   sfnewton.cpp:1312-1330   initialize_iteration
 *   sfnewton.cpp:1433-1521   newhessian (pseudo-Hessian branch)
 *   sfnewton.cpp:1578-1580   newfunction
 *   sfnewton.cpp:1916-1940   direction (Newton branch)
 *   sfnewton.cpp:1946-2038   zero, linesearch, linecriterion, stepchange,
 *                            newtrustregion
 *   sfnewton.cpp:2040-2056   residue (the reported accuracy), newdirection
 *   newttool.cpp:3-10,12-32,49-65,67-75,170-210,212-243
 *                            innerproduct, multiply, norm2, signdeterminant,
 *                            updatpos, gausa, gausb
 *   SF_Solve.cpp:659-760     inneriteration (reset rules)
 * Defaults: sfnewton.cpp:11-37 (linesearchlimit 20, max accuracy for Hessian
 * scaling 0.1) and the values SF_Solve reads for a `newton` block with the
 * default method and line search (SF_Solve.cpp:1836-1963, 2168-2176):
 * deltamax 0.5, deltamin 0, reset_hessian_criterion 1e3, small_alpha 1e-5,
 * max_n_small_alpha 50, reverse_direction_range 50, min_accuracy_for_hessian
 * 0 (the full-Hessian branches never fire).
 *
 * The factorisation is stored as the reference stores it: one float n x n
 * array h, element (r, c) at h[r + n c]; the diagonal holds D, the strict
 * lower part the unit-lower L, the strict upper part the unit-upper U, and
 * H = L D U.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define HE(h, n, r, c) (h)[(size_t)(r) + (size_t)(n) * (size_t)(c)]

static double dot(const double* a, const double* b, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static double nrm(const double* a, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * a[i];
  return sqrt(s);
}

/* v = alpha L D U w  (newttool.cpp:12-32) */
void orc_ldu_multiply(const float* h, int n, const double* w, double alpha, double* v) {
  double* t = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = i + 1; j < n; j++) s += w[j] * HE(h, n, i, j);
    t[i] = (s + w[i]) * HE(h, n, i, i);
    s = 0;
    for (int j = 0; j < i; j++) s += t[j] * HE(h, n, i, j);
    v[i] = alpha * (s + t[i]);
  }
  free(t);
}

/* +1 or -1: the sign of det D (newttool.cpp:67-75) */
int orc_ldu_sign(const float* h, int n) {
  int s = 1;
  for (int i = 0; i < n; i++)
    if (HE(h, n, i, i) < 0) s = -s;
  return s;
}

/* L D U <- L D U + alpha w v^T in place; w and v are overwritten
 * (newttool.cpp:170-210) */
void orc_ldu_update(float* h, int n, double* w, double* v, double alpha) {
  for (int i = 0; i < n; i++) {
    const double vi = v[i];
    double d = HE(h, n, i, i);
    double b = d + (alpha * w[i]) * vi;
    HE(h, n, i, i) = (float)b;
    /* the reference divides by the double b, not by the stored float */
    d /= b;
    const double c = vi * alpha / b;
    b = w[i] * alpha / b;
    alpha *= d;
    for (int j = i + 1; j < n; j++) {
      const double wj = w[j];
      w[j] -= w[i] * HE(h, n, j, i);
      HE(h, n, j, i) *= d;
      HE(h, n, j, i) += c * wj;
      const double vj = v[j];
      v[j] -= vi * HE(h, n, i, j);
      HE(h, n, i, j) *= d;
      HE(h, n, i, j) += b * vj;
    }
  }
}

/* p = -(L D U)^-1 g: forward substitution with L (gausa, newttool.cpp:212-232)
 * then D and U (gausb, newttool.cpp:234-243) */
void orc_ldu_solve(const float* h, int n, const double* g, double* p) {
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < i; j++) s += HE(h, n, i, j) * p[j];
    p[i] = -g[i] - s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = 0;
    for (int j = i + 1; j < n; j++) s += HE(h, n, i, j) * p[j];
    p[i] = p[i] / HE(h, n, i, i) - s;
  }
}

enum { REVERSE_RANGE = 50 };

typedef struct {
  orc_resid_fn f;
  void* ctx;
  int n;
  double *x, *x0, *g, *g0, *p, *p0, *y, *hp;
  float* h;
  double alpha, trustregion, trustfactor, deltamax, deltamin, accuracy, minimum;
  int iterations, lineiterations, linesearchlimit, resetiteration, trouble, newtondirection, samehessian;
  /* SF_Solve's reset state */
  double min_accuracy_so_far, reset_criterion, small_alpha;
  int small_alpha_count, max_small_alpha;
  unsigned char reverse[REVERSE_RANGE];
  orc_newton_stats* st;
} Newton;

static void newgradient(Newton* N) {
  N->f(N->ctx, N->x, N->g);
  N->st->functioncalls++;
}

/* startderivatives / resethessian: H = (1 + |g|) I, in float */
static void resethessian(Newton* N) {
  const int n = N->n;
  N->trouble = 0;
  const float diag = (float)(1 + nrm(N->g, n));
  memset(N->h, 0, sizeof(float) * (size_t)n * (size_t)n);
  for (int i = 0; i < n; i++) HE(N->h, n, i, i) = diag;
  N->resetiteration = N->iterations;
  N->st->hessian_resets++;
}

static void setdeltamax(Newton* N, double d) {
  if (d > N->deltamin) N->deltamax = d;
}

static void newhessian(Newton* N) {
  const int n = N->n;
  const double dmin = 1.0 / pow(2.0, 53);
  if (N->samehessian) return;
  if (N->alpha == 0 || N->iterations == 0) {
    resethessian(N);
    return;
  }
  double py = 0;
  for (int i = 0; i < n; i++) {
    const double dg = N->g[i] - N->g0[i];
    N->y[i] = dg;
    py += N->p[i] * dg;
  }
  if (!N->newtondirection)
    orc_ldu_multiply(N->h, n, N->p, 1, N->hp);
  else
    for (int i = 0; i < n; i++) N->hp[i] = -N->g0[i];
  double php = dot(N->p, N->hp, n);
  double theta = py / (10 * dmin + N->alpha * php);
  if (theta > 0 && N->iterations == N->resetiteration + 1 && N->accuracy > 0.1) {
    /* Hessian scaling on the first step after a reset */
    N->alpha *= theta;
    py /= theta;
    php /= theta;
    for (int i = 0; i < n; i++) {
      N->p[i] /= theta;
      HE(N->h, n, i, i) *= theta;
    }
  }
  N->trustfactor *= (4 / (pow(theta - 1, 2) + 1) + 0.5);
  if (n > 1) {
    const double np = nrm(N->p, n);
    double sum = N->alpha * (np * np);
    theta = fabs(py / (N->alpha * php));
    if (theta < .01)
      sum /= 0.8;
    else if (theta > 100)
      sum *= theta / 50;
    for (int i = 0; i < n; i++) N->y[i] -= N->alpha * N->hp[i];
    orc_ldu_update(N->h, n, N->y, N->p, 1.0 / sum);
    N->trouble -= orc_ldu_sign(N->h, n);
    if (N->trouble < 0)
      N->trouble = 0;
    else if (N->trouble >= 3)
      resethessian(N);
  }
}

/* SF_Solve::inneriteration, the rules it applies before each direction */
static void inneriteration(Newton* N) {
  const double acc = N->accuracy;
  if (N->iterations > 0) N->samehessian = 0;
  if (acc < N->min_accuracy_so_far && N->iterations > 0 && acc == fabs(acc)) N->min_accuracy_so_far = acc;
  if (acc > N->min_accuracy_so_far * N->reset_criterion && acc == fabs(acc)) {
    resethessian(N);
    N->min_accuracy_so_far *= 1.5;
    if (N->deltamax > 0.005) setdeltamax(N, 0.9 * N->deltamax);
    memset(N->reverse, 0, sizeof N->reverse);
  }
  if (N->alpha < N->small_alpha)
    N->small_alpha_count++;
  else
    N->small_alpha_count = 0;
  if (N->small_alpha_count == N->max_small_alpha) {
    N->small_alpha_count = 0;
    memset(N->reverse, 0, sizeof N->reverse);
    resethessian(N);
    if (N->deltamax > 0.005) setdeltamax(N, 0.9 * N->deltamax);
  }
  N->reverse[N->iterations % REVERSE_RANGE] = !N->newtondirection;
}

static void direction(Newton* N) {
  const int n = N->n;
  if (N->iterations == 0 && !N->samehessian) newhessian(N);
  orc_ldu_solve(N->h, n, N->g, N->p);
  N->newtondirection = orc_ldu_sign(N->h, n) > 0;
  if (!N->newtondirection) {
    for (int i = 0; i < n; i++) N->p[i] = -N->p[i];
    N->st->reversed_directions++;
  }
}

static void newdirection(Newton* N) {
  const int n = N->n;
  inneriteration(N);
  memcpy(N->p0, N->p, sizeof(double) * (size_t)n);
  direction(N);
  N->accuracy = sqrt(nrm(N->p, n) * nrm(N->g, n) / (1 + nrm(N->x, n)));
}

/* x = x0 + a p and its residual; the fifth call of a line search first
 * re-evaluates x0 to detect a residual that changes between calls */
static int zero(Newton* N, double a) {
  const int n = N->n;
  int valid = 1;
  N->lineiterations++;
  if (N->lineiterations == 5) {
    int timedep = 0;
    memcpy(N->x, N->x0, sizeof(double) * (size_t)n);
    newgradient(N);
    for (int i = 0; i < n && valid && !timedep; i++) {
      if (N->g[i] != N->g0[i]) {
        timedep = 1;
        N->st->time_dependent = 1;
      } else if (!isfinite(N->g[i])) {
        valid = 0;
        a = -a;
      }
    }
  }
  for (int i = 0; i < n; i++) N->x[i] = N->x0[i] + a * N->p[i];
  valid = 1;
  newgradient(N);
  for (int i = 0; i < n && valid; i++)
    if (!isfinite(N->g[i])) {
      valid = 0;
      N->g[i] = 1;
    }
  const double ng = nrm(N->g, n);
  N->minimum = ng * ng;
  return valid;
}

static double linecriterion(Newton* N) {
  const int n = N->n;
  double normg = nrm(N->g0, n);
  const double gg0 = dot(N->g, N->g0, n) / normg / normg;
  normg = pow(nrm(N->g, n) / normg, 2);
  if ((gg0 > 1 || normg > 1) && normg - gg0 * fabs(gg0) < 0.2) normg = 1.5 * normg;
  if (gg0 < 0 && normg < 2) return 1;
  if (normg > 10) return .01;
  return 0.4 * (1 + 0.75 * normg) / (normg - gg0 * fabs(gg0) + 0.1);
}

static double stepchange(Newton* N) {
  double crit = linecriterion(N), change = crit;
  while (crit < 0.35 && N->lineiterations < N->linesearchlimit) {
    N->alpha /= 4;
    zero(N, N->alpha);
    crit = linecriterion(N);
    change = 1;
  }
  return change;
}

static void newtrustregion(Newton* N) {
  const double np0 = nrm(N->p0, N->n);
  if (np0 > 0 && N->trustregion > 2 * N->alpha * np0) N->trustregion = 2 * N->alpha * np0;
  N->trustregion *= N->trustfactor;
  N->trustfactor = 1.0;
  if (N->trustregion > N->deltamax) N->trustregion = N->deltamax;
  if (N->trustregion < N->deltamin) N->trustregion = N->deltamin;
}

void orc_newton_defaults(orc_newton_opts* o) {
  o->tolerance = 1e-7;
  o->iterationlimit = 1000;
  o->deltamax = 0.5;
  o->deltamin = 0;
  o->linesearchlimit = 20;
  o->reset_hessian_criterion = 1e3;
  o->small_alpha = 1e-5;
  o->max_n_small_alpha = 50;
}

int orc_newton(orc_resid_fn f, void* ctx, double* x, int n, const orc_newton_opts* o, orc_newton_stats* st) {
  if (n < 2) return -1;
  memset(st, 0, sizeof *st);
  Newton N;
  memset(&N, 0, sizeof N);
  N.f = f;
  N.ctx = ctx;
  N.n = n;
  N.st = st;
  N.x = x;
  double* buf = (double*)calloc((size_t)n * 7, sizeof(double));
  N.x0 = buf;
  N.g = buf + n;
  N.g0 = buf + 2 * (size_t)n;
  N.p = buf + 3 * (size_t)n;
  N.p0 = buf + 4 * (size_t)n;
  N.y = buf + 5 * (size_t)n;
  N.hp = buf + 6 * (size_t)n;
  N.h = (float*)calloc((size_t)n * (size_t)n, sizeof(float));
  N.deltamax = o->deltamax;
  N.deltamin = o->deltamin > 0 && o->deltamin <= o->deltamax ? o->deltamin : 0;
  N.linesearchlimit = o->linesearchlimit;
  N.reset_criterion = o->reset_hessian_criterion;
  N.small_alpha = o->small_alpha;
  N.max_small_alpha = o->max_n_small_alpha;
  N.min_accuracy_so_far = 1e30;
  /* initialize_iteration */
  N.alpha = 1;
  N.trustregion = N.deltamax;
  N.trustfactor = 1;
  const double tol = o->tolerance;

  newgradient(&N);
  {
    const double ng = nrm(N.g, n);
    N.minimum = ng * ng;
  }
  newdirection(&N);
  double normg = sqrt(N.minimum);
  while ((tol < N.accuracy || tol * 10 < normg) && N.iterations < o->iterationlimit &&
         N.accuracy == fabs(N.accuracy)) {
    N.iterations++;
    N.lineiterations = 0;
    newtrustregion(&N);
    const double alphabound = N.trustregion / (nrm(N.p, n) + 1 / pow(2.0, 53));
    memcpy(N.x0, N.x, sizeof(double) * (size_t)n);
    memcpy(N.g0, N.g, sizeof(double) * (size_t)n);
    N.alpha = alphabound < 1 ? alphabound : 1;
    zero(&N, N.alpha);
    N.trustfactor *= stepchange(&N);
    N.trustfactor *= N.alpha / alphabound;
    newhessian(&N);
    newdirection(&N);
    normg = sqrt(N.minimum);
  }
  st->iterations = N.iterations;
  st->accuracy = N.accuracy;
  st->residual_norm = normg;
  st->converged = !(tol < N.accuracy || tol * 10 < normg) && N.accuracy == fabs(N.accuracy);
  free(buf);
  free(N.h);
  return 0;
}

static void sys_resid(void* ctx, const double* x, double* g) { orc_sys_residual((orc_sys*)ctx, x, g); }

int orc_newton_sys(orc_sys* s, double* x, int n, const orc_newton_opts* o, orc_newton_stats* st) {
  return orc_newton(sys_resid, s, x, n, o, st);
}
