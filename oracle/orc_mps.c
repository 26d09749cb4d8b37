/*
 * orc_mps.c -- ORACLE (test infrastructure only): restatement of the
 * reference's standalone 3-D melt SCF, MPS.cpp, for an A(N1)-B(N2) diblock
 * melt on a periodic simple-cubic box.
 *
 * Follows MPS.cpp:18-284 step by step:
 *   halo copies bx/by/bz, b_x/b_y/b_z   MPS.cpp:18-74, 138-148
 *   Side (6-neighbour average)         MPS.cpp:150-156
 *   Propagate / ComputePhi             MPS.cpp:158-185
 *   ComputeG (residual)                MPS.cpp:187-198
 *   Ax (DIIS coefficients), DIIS       MPS.cpp:200-251
 *   Helmholtz, SCF                     MPS.cpp:253-284
 * Two constants are parameters instead of the reference's integer divisions:
 * n (MPS.cpp:301, Mx*My*Mz/N in int arithmetic) and the bulk fractions in
 * PutAlpha (MPS.cpp:193-194, N2/N and N1/N in int arithmetic, i.e. 0).
 * Ax solves A X = 1 for the symmetric DIIS matrix by a Jacobi eigen-
 * decomposition (X = Q diag(1/l) Q^T 1), which equals the reference's SVD
 * pseudo-inverse V S^-1 U^T 1 for a non-singular A.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

struct orc_mps {
  int Mx, My, Mz, N1, N2, N, n_seg;
  long M, jx, jy, iv;
  double chi, n, phib_a, phib_b, GN;
  int bx1, by1, bz1, bxm, bym, bzm;
  double *phi, *phitot, *G1, *alpha, *Gg_f, *Gg_b, *phi_side, *g;
};

static void halo_x(orc_mps* p, double* P) {
  long jx = p->jx, jy = p->jy;
  for (int y = 0; y < p->My + 2; y++)
    for (int z = 0; z < p->Mz + 2; z++) {
      long i = jy * y + z;
      P[i] = P[p->bx1 * jx + i];
      P[jx * (p->Mx + 1) + i] = P[jx * p->bxm + i];
    }
}
static void halo_y(orc_mps* p, double* P) {
  long jx = p->jx, jy = p->jy;
  for (int x = 0; x < p->Mx + 2; x++)
    for (int z = 0; z < p->Mz + 2; z++) {
      long i = jx * x + z;
      P[i] = P[jy * p->by1 + i];
      P[jy * (p->My + 1) + i] = P[jy * p->bym + i];
    }
}
static void halo_z(orc_mps* p, double* P) {
  long jx = p->jx, jy = p->jy;
  for (int x = 0; x < p->Mx + 2; x++)
    for (int y = 0; y < p->My + 2; y++) {
      long i = jx * x + jy * y;
      P[i] = P[i + p->bz1];
      P[i + p->Mz + 1] = P[i + p->bzm];
    }
}
static void set_boundaries(orc_mps* p, double* P) {
  halo_x(p, P);
  halo_y(p, P);
  halo_z(p, P);
}
static void remove_boundaries(orc_mps* p, double* P) {
  long jx = p->jx, jy = p->jy;
  for (int y = 0; y < p->My + 2; y++)
    for (int z = 0; z < p->Mz + 2; z++) P[jy * y + z] = P[jx * (p->Mx + 1) + jy * y + z] = 0;
  for (int x = 0; x < p->Mx + 2; x++)
    for (int z = 0; z < p->Mz + 2; z++) P[jx * x + z] = P[jx * x + jy * (p->My + 1) + z] = 0;
  for (int x = 0; x < p->Mx + 2; x++)
    for (int y = 0; y < p->My + 2; y++) P[jx * x + jy * y] = P[jx * x + jy * y + p->Mz + 1] = 0;
}

/* Side: X_side = (sum of the six shifted copies of X)/6 over the whole array,
 * shifted copies clipped at the array ends (MPS.cpp:150-156). */
static void side(orc_mps* p, double* Xs, double* X) {
  long M = p->M, jx = p->jx, jy = p->jy;
  set_boundaries(p, X);
#pragma omp parallel for
  for (long i = 0; i < M; i++) {
    double s = 0;
    if (i >= jx) s += X[i - jx];
    if (i + jx < M) s += X[i + jx];
    if (i >= jy) s += X[i - jy];
    if (i + jy < M) s += X[i + jy];
    if (i >= 1) s += X[i - 1];
    if (i + 1 < M) s += X[i + 1];
    Xs[i] = s * (1.0 / 6.0);
  }
}

static void propagate(orc_mps* p, double* G, const double* g1, int s_from, int s_to) {
  double* gs = G + p->M * (s_to - 1);
  double* gs_1 = G + p->M * (s_from - 1);
  side(p, gs, gs_1);
#pragma omp parallel for
  for (long i = 0; i < p->M; i++) gs[i] = gs[i] * g1[i];
}

orc_mps* orc_mps_new(int mx, int my, int mz, int n1, int n2, double chi, double n_chains,
                     double phib_a, double phib_b) {
  orc_mps* p = calloc(1, sizeof(orc_mps));
  p->Mx = mx;
  p->My = my;
  p->Mz = mz;
  p->N1 = n1;
  p->N2 = n2;
  p->N = n1 + n2;
  p->n_seg = 2;
  p->chi = chi;
  p->n = n_chains;
  p->phib_a = phib_a;
  p->phib_b = phib_b;
  p->jx = (long)(my + 2) * (mz + 2);
  p->jy = mz + 2;
  p->M = (long)(mx + 2) * (my + 2) * (mz + 2);
  p->bx1 = mx;
  p->bxm = 1;
  p->by1 = my;
  p->bym = 1;
  p->bz1 = mz;
  p->bzm = 1; /* periodic, MPS.cpp:308 */
  p->iv = 2 * p->M;
  p->phi = calloc(p->iv, sizeof(double));
  p->phitot = calloc(p->M, sizeof(double));
  p->G1 = calloc(p->iv, sizeof(double));
  p->alpha = calloc(p->M, sizeof(double));
  p->Gg_f = calloc(p->M * p->N, sizeof(double));
  p->Gg_b = calloc(p->M * 2, sizeof(double));
  p->phi_side = calloc(p->iv, sizeof(double));
  p->g = calloc(p->iv, sizeof(double));
  return p;
}

void orc_mps_free(orc_mps* p) {
  if (!p) return;
  free(p->phi);
  free(p->phitot);
  free(p->G1);
  free(p->alpha);
  free(p->Gg_f);
  free(p->Gg_b);
  free(p->phi_side);
  free(p->g);
  free(p);
}

long orc_mps_M(const orc_mps* p) { return p->M; }
double orc_mps_gn(const orc_mps* p) { return p->GN; }
void orc_mps_get_phi(const orc_mps* p, double* phi) { memcpy(phi, p->phi, p->iv * sizeof(double)); }

/* MPS.cpp:164-185 */
static void compute_phi(orc_mps* p, const double* u) {
  long M = p->M;
  int N = p->N, N1 = p->N1;
  for (int i = 0; i < p->n_seg; i++) {
    for (long k = 0; k < M; k++) p->G1[i * M + k] = exp(-u[i * M + k]);
    set_boundaries(p, p->G1 + i * M);
  }
  memcpy(p->Gg_f, p->G1, M * sizeof(double));
  for (int s = 2; s <= N1; s++) propagate(p, p->Gg_f, p->G1, s - 1, s);
  for (int s = N1 + 1; s <= N; s++) propagate(p, p->Gg_f, p->G1 + M, s - 1, s);
  memcpy(p->Gg_b + (N % 2) * M, p->G1 + M, M * sizeof(double));
  for (long k = 0; k < M; k++) p->phi[M + k] = p->Gg_f[(long)(N - 1) * M + k] * p->Gg_b[(N % 2) * M + k];
  for (int s = N - 1; s > N1; s--) {
    propagate(p, p->Gg_b, p->G1 + M, ((s + 1) % 2) + 1, (s % 2) + 1);
    for (long k = 0; k < M; k++) p->phi[M + k] += p->Gg_f[(long)(s - 1) * M + k] * p->Gg_b[(s % 2) * M + k];
  }
  for (long k = 0; k < M; k++) p->phi[k] = 0;
  for (int s = N1; s >= 1; s--) {
    propagate(p, p->Gg_b, p->G1, ((s + 1) % 2) + 1, (s % 2) + 1);
    for (long k = 0; k < M; k++) p->phi[k] += p->Gg_f[(long)(s - 1) * M + k] * p->Gg_b[(s % 2) * M + k];
  }
  double* gN = p->Gg_f + (long)(N - 1) * M;
  remove_boundaries(p, gN);
  double GN = 0;
  for (long k = 0; k < M; k++) GN += gN[k];
  p->GN = GN;
  for (int i = 0; i < p->n_seg; i++)
    for (long k = 0; k < M; k++) p->phi[i * M + k] *= (p->n / GN) / p->G1[i * M + k];
}

/* MPS.cpp:187-198 */
void orc_mps_compute_g(orc_mps* p, const double* x, double* g) {
  long M = p->M;
  compute_phi(p, x);
  for (long k = 0; k < M; k++) p->phitot[k] = p->phi[k] + p->phi[M + k];
  for (int i = 0; i < p->n_seg; i++) side(p, p->phi_side + i * M, p->phi + i * M);
  memcpy(g, x, p->iv * sizeof(double));
  for (long k = 0; k < M; k++) g[k] = g[k] - p->chi * (p->phi_side[M + k] / p->phitot[k] - p->phib_b);
  for (long k = 0; k < M; k++) g[M + k] = g[M + k] - p->chi * (p->phi_side[k] / p->phitot[k] - p->phib_a);
  for (long k = 0; k < M; k++) p->alpha[k] = (g[k] + g[M + k]) * (1.0 / p->n_seg);
  for (int i = 0; i < p->n_seg; i++) {
    for (long k = 0; k < M; k++) g[i * M + k] = g[i * M + k] - p->alpha[k] + 1 / p->phitot[k] - 1;
    remove_boundaries(p, g + i * M);
  }
}

/* MPS.cpp:253-261. Uses phi and phi_side of the last compute_g; x is u. */
double orc_mps_helmholtz(orc_mps* p, const double* x) {
  long M = p->M;
  remove_boundaries(p, p->phi);
  remove_boundaries(p, p->phi + M);
  double uint_ = 0, ent = 0;
  for (long k = 0; k < M; k++) uint_ += p->phi_side[k] * p->phi[M + k];
  uint_ *= p->chi;
  for (long k = 0; k < M; k++) ent += p->phi[M + k] * x[M + k];
  ent -= p->n * log(p->n * p->N / p->GN);
  double e2 = 0;
  for (long k = 0; k < M; k++) e2 += p->phi[k] * x[k];
  ent += e2;
  return uint_ - ent;
}

/* Jacobi eigen-decomposition of a symmetric n x n matrix (row-major). */
static void jacobi_eig(double* A, int n, double* Q, double* w) {
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Q[i * n + j] = i == j;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p * n + q];
        if (apq == 0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double th = (aqq - app) / (2 * apq);
        double t = (th >= 0 ? 1 : -1) / (fabs(th) + sqrt(th * th + 1));
        double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double qkp = Q[k * n + p], qkq = Q[k * n + q];
          Q[k * n + p] = c * qkp - s * qkq;
          Q[k * n + q] = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = A[i * n + i];
}

void orc_diis_coeffs(const double* Ain, int n, double* X) {
  double A[64 * 64], Q[64 * 64], w[64];
  memcpy(A, Ain, sizeof(double) * n * n);
  jacobi_eig(A, n, Q, w);
  for (int i = 0; i < n; i++) X[i] = 0;
  for (int k = 0; k < n; k++) {
    double qt1 = 0;
    for (int i = 0; i < n; i++) qt1 += Q[i * n + k];
    double c = qt1 / w[k];
    for (int i = 0; i < n; i++) X[i] += Q[i * n + k] * c;
  }
}

static double dot(const double* a, const double* b, long n) {
  double r = 0;
  for (long i = 0; i < n; i++) r += a[i] * b[i];
  return r;
}

/* MPS.cpp:228-284: DIIS with memory m and step eta from x (in/out). */
int orc_mps_scf(orc_mps* p, double* x, double tol, int max_iter, int m, double eta,
                int* iterations, double* final_error) {
  long iv = p->iv;
  double* x0 = calloc(iv, sizeof(double));
  double* xR = calloc(iv * m, sizeof(double));
  double* x_x0 = calloc(iv * m, sizeof(double));
  double* Aij = calloc(m * m, sizeof(double));
  double* Apij = calloc(m * m, sizeof(double));
  double* Ci = calloc(m, sizeof(double));
  double* g = calloc(iv, sizeof(double));
  int it = 0, k_diis = 1, k = 0;
  /* MPS.cpp:263 starts from x = x0 = 0; from another start x0 is that start,
   * so x_x0[0] is the first step in both cases */
  memcpy(x0, x, iv * sizeof(double));
  orc_mps_compute_g(p, x, g);
  for (long i = 0; i < iv; i++) x[i] += -eta * g[i];
  for (long i = 0; i < iv; i++) x_x0[i] = x[i] - x0[i];
  memcpy(xR, x, iv * sizeof(double));
  double error = sqrt(dot(g, g, iv));
  while (error > tol && it < max_iter) {
    it++;
    memcpy(x0, x, iv * sizeof(double));
    orc_mps_compute_g(p, x, g);
    k = it % m;
    k_diis++;
    for (long i = 0; i < iv; i++) x[i] += -eta * g[i];
    memcpy(xR + k * iv, x, iv * sizeof(double));
    for (long i = 0; i < iv; i++) x_x0[k * iv + i] = x[i] - x0[i];
    /* DIIS(), MPS.cpp:228-251 */
    if (k_diis > m) {
      k_diis = m;
      for (int i = 1; i < m; i++)
        for (int j = 1; j < m; j++) Aij[m * (i - 1) + j - 1] = Aij[m * i + j];
    }
    for (int i = 0; i < k_diis; i++) {
      int posi = k - k_diis + 1 + i;
      if (posi < 0) posi += m;
      Aij[i + m * (k_diis - 1)] = Aij[k_diis - 1 + m * i] = dot(x_x0 + posi * iv, x_x0 + k * iv, iv);
    }
    for (int i = 0; i < k_diis; i++)
      for (int j = 0; j < k_diis; j++) Apij[j + k_diis * i] = Aij[j + m * i];
    orc_diis_coeffs(Apij, k_diis, Ci);
    double normC = 0;
    for (int i = 0; i < k_diis; i++) normC += Ci[i];
    for (int i = 0; i < k_diis; i++) Ci[i] = Ci[i] / normC;
    memset(x, 0, iv * sizeof(double));
    for (int i = 0; i < k_diis; i++) {
      int posi = k - k_diis + 1 + i;
      if (posi < 0) posi += m;
      for (long q = 0; q < iv; q++) x[q] += Ci[i] * xR[posi * iv + q];
    }
    error = sqrt(dot(g, g, iv));
  }
  if (iterations) *iterations = it;
  if (final_error) *final_error = error;
  free(x0);
  free(xR);
  free(x_x0);
  free(Aij);
  free(Apij);
  free(Ci);
  free(g);
  return error <= tol ? 0 : SFX_E_NOTCONVERGED;
}
