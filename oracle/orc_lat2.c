/*
 * orc_lat2.c -- ORACLE (test infrastructure only): restatement of the
 * second-order (semiflexible) lattices: Lat3D2ndO.cpp (six components),
 * Lat2D2ndO.cpp and Lat2DCyl2ndO.cpp (five), Lat1D2ndO.cpp, Lat1DCyl2ndO.cpp
 * and Lat1DSphere2ndO.cpp (three), each in its own lumped form.
 *
 * A propagator has six direction components k (k*M + site): the bond into the
 * site points along e_k with e = {+x, +y, +z, -z, -y, -x}, so component 5-k is
 * the reversal of k. Transfer weights T_kk = S, T_k,5-k = 0, others
 * L = (1-S)/4 (Lat3D2ndO.cpp:70, :108-145).
 *   PropagateF  G^k_s(r) = g(r) sum_j T_kj G^j_{s-1}(r - e_k)   (:64-152)
 *   PropagateB  G^k(r)   = g(r) sum_j T_jk G^j(r + e_j)          (:597-678)
 *   halos: a mirror bound swaps the +/- components of its axis, a periodic
 *   bound copies plainly (SetBx1(P,Q) etc., Lat3D.cpp:931-972, :1276-1290)
 *   ConnectG    phi += sum_k A^k B^k                            (:895-955)
 *   ComputeLnGN ln(sum_k sum_interior G^k / 6)                  (:996-1025)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static int swap_axis(int k, int axis) {
  /* +x(0) <-> -x(5), +y(1) <-> -y(4), +z(2) <-> -z(3) */
  int a = (k == 0 || k == 5) ? 0 : (k == 1 || k == 4) ? 1 : 2;
  return a == axis ? 5 - k : k;
}

/* halo fill of all six components with mirror swaps */
void orc_setb6(const orc_lat* L, double* P) {
  long M = L->M;
  for (int x = 0; x < L->MX; x++)
    for (int y = 0; y < L->MY; y++)
      for (int z = 0; z < L->MZ; z++) {
        int c[3] = {x, y, z}, s[3] = {x, y, z};
        int halo = 0;
        int mir[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++) {
          if (c[a] == 0) {
            s[a] = L->b1[a];
            halo = 1;
            mir[a] = L->code1[a] != SFX_BOUND_PERIODIC;
          } else if (c[a] == L->n[a] + 1) {
            s[a] = L->b2[a];
            halo = 1;
            mir[a] = L->code2[a] != SFX_BOUND_PERIODIC;
          }
        }
        if (!halo) continue;
        long i = ((long)x * L->MY + y) * L->MZ + z;
        long src = ((long)s[0] * L->MY + s[1]) * L->MZ + s[2];
        double v[6];
        for (int k = 0; k < 6; k++) {
          int kk = k;
          for (int a = 0; a < 3; a++)
            if (mir[a]) kk = swap_axis(kk, a);
          v[k] = P[kk * M + src];
        }
        for (int k = 0; k < 6; k++) P[k * M + i] = v[k];
      }
}

static double T(int k, int j, double S) {
  if (k == j) return S;
  if (j == 5 - k) return 0.0;
  return (1.0 - S) / 4.0;
}

/* ---- 1-D and 2-D forms ----------------------------------------------------
 * Lat1D2ndO: components {+x, in-layer (each of the four), -x}.
 * Lat2D2ndO: components {+x, +y, perpendicular (each of the two), -y, -x}.
 * Halo rules of the (P,Q) setters: the lower bound swaps the pair for surface
 * and mirror1 and copies plainly otherwise (`bound1 > 2`, Lat1DFlat.cpp:658-666,
 * Lat2DFlat.cpp:1343-1355, 1382-1393); the upper bound copies plainly only
 * when periodic (`bound2 < numLayers-1`, Lat1DFlat.cpp:673-682,
 * Lat2DFlat.cpp:1363-1375, 1401-1410). Copies are sequential, as there. */
static int lower_swaps(const orc_lat* L, int a) {
  return L->code1[a] == SFX_BOUND_SURFACE || L->code1[a] == SFX_BOUND_MIRROR1;
}
static int upper_swaps(const orc_lat* L, int a) { return L->code2[a] != SFX_BOUND_PERIODIC; }

/* x-halo rows (all y) from row b; swap pair (p, q) or plain */
static void rows(const orc_lat* L, double* p, double* q, int dst, int src, int swap) {
  long jx = L->MY;
  for (int y = 0; y < L->MY; y++) {
    if (swap && q) {
      p[dst * jx + y] = q[src * jx + y];
      q[dst * jx + y] = p[src * jx + y];
    } else {
      p[dst * jx + y] = p[src * jx + y];
      if (q) q[dst * jx + y] = q[src * jx + y];
    }
  }
}
static void cols(const orc_lat* L, double* p, double* q, int dst, int src, int swap) {
  long jx = L->MY;
  for (int x = 0; x < L->MX; x++) {
    if (swap && q) {
      p[x * jx + dst] = q[x * jx + src];
      q[x * jx + dst] = p[x * jx + src];
    } else {
      p[x * jx + dst] = p[x * jx + src];
      if (q) q[x * jx + dst] = q[x * jx + src];
    }
  }
}

static void setb3(const orc_lat* L, double* P) {
  long M = L->M;
  double *p0 = P, *p1 = P + M, *p2 = P + 2 * M;
  int top = L->MX - 1;
  /* SetBx1(gz0,gz2); SetBx1(gz1); SetBxm(gz0,gz2); SetBxm(gz1) */
  rows(L, p0, p2, 0, L->b1[0], lower_swaps(L, 0));
  rows(L, p1, NULL, 0, L->b1[0], 0);
  rows(L, p0, p2, top, L->b2[0], upper_swaps(L, 0));
  rows(L, p1, NULL, top, L->b2[0], 0);
}

static void setb5(const orc_lat* L, double* P) {
  long M = L->M;
  double* p[5];
  for (int k = 0; k < 5; k++) p[k] = P + k * M;
  int tx = L->MX - 1, ty = L->MY - 1;
  /* Lat2D2ndO.cpp:47-50: x lower, y lower, x upper, y upper */
  rows(L, p[0], p[4], 0, L->b1[0], lower_swaps(L, 0));
  for (int k = 1; k <= 3; k++) rows(L, p[k], NULL, 0, L->b1[0], 0);
  cols(L, p[1], p[3], 0, L->b1[1], lower_swaps(L, 1));
  cols(L, p[0], NULL, 0, L->b1[1], 0);
  cols(L, p[2], NULL, 0, L->b1[1], 0);
  cols(L, p[4], NULL, 0, L->b1[1], 0);
  rows(L, p[0], p[4], tx, L->b2[0], upper_swaps(L, 0));
  for (int k = 1; k <= 3; k++) rows(L, p[k], NULL, tx, L->b2[0], 0);
  cols(L, p[1], p[3], ty, L->b2[1], upper_swaps(L, 1));
  cols(L, p[0], NULL, ty, L->b2[1], 0);
  cols(L, p[2], NULL, ty, L->b2[1], 0);
  cols(L, p[4], NULL, ty, L->b2[1], 0);
}

/* Lat1D2ndO.cpp:32-62 (forward) and :117-145 (backward) */
static void propagate_1d(const orc_lat* L, double* Gin, const double* g, double* Gout, double S, int backward) {
  long M = L->M;
  double Lw = (1.0 - S) / 4.0;
  setb3(L, Gin);
  const double *a0 = Gin, *a1 = Gin + M, *a2 = Gin + 2 * M;
  for (long z = 1; z <= L->n[0]; z++) {
    double o0, o1, o2;
    if (!backward) {
      o0 = S * a0[z - 1] + 4 * Lw * a1[z - 1];
      o1 = Lw * a0[z] + (2 * Lw + S) * a1[z] + Lw * a2[z];
      o2 = 4 * Lw * a1[z + 1] + S * a2[z + 1];
    } else {
      o0 = S * a0[z + 1] + 4 * Lw * a1[z];
      o1 = Lw * a0[z + 1] + (2 * Lw + S) * a1[z] + Lw * a2[z - 1];
      o2 = S * a2[z - 1] + 4 * Lw * a1[z];
    }
    Gout[z] = o0 * g[z];
    Gout[M + z] = o1 * g[z];
    Gout[2 * M + z] = o2 * g[z];
  }
  setb3(L, Gout);
}

/* Lat2D2ndO.cpp:31-92 (forward) and :167-219 (backward) */
static void propagate_2d(const orc_lat* L, double* Gin, const double* g, double* Gout, double S, int backward) {
  long M = L->M, jx = L->MY;
  double Lw = (1.0 - S) / 4.0;
  setb5(L, Gin);
  const double* a[5];
  for (int k = 0; k < 5; k++) a[k] = Gin + k * M;
  for (int x = 1; x <= L->n[0]; x++)
    for (int y = 1; y <= L->n[1]; y++) {
      long i = x * jx + y;
      double o[5];
      if (!backward) {
        long p = i - jx, q = i - 1, r = i + 1, u = i + jx;
        o[0] = S * a[0][p] + Lw * a[1][p] + 2 * Lw * a[2][p] + Lw * a[3][p];
        o[1] = Lw * a[0][q] + S * a[1][q] + 2 * Lw * a[2][q] + Lw * a[4][q];
        o[2] = Lw * a[0][i] + Lw * a[1][i] + S * a[2][i] + Lw * a[3][i] + Lw * a[4][i];
        o[3] = Lw * a[0][r] + 2 * Lw * a[2][r] + S * a[3][r] + Lw * a[4][r];
        o[4] = Lw * a[1][u] + 2 * Lw * a[2][u] + Lw * a[3][u] + S * a[4][u];
      } else {
        o[0] = S * a[0][i + jx] + Lw * a[1][i + 1] + 2 * Lw * a[2][i] + Lw * a[3][i - 1];
        o[1] = Lw * a[0][i + jx] + S * a[1][i + 1] + 2 * Lw * a[2][i] + Lw * a[4][i - jx];
        o[2] = Lw * a[0][i + jx] + Lw * a[1][i + 1] + S * a[2][i] + Lw * a[3][i - 1] + Lw * a[4][i - jx];
        o[3] = Lw * a[0][i + jx] + 2 * Lw * a[2][i] + S * a[3][i - 1] + Lw * a[4][i - jx];
        o[4] = Lw * a[1][i + 1] + 2 * Lw * a[2][i] + Lw * a[3][i - 1] + S * a[4][i - jx];
      }
      for (int k = 0; k < 5; k++) Gout[k * M + i] = o[k] * g[i];
    }
  setb5(L, Gout);
}

/* ---- curved forms --------------------------------------------------------
 * Lat1DCyl2ndO.cpp and Lat1DSphere2ndO.cpp (three components; the sphere is
 * the same code over the sphere's tables) and Lat2DCyl2ndO.cpp (five). A
 * radial step is split by per-layer fractions l1(x) = 6 lambda_+1(x) and
 * l_1(x) = 6 lambda_-1(x) (Lat1DCyl2ndO.cpp:22-25, Lat2DCyl2ndO.cpp:19-23):
 * of a bond that would cross into the neighbouring layer the part l crosses
 * and the part 1 - l stays in the layer and continues with the reversed
 * radial component (LReflect/UReflect, Lat1DCyl2ndO.cpp:52-62,
 * Lat2DCyl2ndO.cpp:47-61). Halos as the flat forms. */

/* Lat1DCyl2ndO.cpp:64-94 (forward, PropagateF) and :149-181 (PropagateB) */
static void propagate_1d_curved(const orc_lat* L, double* Gin, const double* g, double* Gout, double S,
                                int backward) {
  long M = L->M;
  double Lw = (1.0 - S) / 4.0;
  setb3(L, Gin);
  const double *a0 = Gin, *a1 = Gin + M, *a2 = Gin + 2 * M;
  for (long z = 1; z <= L->n[0]; z++) {
    double up = 6 * L->lp[z], dn = 6 * L->lm[z]; /* l1[z], l_1[z] */
    double o0, o1, o2;
    if (!backward) {
      o0 = S * (dn * a0[z - 1] + (1 - dn) * a2[z]) + 4 * Lw * (dn * a1[z - 1] + (1 - dn) * a1[z]);
      o1 = Lw * a0[z] + (2 * Lw + S) * a1[z] + Lw * a2[z];
      o2 = 4 * Lw * (up * a1[z + 1] + (1 - up) * a1[z]) + S * (up * a2[z + 1] + (1 - up) * a0[z]);
    } else {
      double out = up * a0[z + 1] + (1 - up) * a2[z]; /* UReflect(H, gz0, 0) */
      double in = dn * a2[z - 1] + (1 - dn) * a0[z];  /* LReflect(H, gz2, 2) */
      o0 = S * out + 4 * Lw * a1[z];
      o1 = Lw * out + (2 * Lw + S) * a1[z] + Lw * in;
      o2 = S * in + 4 * Lw * a1[z];
    }
    Gout[z] = o0 * g[z];
    Gout[M + z] = o1 * g[z];
    Gout[2 * M + z] = o2 * g[z];
  }
  setb3(L, Gout);
}

/* Lat2DCyl2ndO.cpp:63-121 (forward) and :198-255 (backward). PropagateB
 * computes the outward split UReflect(H,gz0,0) (:223) but then reads the
 * plain outward neighbour gz0+jx (:224-227), so on backward sweeps only the
 * inward radial step is split; restated as it is. */
int orc_cyl2d_split_outward = 0; /* tests only: 1 applies the outward split on backward sweeps too */
void orc_set_cyl2d_split_outward(int on) { orc_cyl2d_split_outward = on; }

static void propagate_2d_cyl(const orc_lat* L, double* Gin, const double* g, double* Gout, double S,
                             int backward) {
  long M = L->M, jx = L->MY;
  double Lw = (1.0 - S) / 4.0;
  setb5(L, Gin);
  const double* a[5];
  for (int k = 0; k < 5; k++) a[k] = Gin + k * M;
  for (int x = 1; x <= L->n[0]; x++) {
    double up = 6 * L->lp[x], dn = 6 * L->lm[x];
    for (int y = 1; y <= L->n[1]; y++) {
      long i = x * jx + y;
      double o[5];
      if (!backward) {
        long p = i - jx, q = i - 1, r = i + 1, u = i + jx;
        o[0] = S * (dn * a[0][p] + (1 - dn) * a[4][i]) + Lw * (dn * a[1][p] + (1 - dn) * a[1][i]) +
               2 * Lw * (dn * a[2][p] + (1 - dn) * a[2][i]) + Lw * (dn * a[3][p] + (1 - dn) * a[3][i]);
        o[1] = Lw * a[0][q] + S * a[1][q] + 2 * Lw * a[2][q] + Lw * a[4][q];
        o[2] = Lw * a[0][i] + Lw * a[1][i] + S * a[2][i] + Lw * a[3][i] + Lw * a[4][i];
        o[3] = Lw * a[0][r] + 2 * Lw * a[2][r] + S * a[3][r] + Lw * a[4][r];
        o[4] = Lw * (up * a[1][u] + (1 - up) * a[1][i]) + 2 * Lw * (up * a[2][u] + (1 - up) * a[2][i]) +
               Lw * (up * a[3][u] + (1 - up) * a[3][i]) + S * (up * a[4][u] + (1 - up) * a[0][i]);
      } else {
        double h = dn * a[4][i - jx] + (1 - dn) * a[0][i]; /* LReflect(H, gz4, 4) */
        double w = a[0][i + jx];                           /* gz0+jx, not UReflect's H */
        if (orc_cyl2d_split_outward) w = up * a[0][i + jx] + (1 - up) * a[4][i];
        o[0] = S * w + Lw * a[1][i + 1] + 2 * Lw * a[2][i] + Lw * a[3][i - 1];
        o[1] = Lw * w + S * a[1][i + 1] + 2 * Lw * a[2][i] + Lw * h;
        o[2] = Lw * w + Lw * a[1][i + 1] + S * a[2][i] + Lw * a[3][i - 1] + Lw * h;
        o[3] = Lw * w + 2 * Lw * a[2][i] + S * a[3][i - 1] + Lw * h;
        o[4] = Lw * a[1][i + 1] + 2 * Lw * a[2][i] + Lw * a[3][i - 1] + S * h;
      }
      for (int k = 0; k < 5; k++) Gout[k * M + i] = o[k] * g[i];
    }
  }
  setb5(L, Gout);
}

int orc_ncomp2(const orc_lat* L) { return L->grad == 1 ? 3 : L->grad == 2 ? 5 : 6; }

/* multiplicity of a lumped component in ConnectG and ComputeLnGN
 * (Lat1D2ndO.cpp ConnectG: 1, 4, 1; Lat2D2ndO.cpp:270-282: 1, 1, 2, 1, 1) */
double orc_wcomp2(const orc_lat* L, int k) {
  if (L->grad == 1) return k == 1 ? 4.0 : 1.0;
  if (L->grad == 2) return k == 2 ? 2.0 : 1.0;
  return 1.0;
}

void orc_propagate2(const orc_lat* L, double* Gin, const double* g, double* Gout, double S, int backward) {
  if (L->grad == 1) {
    if (L->geom != SFX_FLAT)
      propagate_1d_curved(L, Gin, g, Gout, S, backward);
    else
      propagate_1d(L, Gin, g, Gout, S, backward);
    return;
  }
  if (L->grad == 2) {
    if (L->geom == SFX_CYLINDER)
      propagate_2d_cyl(L, Gin, g, Gout, S, backward);
    else
      propagate_2d(L, Gin, g, Gout, S, backward);
    return;
  }
  long M = L->M, jx = L->jx, jy = L->jy;
  long off[6] = {jx, jy, 1, -1, -jy, -jx}; /* e_k as a site offset */
  orc_setb6(L, Gin);
  for (int x = 1; x <= L->n[0]; x++)
    for (int y = 1; y <= L->n[1]; y++)
      for (int z = 1; z <= L->n[2]; z++) {
        long r = (x * (long)L->MY + y) * L->MZ + z;
        for (int k = 0; k < 6; k++) {
          double s = 0;
          for (int j = 0; j < 6; j++) {
            if (!backward)
              s += Gin[j * M + r - off[k]] * (g[r] * T(k, j, S));
            else
              s += Gin[j * M + r + off[j]] * (g[r] * T(j, k, S));
          }
          Gout[k * M + r] = s;
        }
      }
  orc_setb6(L, Gout);
}

double orc_lngn2(const orc_lat* L, const double* G) {
  double value = 0;
  if (L->grad == 1 && L->geom != SFX_FLAT) {
    /* Lat1DCyl2ndO.cpp:319-330: layer volumes L[z], mirror2 edges halved */
    for (int z = 1; z <= L->n[0]; z++) {
      double s = 0;
      for (int k = 0; k < 3; k++) s += orc_wcomp2(L, k) * G[k * L->M + z];
      double w = L->L[z];
      if (z == 1 && L->code1[0] == SFX_BOUND_MIRROR2) w *= 0.5;
      if (z == L->n[0] && L->code2[0] == SFX_BOUND_MIRROR2) w *= 0.5;
      value += w * s;
    }
    return log(value / 6.0);
  }
  if (L->grad == 1) { /* plain sum over the layers (Lat1D2ndO.cpp ComputeLnGN) */
    for (int z = 1; z <= L->n[0]; z++)
      for (int k = 0; k < 3; k++) value += orc_wcomp2(L, k) * G[k * L->M + z];
    return log(value / 6.0);
  }
  if (L->grad == 2) { /* Lat2D2ndO.cpp:387-439: mirror2 edges count half */
    for (int x = 1; x <= L->n[0]; x++)
      for (int y = 1; y <= L->n[1]; y++) {
        long i = x * (long)L->MY + y;
        double w = 1.0;
        if (x == 1 && L->code1[0] == SFX_BOUND_MIRROR2) w *= 0.5;
        if (x == L->n[0] && L->code2[0] == SFX_BOUND_MIRROR2) w *= 0.5;
        if (y == 1 && L->code1[1] == SFX_BOUND_MIRROR2) w *= 0.5;
        if (y == L->n[1] && L->code2[1] == SFX_BOUND_MIRROR2) w *= 0.5;
        if (L->geom == SFX_CYLINDER) w *= L->L[x]; /* Lat2DCyl2ndO.cpp:410-462 */
        double s = 0;
        for (int k = 0; k < 5; k++) s += orc_wcomp2(L, k) * G[k * L->M + i];
        value += w * s;
      }
    return log(value / 6.0);
  }
  double* c = malloc(L->M * sizeof(double));
  for (int k = 0; k < 6; k++) {
    memcpy(c, G + k * L->M, L->M * sizeof(double));
    orc_subtract_boundaries(L, c);
    for (long i = 0; i < L->M; i++)
      if (orc_lat_is_interior(L, i)) value += c[i];
  }
  free(c);
  return log(value / 6.0);
}

/* ConnectG with PHI (Lat3D2ndO.cpp:905-931): out += A^k B^k for every k, and
 * OUT + kk*M, kk the axis of k (0 for k = 0, 5; 1 for 1, 4; 2 for 2, 3). */
void orc_connect2(const orc_lat* L, const double* A, const double* B, double* out, double* OUT) {
  long M = L->M;
  for (int k = 0; k < 6; k++) {
    int kk = (k == 0 || k == 5) ? 0 : (k == 1 || k == 4) ? 1 : 2;
    for (long i = 0; i < M; i++) {
      out[i] += A[k * M + i] * B[k * M + i];
      if (OUT) OUT[kk * M + i] += A[k * M + i] * B[k * M + i];
    }
  }
}

/* CorrectDoubleCountG with PHI (Lat3D2ndO.cpp:988-994, tools.cpp:86-90) */
void orc_correct2(const orc_lat* L, double* in, double* IN, const double* G) {
  long M = L->M;
  for (long i = 0; i < M; i++) in[i] = G[i] > 0 ? in[i] / G[i] : 0.0;
  if (IN)
    for (int k = 0; k < 3; k++)
      for (long i = 0; i < M; i++) IN[k * M + i] = G[i] > 0 ? IN[k * M + i] / G[i] : 0.0;
}
