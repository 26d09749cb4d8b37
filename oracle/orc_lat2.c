/*
 * orc_lat2.c -- ORACLE (test infrastructure only): restatement of the
 * second-order (semiflexible) simple-cubic lattice, Lat3D2ndO.cpp.
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

void orc_propagate2(const orc_lat* L, double* Gin, const double* g, double* Gout, double S, int backward) {
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
