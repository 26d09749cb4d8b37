/*
 * orc_lattice.c -- ORACLE (test infrastructure only): FP64 CPU restatement of
 * the reference's lattice layer. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the
 * timed CPU baseline; the product path never does.
 *
 * Site arrays use the reference layout shifted to 0-based indices:
 * i = (x*MY + y)*MZ + z over the halo-padded box (see include/sfx.h).
 *
 * Follows (file:line in the reference):
 *   boundary encoding   Lat1DFlat.cpp:26-88, Lat2DFlat.cpp (ctor), Lat3D.cpp:69-208
 *   SetBoundaries       Lat1DFlat.cpp:396-399, Lat3D.cpp:1510-1562
 *   1-D PropagateG      Lat1DFlat1stO.cpp:101-113, Lat1DCyl1stO.cpp:13-26
 *   curved tables       Lat1DCylinder.cpp:370-414, Lat1DSphere.cpp:240-285
 *   3-D PropagateG      Lat3D1stO.cpp:53-70/:120-175, Lat3DFCC1stO.cpp:55-98,
 *                       Lat3DHEX1stO.cpp:55-98
 *   ComputeLnGN         Lat1DFlat1stO.cpp:301-316, Lat1DCyl1stO.cpp:258-270,
 *                       Lat3D1stO.cpp:345-370
 *   SubtractBoundaries  Lat1DFlat.cpp:409-415, Lat3D.cpp:1591-1640
 *   force PropagateG    Lat1DFlat1stO.cpp:110-130, Lat3D1stO.cpp:94-165,
 *                       Lat3DFCC1stO.cpp:102-145, Lat3DHEX1stO.cpp:102-145
 *   NormPhiFree/Restr   Lat1DFlat1stO.cpp:318-331, Lat3D2ndO.cpp:1027-1047
 *   SideFraction/Sides  Lat1DFlat.cpp:510-524, Lat1DCylinder.cpp:160-171,
 *                       Lat3D.cpp:1729-1784
 *   2-D lattices        Lat2DFlat.cpp, Lat2DFlat1stO.cpp:31-52,
 *                       Lat2DCylinder.cpp:7-55, Lat2DCyl1stO.cpp:16-39
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* halo source coordinate (0-based) of the lower / upper halo of one gradient
 * with n interior layers: Lat1DFlat.cpp:81-88 (1-D/2-D) and Lat3D.cpp:203-208
 * give the same mapping once shifted to 0-based coordinates. */
static int lower_src(int code, int n) {
  switch (code) {
    case SFX_BOUND_SURFACE: return 0;      /* the frozen surface layer itself */
    case SFX_BOUND_MIRROR1: return 1;
    case SFX_BOUND_MIRROR2: return 2;
    case SFX_BOUND_PERIODIC: return n;
  }
  return -1;
}
static int upper_src(int code, int n) {
  switch (code) {
    case SFX_BOUND_SURFACE: return n + 1;
    case SFX_BOUND_MIRROR1: return n;
    case SFX_BOUND_MIRROR2: return n - 1;
    case SFX_BOUND_PERIODIC: return 1;
  }
  return -1;
}

static void curved_tables(orc_lat* L) {
  /* Lat1DCylinder.cpp:370-399 (D = dimensions-1 = 2) and
   * Lat1DSphere.cpp:240-270 (D = dimensions = 3); 1-based z -> i = z-1. */
  int M = L->MX;
  double* V = calloc(M, sizeof(double));
  double* S = calloc(M, sizeof(double));
  double off = L->offset, adj = L->ladj;
  int sphere = L->geom == SFX_SPHERE;
  double D = sphere ? L->dimensions : L->dimensions - 1;
  if (sphere) {
    S[0] = off * off * M_PI * 4;
    V[0] = (4.0 / 3.0) * pow(off, 3) * M_PI;
  } else {
    S[0] = off * M_PI * 2;
    V[0] = off * off * M_PI;
  }
  L->L[0] = V[0];
  for (int z = 2; z <= M; z++) {
    int i = z - 1;
    double x = z + off - 2 + adj;
    if (z == M - 1 && L->adjust_outer) x += -adj + 1;
    V[i] = (M_PI * pow(2, (D - 1)) / D) * pow(x, D);
    L->L[i] = V[i] - V[i - 1];
    S[i] = pow(2, (D - 1)) * M_PI * pow(x, (D - 1));
    L->lm[i] = L->l1 * S[i - 1] / L->L[i];
    L->lp[i] = L->l1 * S[i] / L->L[i];
    L->lmc[i] = L->l1chi * S[i - 1] / L->L[i];
    L->lpc[i] = L->l1chi * S[i] / L->L[i];
  }
  if (sphere) {
    if (L->L[0] > 4.0 * M_PI / 3.0) L->L[0] = V[0] - pow((off - 1), 3) * M_PI * 4.0 / 3.0;
  } else {
    if (L->L[0] > M_PI) L->L[0] = V[0] - (off - 1) * (off - 1) * M_PI;
  }
  if (L->L[0] > 0) {
    L->lp[0] = L->L[1] * L->lm[1] / L->L[0];
    L->lpc[0] = L->L[1] * L->lmc[1] / L->L[0];
  }
  free(V);
  free(S);
}

static void cyl2d_tables(orc_lat* L) {
  /* Lat2DCylinder.cpp:7-55: radial gradient x, axial y; ring volume L[x],
   * surfaces S[x], lambda11[x] = l1*S[x]/L[x], lambda1_1[x] = l1*S[x-1]/L[x].
   * Restated for offset_first_layer = 0 and layer_adjustment = 1. */
  int MX = L->MX;
  double* V = calloc(MX, sizeof(double));
  double* S = calloc(MX, sizeof(double));
  double off = L->offset;
  S[0] = off * M_PI * 2;
  V[0] = off * off * M_PI;
  L->L[0] = V[0];
  for (int x = 2; x <= MX; x++) {
    int i = x - 1;
    double r = x + off - 2 + L->ladj;
    V[i] = M_PI * r * r;
    L->L[i] = V[i] - V[i - 1];
    S[i] = 2 * M_PI * r;
    L->lm[i] = L->l1 * S[i - 1] / L->L[i];
    L->lp[i] = L->l1 * S[i] / L->L[i];
    L->lmc[i] = L->l1chi * S[i - 1] / L->L[i];
    L->lpc[i] = L->l1chi * S[i] / L->L[i];
  }
  L->lp[MX - 1] = 0; /* lambda11[MX] = 0 */
  L->lpc[MX - 1] = 0;
  if (L->L[0] > M_PI) L->L[0] = V[0] - (off - 1) * (off - 1) * M_PI;
  if (L->L[0] > 0) {
    L->lp[0] = L->L[1] * L->lm[1] / L->L[0];
    L->lpc[0] = L->L[1] * L->lmc[1] / L->L[0];
  }
  free(V);
  free(S);
}

orc_lat* orc_lat_new(const sfx_lattice_desc* d) {
  if (!d || d->gradients < 1 || d->gradients > 3) return NULL;
  orc_lat* L = calloc(1, sizeof(orc_lat));
  L->grad = d->gradients;
  L->geom = d->geometry;
  L->type = d->lattice_type;
  L->order = d->order ? d->order : 1;
  for (int k = 0; k < 3; k++) L->n[k] = k < L->grad ? d->n_layers[k] : 0;
  L->MX = L->n[0] + 2;
  L->MY = L->grad >= 2 ? L->n[1] + 2 : 1;
  L->MZ = L->grad >= 3 ? L->n[2] + 2 : 1;
  L->M = (long)L->MX * L->MY * L->MZ;
  L->jx = (long)L->MY * L->MZ;
  L->jy = L->MZ;
  for (int k = 0; k < L->grad; k++) {
    int lo = d->lowerbound[k], up = d->upperbound[k];
    /* one periodic side makes the other periodic (Lat1DFlat.cpp:63-80) */
    if (lo == SFX_BOUND_PERIODIC || up == SFX_BOUND_PERIODIC) lo = up = SFX_BOUND_PERIODIC;
    L->code1[k] = lo;
    L->code2[k] = up;
    L->b1[k] = lower_src(lo, L->n[k]);
    L->b2[k] = upper_src(up, L->n[k]);
  }
  L->dimensions = 3;
  L->l1 = d->lambda;
  L->l1chi = d->lambda_chi > 0 ? d->lambda_chi : d->lambda;
  L->ladj = d->layer_adjustment > 0 ? d->layer_adjustment : 1.0;
  L->adjust_outer = d->layer_adjustment > 0 && d->layer_adjustment != 1.0;
  L->offset = d->offset_first_layer;
  if (L->grad == 3) {
    /* Lat3D.cpp:226-245: standard -> l1 = 1/6; FCC stencil 1/12 with 27-point
     * contact average (l0=l1=l2=l3=1/27); HEX 1/12. */
    L->l1 = 1.0 / 6.0;
  }
  if (L->grad == 2) {
    /* Lat2DFlat: l1x = l1y = l1, l0 = 1 - 4*l1 (standard lattice). */
    L->l1x = L->l1;
    L->l1y = L->l1;
    L->l0_2d = 1 - 4 * L->l1;
    L->l1chi = L->l1; /* 2-D lattices have no lambda_chi */
  }
  if ((L->grad == 1 && L->geom != SFX_FLAT) || (L->grad == 2 && L->geom == SFX_CYLINDER)) {
    L->L = calloc(L->MX, sizeof(double));
    L->lm = calloc(L->MX, sizeof(double));
    L->lp = calloc(L->MX, sizeof(double));
    L->lmc = calloc(L->MX, sizeof(double));
    L->lpc = calloc(L->MX, sizeof(double));
    if (L->grad == 1)
      curved_tables(L);
    else
      cyl2d_tables(L);
  }
  return L;
}

void orc_lat_free(orc_lat* L) {
  if (!L) return;
  free(L->L);
  free(L->lm);
  free(L->lp);
  free(L->lmc);
  free(L->lpc);
  free(L);
}

long orc_lat_M(const orc_lat* L) { return L->M; }

int orc_lat_is_interior(const orc_lat* L, long i) {
  int z = (int)(i % L->MZ);
  int y = (int)((i / L->MZ) % L->MY);
  int x = (int)(i / L->jx);
  if (x < 1 || x > L->n[0]) return 0;
  if (L->grad >= 2 && (y < 1 || y > L->n[1])) return 0;
  if (L->grad >= 3 && (z < 1 || z > L->n[2])) return 0;
  return 1;
}

/* SetBoundaries: every halo site takes the value of the site whose
 * coordinates are remapped per gradient (faces, edges and corners alike,
 * Lat3D.cpp:1510-1562; 1-D Lat1DFlat.cpp:396-399). A surface halo keeps its
 * own value. */
void orc_setb(const orc_lat* L, double* P) {
  int MX = L->MX, MY = L->MY, MZ = L->MZ;
  for (int x = 0; x < MX; x++)
    for (int y = 0; y < MY; y++)
      for (int z = 0; z < MZ; z++) {
        int sx = x, sy = y, sz = z;
        if (x == 0) sx = L->b1[0];
        if (x == MX - 1) sx = L->b2[0];
        if (L->grad >= 2) {
          if (y == 0) sy = L->b1[1];
          if (y == MY - 1) sy = L->b2[1];
        }
        if (L->grad >= 3) {
          if (z == 0) sz = L->b1[2];
          if (z == MZ - 1) sz = L->b2[2];
        }
        if (sx != x || sy != y || sz != z)
          P[((long)x * MY + y) * MZ + z] = P[((long)sx * MY + sy) * MZ + sz];
      }
}

/* removeboundaries: zero every halo site (Lat3D.cpp:1566-1585). */
void orc_zero_halos(const orc_lat* L, double* P) {
  for (long i = 0; i < L->M; i++)
    if (!orc_lat_is_interior(L, i)) P[i] = 0;
}

/* 1st-order propagation Gout = G * stencil(Gin). Gin's halos are set first
 * (the reference propagators read the input halos). Gout's halos are set
 * afterwards. */
void orc_propagate(const orc_lat* L, double* Gin, const double* g, double* Gout) {
  orc_setb(L, Gin);
  if (L->grad == 1) {
    int M = L->MX;
    double a, b = Gin[0], c = Gin[1];
    for (int i = 1; i < M - 1; i++) {
      a = b;
      b = c;
      c = Gin[i + 1];
      if (L->lm)
        Gout[i] = g[i] * (b + L->lm[i] * (a - b) + L->lp[i] * (c - b));
      else
        Gout[i] = g[i] * (b + L->l1 * (a - b - b + c));
    }
  } else if (L->grad == 2) {
    long jx = L->jx;
    for (int x = 1; x <= L->n[0]; x++)
      for (int y = 1; y <= L->n[1]; y++) {
        long i = x * jx + y;
        double a = Gin[i - 1], b = Gin[i], c = Gin[i + 1];
        if (L->lm) /* Lat2DCyl1stO.cpp:24-37 */
          Gout[i] = g[i] * (b + L->lm[x] * (Gin[i - jx] - b) + L->lp[x] * (Gin[i + jx] - b) +
                            L->l1 * (a + c - b - b));
        else /* Lat2DFlat1stO.cpp:44 */
          Gout[i] = g[i] * (L->l0_2d * b + L->l1y * (a + c) + L->l1x * (Gin[i - jx] + Gin[i + jx]));
      }
  } else {
    long jx = L->jx, jy = L->jy, jz = 1;
    for (int x = 1; x <= L->n[0]; x++)
      for (int y = 1; y <= L->n[1]; y++)
        for (int z = 1; z <= L->n[2]; z++) {
          long r = (x * (long)L->MY + y) * L->MZ + z;
          double s;
          if (L->type == SFX_LAT_FCC) {
            s = g[r] * Gin[r - jx - jy];
            s += g[r] * Gin[r + jx + jy];
            s += g[r] * Gin[r - jx + jy];
            s += g[r] * Gin[r + jx - jy];
            s += g[r] * Gin[r - jy - jz];
            s += g[r] * Gin[r + jy + jz];
            s += g[r] * Gin[r - jy + jz];
            s += g[r] * Gin[r + jy - jz];
            s += g[r] * Gin[r - jx - jz];
            s += g[r] * Gin[r + jx + jz];
            s += g[r] * Gin[r - jx + jz];
            s += g[r] * Gin[r + jx - jz];
            Gout[r] = s * (1.0 / 12.0);
          } else if (L->type == SFX_LAT_HEX) {
            s = g[r] * Gin[r - jx];
            s += g[r] * Gin[r + jx];
            s += g[r] * Gin[r - jx + jy];
            s += g[r] * Gin[r + jy];
            s += g[r] * Gin[r - jy];
            s += g[r] * Gin[r + jx - jy];
            s += g[r] * Gin[r - jz + jy];
            s += g[r] * Gin[r - jz];
            s += g[r] * Gin[r + jx - jz];
            s += g[r] * Gin[r - jx + jz];
            s += g[r] * Gin[r + jz];
            s += g[r] * Gin[r - jy + jz];
            Gout[r] = s * (1.0 / 12.0);
          } else {
            s = Gin[r - jx] * g[r];
            s += Gin[r + jx] * g[r];
            s += Gin[r - jy] * g[r];
            s += Gin[r + jy] * g[r];
            s += Gin[r - jz] * g[r];
            s += Gin[r + jz] * g[r];
            Gout[r] = s * (1.0 / 6.0);
          }
        }
  }
  orc_setb(L, Gout);
}

/* SubtractBoundaries: zero halos (non-surface), halve mirror2 edge layers. */
void orc_subtract_boundaries(const orc_lat* L, double* P) {
  if (L->grad == 1) {
    int M = L->MX;
    if (L->code1[0] != SFX_BOUND_SURFACE) P[0] = 0;
    if (L->code1[0] == SFX_BOUND_MIRROR2) P[1] /= 2;
    if (L->code2[0] != SFX_BOUND_SURFACE) P[M - 1] = 0;
    if (L->code2[0] == SFX_BOUND_MIRROR2) P[M - 2] /= 2;
    return;
  }
  for (long i = 0; i < L->M; i++) {
    int z = (int)(i % L->MZ), y = (int)((i / L->MZ) % L->MY), x = (int)(i / L->jx);
    int c[3] = {x, y, z};
    int interior = orc_lat_is_interior(L, i);
    if (!interior) {
      /* a halo coordinate that is not a surface zeroes the site (the x and y
       * tests of Lat2DFlat::SubtractBoundaries apply one after the other);
       * sites whose halo coordinates are all surfaces keep their value */
      int zero = L->grad == 3;
      for (int k = 0; k < L->grad; k++) {
        if (c[k] == 0 && L->code1[k] != SFX_BOUND_SURFACE) zero = 1;
        if (c[k] == L->n[k] + 1 && L->code2[k] != SFX_BOUND_SURFACE) zero = 1;
      }
      if (zero) P[i] = 0;
      continue;
    }
    for (int k = 0; k < L->grad; k++) {
      if (c[k] == 1 && L->code1[k] == SFX_BOUND_MIRROR2) P[i] /= 2;
      if (c[k] == L->n[k] && L->code2[k] == SFX_BOUND_MIRROR2) P[i] /= 2;
    }
  }
}

/* MultiplyWithLatticeSites (Lat1DFlat.cpp:265-269, Lat1DCylinder.cpp:126-131,
 * 2-D cylinder ring sizes, 3-D no-op). */
void orc_multiply_sites(const orc_lat* L, double* P) {
  if (L->grad == 1) {
    if (L->L) {
      for (int i = 0; i < L->MX; i++) P[i] *= L->L[i];
    } else {
      P[1] *= L->ladj;
      P[L->MX - 2] *= 2 - L->ladj;
    }
  } else if (L->grad == 2 && L->L) {
    for (long i = 0; i < L->M; i++) P[i] *= L->L[i / L->jx];
  }
}

/* ComputeLnGN (Lat1DFlat1stO.cpp:301-316, Lat1DCyl1stO.cpp:258-270,
 * Lat3D1stO.cpp:345-370). G is not modified. */
double orc_lngn(const orc_lat* L, const double* G) {
  if (L->grad == 1) {
    int M = L->MX;
    double value;
    if (L->L) {
      value = G[1] * L->L[1];
      if (L->code1[0] == SFX_BOUND_MIRROR2) value /= 2;
      for (int i = 2; i <= M - 3; i++) value += G[i] * L->L[i];
      if (L->code2[0] == SFX_BOUND_MIRROR2)
        value += G[M - 2] * L->L[M - 2] / 2;
      else if (M > 3)
        value += G[M - 2] * L->L[M - 2];
    } else {
      value = G[1] * L->ladj;
      if (L->code1[0] == SFX_BOUND_MIRROR2) value /= 2;
      for (int i = 2; i <= M - 3; i++) value += G[i];
      if (L->code2[0] == SFX_BOUND_MIRROR2)
        value += G[M - 2] / 2;
      else if (M > 3)
        value += G[M - 2] * (2 - L->ladj);
    }
    return log(value);
  }
  double* c = malloc(L->M * sizeof(double));
  memcpy(c, G, L->M * sizeof(double));
  orc_subtract_boundaries(L, c);
  orc_multiply_sites(L, c);
  double value = 0;
  for (long i = 0; i < L->M; i++)
    if (orc_lat_is_interior(L, i)) value += c[i];
  free(c);
  return log(value);
}

/* 2-D SideFraction on any site, following the branch order of
 * Lat2DFlat.cpp:996-1053 and Lat2DCylinder.cpp:205-320: corners remap x (if
 * not a surface) else y; then the y edges, then the x edges; a surface edge has
 * its own formula; everything else falls through to the interior formula. The
 * cylinder keeps the original x (its lambda tables) and the original value. */
static double side2d(const orc_lat* L, const double* P, int x, int y) {
  const int MX = L->MX, MY = L->MY;
  const int ox = x;
  const double inz = P[x * MY + y];
  const double l0 = L->l0_2d, l1x = L->l1x, l1y = L->l1y, l1 = L->l1;
  const int cyl = L->lm != NULL;
#define IN(X, Y) P[(X) * MY + (Y)]
#define LAM0(X) (1 - 2 * l1 - L->lp[X] - L->lm[X])
  int sx1 = L->code1[0] == SFX_BOUND_SURFACE, sx2 = L->code2[0] == SFX_BOUND_SURFACE;
  int sy1 = L->code1[1] == SFX_BOUND_SURFACE, sy2 = L->code2[1] == SFX_BOUND_SURFACE;
  int cx = x == 0 || x == MX - 1, cy = y == 0 || y == MY - 1;
  if (cx && cy) {
    int surfx = x == 0 ? sx1 : sx2, surfy = y == 0 ? sy1 : sy2;
    if (!surfx)
      x = x == 0 ? L->b1[0] : L->b2[0];
    else if (!surfy)
      y = y == 0 ? L->b1[1] : L->b2[1];
    else {
      int dx = x == 0 ? 1 : -1, dy = y == 0 ? 1 : -1;
      if (cyl) return LAM0(x) * inz + l1 * IN(x, y + dy) + (x == 0 ? L->lp[x] : L->lm[x]) * IN(x + dx, y);
      return l0 * IN(x, y) + l1y * IN(x, y + dy) + l1x * IN(x + dx, y);
    }
  }
  if (y == 0) {
    if (sy1) {
      if (cyl) return LAM0(x) * inz + L->lp[x] * IN(x + 1, y) + L->lm[x] * IN(x - 1, y) + l1 * IN(x, y + 1);
      return l0 * IN(x, y) + l1x * (IN(x - 1, y) + IN(x + 1, y)) + l1y * IN(x, y + 1);
    }
    y = L->b1[1];
  } else if (y == MY - 1) {
    if (sy2) {
      if (cyl) return LAM0(x) * inz + L->lp[x] * IN(x + 1, y) + L->lm[x] * IN(x - 1, y) + l1 * IN(x, y - 1);
      return l0 * IN(x, y) + l1x * (IN(x - 1, y) + IN(x + 1, y)) + l1y * IN(x, y - 1);
    }
    y = L->b2[1];
  } else if (x == 0) {
    if (sx1) {
      if (cyl) return LAM0(0) * inz + l1 * (IN(x, y - 1) + IN(x, y + 1)) + L->lp[0] * IN(x + 1, y);
      return l0 * IN(x, y) + l1y * (IN(x, y - 1) + IN(x, y + 1)) + l1x * IN(x + 1, y);
    }
    x = L->b1[0];
  } else if (x == MX - 1) {
    if (sx2) {
      if (cyl) return LAM0(MX - 1) * inz + l1 * (IN(x, y - 1) + IN(x, y + 1)) + L->lm[MX - 1] * IN(x - 1, y);
      return l0 * IN(x, y) + l1y * (IN(x, y - 1) + IN(x, y + 1)) + l1x * IN(x - 1, y);
    }
    x = L->b2[0];
  }
  if (cyl)
    return inz + L->lp[ox] * (IN(x + 1, y) - inz) + L->lm[ox] * (IN(x - 1, y) - inz) +
           l1 * (IN(x, y - 1) + IN(x, y + 1) - inz - inz);
  return l0 * IN(x, y) + l1x * (IN(x - 1, y) + IN(x + 1, y)) + l1y * (IN(x, y - 1) + IN(x, y + 1));
#undef IN
#undef LAM0
}

/* Contact average <P> on every interior site (1-D/2-D SideFraction after
 * SetBoundaries, SF_SegmentList.cpp:521-528; 3-D Lat3D::Sides). P's halos are
 * set first. Halos of out are set with SetBoundaries. */
void orc_side(const orc_lat* L, double* P, double* out) {
  orc_setb(L, P);
  if (L->grad == 1) {
    int M = L->MX;
    for (int i = 0; i < M; i++) {
      int z = i;
      if (z == 0) {
        if (L->code1[0] == SFX_BOUND_SURFACE) {
          double lp = L->lpc ? L->lpc[0] : L->l1chi;
          out[i] = P[0] + lp * (-P[0] - P[0] + P[1]);
          continue;
        }
        z = L->b1[0];
      }
      if (z == M - 1) {
        if (L->code2[0] == SFX_BOUND_SURFACE) {
          double lm = L->lmc ? L->lmc[M - 1] : L->l1chi;
          out[i] = P[M - 1] + lm * (-P[M - 1] - P[M - 1] + P[M - 2]);
          continue;
        }
        z = L->b2[0];
      }
      if (L->lmc)
        out[i] = P[z] + L->lpc[z] * (P[z + 1] - P[z]) + L->lmc[z] * (P[z - 1] - P[z]);
      else
        out[i] = P[z] + L->l1chi * (P[z - 1] - P[z] - P[z] + P[z + 1]);
    }
    return;
  }
  if (L->grad == 2) {
    for (int x = 0; x < L->MX; x++)
      for (int y = 0; y < L->MY; y++) out[x * L->jx + y] = side2d(L, P, x, y);
    return;
  }
  long jx = L->jx, jy = L->jy;
  for (int x = 1; x <= L->n[0]; x++)
    for (int y = 1; y <= L->n[1]; y++)
      for (int z = 1; z <= L->n[2]; z++) {
        long r = (x * (long)L->MY + y) * L->MZ + z;
        double s = 0;
        if (L->type == SFX_LAT_HEX) {
          const double lam = 1.0 / 12.0;
          s = P[r + jx] * lam + P[r - jx] * lam + P[r - jy + jx] * lam + P[r - jy] * lam +
              P[r + jy] * lam + P[r - jx + jy] * lam + P[r - jy + 1] * lam + P[r + 1] * lam +
              P[r - jx + 1] * lam + P[r - 1 + jx] * lam + P[r - 1] * lam + P[r - 1 + jy] * lam;
        } else if (L->type == SFX_LAT_FCC) {
          const double l = 1.0 / 27.0;
          for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dz = -1; dz <= 1; dz++) s += l * P[r + dx * jx + dy * jy + dz];
        } else {
          const double l1 = 1.0 / 6.0;
          s = l1 * P[r + jx] + l1 * P[r - jx] + l1 * P[r + jy] + l1 * P[r - jy] + l1 * P[r + 1] +
              l1 * P[r - 1];
        }
        out[r] = s;
      }
  orc_zero_halos(L, out);
}

/* ---- force ensemble: PropagateG(Gi, G, Gout, f) --------------------------
 * Lat1DFlat1stO.cpp:110-130: norm = l1 (e^-f + e^f - 2) + 1,
 *   f1 = l1 e^f / norm on G(z-1), f_1 = l1 e^-f / norm on G(z+1),
 *   Gout(z) = G(z) (b + f1 (a - b) + f_1 (c - b)).
 * Lat3D1stO.cpp:94-165 (simple cubic): fnorm = 1 / (4 + e^f + e^-f); the
 *   neighbour at z-1 weighs e^f, at z+1 e^-f, the four others 1.
 * Lat3DFCC1stO.cpp:102-145, Lat3DHEX1stO.cpp:102-145: the same weights by the
 *   z offset of each of the twelve neighbours, fnorm = 1 / (4 + 4 e^f + 4 e^-f)
 *   for both lattices. */
void orc_propagate_force(const orc_lat* L, double* Gin, const double* g, double* Gout, double f) {
  orc_setb(L, Gin);
  const double ef = exp(f), emf = exp(-f);
  if (L->grad == 1) {
    const double norm = L->l1 * (exp(-f) + exp(f) - 2) + 1;
    const double f_1 = L->l1 * exp(-f) / norm, f1 = L->l1 * exp(f) / norm;
    int M = L->MX;
    double a, b = Gin[0], c = Gin[1];
    for (int i = 1; i < M - 1; i++) {
      a = b;
      b = c;
      c = Gin[i + 1];
      Gout[i] = g[i] * (b + f1 * (a - b) + f_1 * (c - b));
    }
    orc_setb(L, Gout);
    return;
  }
  long jx = L->jx, jy = L->jy, jz = 1;
  const double fnorm = L->type == SFX_LAT_STANDARD ? 1.0 / (4.0 + ef + emf) : 1.0 / (4.0 + 4.0 * ef + 4.0 * emf);
  for (int x = 1; x <= L->n[0]; x++)
    for (int y = 1; y <= L->n[1]; y++)
      for (int z = 1; z <= L->n[2]; z++) {
        long r = (x * (long)L->MY + y) * L->MZ + z;
        double s;
        if (L->type == SFX_LAT_FCC) {
          s = g[r] * Gin[r - jx - jy];
          s += g[r] * Gin[r + jx + jy];
          s += g[r] * Gin[r + jy - jx];
          s += g[r] * Gin[r + jx - jy];
          s += ef * g[r] * Gin[r - jy - jz];
          s += emf * g[r] * Gin[r + jy + jz];
          s += emf * g[r] * Gin[r + jz - jy];
          s += ef * g[r] * Gin[r + jy - jz];
          s += ef * g[r] * Gin[r - jx - jz];
          s += emf * g[r] * Gin[r + jx + jz];
          s += emf * g[r] * Gin[r + jz - jx];
          s += ef * g[r] * Gin[r + jx - jz];
        } else if (L->type == SFX_LAT_HEX) {
          s = g[r] * Gin[r - jx];
          s += g[r] * Gin[r + jx];
          s += g[r] * Gin[r + jy - jx];
          s += g[r] * Gin[r + jy];
          s += g[r] * Gin[r - jy];
          s += g[r] * Gin[r + jx - jy];
          s += ef * g[r] * Gin[r + jy - jz];
          s += ef * g[r] * Gin[r - jz];
          s += ef * g[r] * Gin[r + jx - jz];
          s += emf * g[r] * Gin[r + jz - jx];
          s += emf * g[r] * Gin[r + jz];
          s += emf * g[r] * Gin[r + jz - jy];
        } else {
          s = Gin[r - jx] * g[r];
          s += Gin[r + jx] * g[r];
          s += Gin[r - jy] * g[r];
          s += Gin[r + jy] * g[r];
          s += ef * Gin[r - jz] * g[r];
          s += emf * Gin[r + jz] * g[r];
        }
        Gout[r] = s * fnorm;
      }
  orc_setb(L, Gout);
}

/* ---- normalisations (Lat1DFlat1stO.cpp:318-331, Lat3D2ndO.cpp:1027-1047) -
 * NormPhiFree: phi *= C on every site (and the three axis parts of PHI).
 * NormPhiRestr: C /= exp(ComputeLnGN(Gi)) first; on a 3-D second-order
 * lattice C /= 6 exp(ComputeLnGN(Gi)) with the six-component ComputeLnGN. */
void orc_norm_phi_free(const orc_lat* L, double* phi, double* PHI, double C) {
  for (long i = 0; i < L->M; i++) phi[i] *= C;
  if (PHI)
    for (long i = 0; i < 3 * L->M; i++) PHI[i] *= C;
}

void orc_norm_phi_restr(const orc_lat* L, double* phi, double* PHI, const double* Gi, double C) {
  if (L->order == 2)
    C /= 6.0 * exp(orc_lngn2(L, Gi));
  else
    C /= exp(orc_lngn(L, Gi));
  orc_norm_phi_free(L, phi, PHI, C);
}
