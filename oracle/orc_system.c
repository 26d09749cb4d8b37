/*
 * orc_system.c -- ORACLE (test infrastructure only): restatement of the
 * reference's SCF function evaluation for linear chains on a 1st-order
 * lattice. Follows (file:line in the reference):
 *   iteration-variable layout   SF_Solve.cpp:39-86 (distinct chi rows share a
 *                               variable), :129-163 (iteration sites)
 *   UpdateSWF                   SF_Solve.cpp:794-852, SetFrozenSWFZero :1362-1374
 *   monomer                     SF_Monomer1stO.cpp:101-133
 *   homopolymer (half store)    SF_Homopol1stO.cpp:552-622
 *   copolymer Matrix1           SF_Copol1stO.cpp:1186-1281
 *   copolymer SymMatrix1        SF_Copol1stO.cpp:1630-1729
 *   solvent closure             SF_MolList1stO.cpp:280-295
 *   segment bulk fractions      SF_MolStructure.cpp:605-612, SF_Segment.cpp:443-455
 *   PutSides / ChemInt(Bulk)    SF_SegmentList.cpp:512-623
 *   GetResiduals                SF_Solve.cpp:527-591
 *   GetFreeEnergy               SF_MolList1stO.cpp:606-635, :815-940
 *   chemIntRef                  SF_MolStructure.cpp:412-424
 *   chemical potential          SF_MolList1stO.cpp:389-452 (first-generation
 *                               freedoms, one state per segment, no reactions)
 *   grand potential             SF_MolList1stO.cpp:637-647 over :1001-1100
 *                               (no charges, no grafted chains, no external
 *                               potential)
 * The melt form (SFX_FORM_MELT) replaces the solvent closure by explicit bulk
 * fractions and the free energy by MPS.cpp:253-261.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

typedef struct {
  int N, freedom, kind; /* kind: 0 monomer, 1 homopolymer, 2 copolymer, 3 symmetric copolymer */
  int* seq;
  double phib, theta, lnGN, lnC, chemIntRef;
  double* phi; /* ns*M, this molecule's contribution per state */
  int rlo[3], rhi[3]; /* range_restricted: the box, 1-based interior coordinates */
  double phi_range;
} omol;

struct orc_sys {
  orc_lat* L;
  long M;
  int ns, nm, form;
  double *chi, *fprof, *melt_phib;
  int* frozen;
  omol* mols;
  int n_itvar;
  int* state_itvar; /* state -> itvar, -1 if frozen */
  int* itvar_rep;   /* itvar -> representative state */
  char* itmask;
  long n_sites;
  long* sites;
  double *swf, *phi, *rho, *phitot, *phib_state, *u;
  double *Gi, *Ginv, *Gtmp;
  int maxN;
  int order;
  double stiffness;
};

static int is_palindrome(const int* s, int N) {
  for (int i = 0; i < N / 2; i++)
    if (s[i] != s[N - 1 - i]) return 0;
  return 1;
}

orc_sys* orc_sys_new(const sfx_lattice_desc* ld, const sfx_system_desc* sd) {
  orc_sys* s = calloc(1, sizeof(orc_sys));
  s->L = orc_lat_new(ld);
  if (!s->L) {
    free(s);
    return NULL;
  }
  long M = s->M = s->L->M;
  int ns = s->ns = sd->n_states;
  s->nm = sd->n_molecules;
  s->form = sd->form;
  s->chi = malloc(sizeof(double) * ns * ns);
  memcpy(s->chi, sd->chi, sizeof(double) * ns * ns);
  s->frozen = calloc(ns, sizeof(int));
  if (sd->frozen) memcpy(s->frozen, sd->frozen, sizeof(int) * ns);
  s->fprof = calloc(ns * M, sizeof(double));
  if (sd->frozen_profile)
    for (int i = 0; i < ns; i++)
      if (s->frozen[i]) memcpy(s->fprof + i * M, sd->frozen_profile + i * M, sizeof(double) * M);
  s->melt_phib = calloc(ns, sizeof(double));
  if (sd->melt_phibulk) memcpy(s->melt_phib, sd->melt_phibulk, sizeof(double) * ns);
  /* iteration variables: liquid states with identical chi rows share one */
  s->state_itvar = malloc(sizeof(int) * ns);
  s->itvar_rep = malloc(sizeof(int) * ns);
  s->n_itvar = 0;
  for (int i = 0; i < ns; i++) {
    s->state_itvar[i] = -1;
    if (s->frozen[i]) continue;
    for (int j = 0; j < i && s->state_itvar[i] < 0; j++) {
      if (s->frozen[j]) continue;
      int eq = 1;
      for (int k = 0; k < ns; k++)
        if (s->chi[i * ns + k] != s->chi[j * ns + k]) eq = 0;
      if (eq) s->state_itvar[i] = s->state_itvar[j];
    }
    if (s->state_itvar[i] < 0) s->state_itvar[i] = s->n_itvar++;
    s->itvar_rep[s->state_itvar[i]] = i;
  }
  /* iteration sites: interior, not covered by a frozen range value > 0.999 */
  s->itmask = calloc(M, 1);
  s->sites = malloc(sizeof(long) * M);
  for (long z = 0; z < M; z++) {
    int ok = orc_lat_is_interior(s->L, z);
    for (int i = 0; ok && i < ns; i++)
      if (s->frozen[i] && s->fprof[i * M + z] > 0.999) ok = 0;
    s->itmask[z] = ok;
    if (ok) s->sites[s->n_sites++] = z;
  }
  s->mols = calloc(s->nm, sizeof(omol));
  s->maxN = 1;
  for (int m = 0; m < s->nm; m++) {
    const sfx_molecule_desc* d = &sd->molecules[m];
    omol* o = &s->mols[m];
    o->N = d->length;
    o->freedom = d->freedom;
    o->phib = d->phibulk;
    o->theta = d->theta;
    for (int a = 0; a < 3; a++) {
      o->rlo[a] = d->range_lo[a];
      o->rhi[a] = d->range_hi[a];
    }
    o->phi_range = d->phi_range;
    o->seq = malloc(sizeof(int) * o->N);
    memcpy(o->seq, d->sequence, sizeof(int) * o->N);
    int homo = 1;
    for (int k = 1; k < o->N; k++)
      if (o->seq[k] != o->seq[0]) homo = 0;
    o->kind = o->N == 1 ? 0 : homo ? 1 : is_palindrome(o->seq, o->N) ? 3 : 2;
    o->phi = calloc(ns * M, sizeof(double));
    if (o->N > s->maxN) s->maxN = o->N;
    /* chemIntRef = 1/2 sum_jk f_j f_k chi_jk over the molecule's composition */
    double ref = 0;
    for (int a = 0; a < ns; a++) {
      double fa = 0;
      for (int k = 0; k < o->N; k++) fa += o->seq[k] == a;
      fa /= o->N;
      if (fa == 0) continue;
      double ci = 0;
      for (int b = 0; b < ns; b++) {
        double fb = 0;
        for (int k = 0; k < o->N; k++) fb += o->seq[k] == b;
        ci += s->chi[a * ns + b] * fb / o->N;
      }
      ref += fa * ci;
    }
    o->chemIntRef = ref / 2;
  }
  s->swf = calloc(ns * M, sizeof(double));
  s->phi = calloc(ns * M, sizeof(double));
  s->rho = calloc(ns * M, sizeof(double));
  s->phitot = calloc(M, sizeof(double));
  s->phib_state = calloc(ns, sizeof(double));
  s->u = calloc(s->n_itvar * M, sizeof(double));
  s->order = ld->order == 2 ? 2 : 1;
  s->stiffness = sd->stiffness;
  int D = s->order == 2 ? 6 : 1;
  s->Gi = calloc((size_t)M * D * s->maxN, sizeof(double));
  s->Ginv = calloc(M * D, sizeof(double));
  s->Gtmp = calloc(M * D, sizeof(double));
  return s;
}

void orc_sys_free(orc_sys* s) {
  if (!s) return;
  for (int m = 0; m < s->nm; m++) {
    free(s->mols[m].seq);
    free(s->mols[m].phi);
  }
  free(s->mols);
  orc_lat_free(s->L);
  free(s->chi);
  free(s->frozen);
  free(s->fprof);
  free(s->melt_phib);
  free(s->state_itvar);
  free(s->itvar_rep);
  free(s->itmask);
  free(s->sites);
  free(s->swf);
  free(s->phi);
  free(s->rho);
  free(s->phitot);
  free(s->phib_state);
  free(s->u);
  free(s->Gi);
  free(s->Ginv);
  free(s->Gtmp);
  free(s);
}

long orc_sys_nvar(const orc_sys* s) {
  return s->form == SFX_FORM_MELT ? (long)s->n_itvar * s->M : s->n_sites * s->n_itvar;
}
long orc_sys_M(const orc_sys* s) { return s->M; }
int orc_sys_n_itvar(const orc_sys* s) { return s->n_itvar; }

/* x -> u[itvar][M] (interior iteration sites only, others 0) */
static void unpack_x(orc_sys* s, const double* x) {
  long M = s->M;
  memset(s->u, 0, sizeof(double) * s->n_itvar * M);
  if (s->form == SFX_FORM_MELT) {
    for (int i = 0; i < s->n_itvar; i++)
      for (long z = 0; z < M; z++)
        if (s->itmask[z]) s->u[i * M + z] = x[i * M + z];
  } else {
    for (long k = 0; k < s->n_sites; k++)
      for (int i = 0; i < s->n_itvar; i++) s->u[i * M + s->sites[k]] = x[k * s->n_itvar + i];
  }
}

static void update_swf(orc_sys* s) {
  long M = s->M;
  for (int t = 0; t < s->ns; t++) {
    double* g = s->swf + t * M;
    if (s->frozen[t]) {
      memset(g, 0, sizeof(double) * M);
      continue;
    }
    int iv = s->state_itvar[t];
    for (long z = 0; z < M; z++) g[z] = s->itmask[z] ? exp(-s->u[iv * M + z]) : 0.0;
    orc_setb(s->L, g);
  }
}

static void connect(orc_sys* s, double* out, const double* A, const double* B) {
  for (long z = 0; z < s->M; z++) out[z] += A[z] * B[z];
}

static void correct_dc(orc_sys* s, double* phi, const double* G) {
  for (long z = 0; z < s->M; z++) {
    if (G[z] > 0)
      phi[z] /= G[z];
    else
      phi[z] = 0;
  }
}

static void scale(orc_sys* s, double* phi, double c) {
  for (long z = 0; z < s->M; z++) phi[z] *= c;
}

/* rangeRestricted (SF_Copol1stO.cpp:1246-1280, SF_Homopol1stO.cpp:604-615,
 * SF_Monomer1stO.cpp): phi_raw is the molecule's un-normalised profile per
 * state times `raw`; its sum over the states and the interior sites of the
 * box is brought to phi_range (NormPhiFree with phiRange / phiPreNorm), then
 * theta = ComputeTheta (SF_Molecule.cpp:707-725: sum of site volume x phi over
 * the interior) and phi_b = theta / G_N. */
static int in_range(const orc_sys* s, const omol* o, long i) {
  const orc_lat* L = s->L;
  if (!orc_lat_is_interior(L, i)) return 0;
  int c[3] = {(int)(i / L->jx), (int)((i / L->MZ) % L->MY), (int)(i % L->MZ)};
  for (int a = 0; a < L->grad; a++)
    if (c[a] < o->rlo[a] || c[a] > o->rhi[a]) return 0;
  return 1;
}

static void range_normalize(orc_sys* s, omol* o, double raw) {
  const long M = s->M;
  double pre = 0;
  for (int t = 0; t < s->ns; t++)
    for (long z = 0; z < M; z++)
      if (in_range(s, o, z)) pre += raw * o->phi[t * M + z];
  const double C = o->phi_range / pre;
  o->lnC = log(C);
  for (int t = 0; t < s->ns; t++) scale(s, o->phi + t * M, raw * C);
  double* tot = calloc(M, sizeof(double));
  for (int t = 0; t < s->ns; t++)
    for (long z = 0; z < M; z++) tot[z] += o->phi[t * M + z];
  orc_subtract_boundaries(s->L, tot);
  orc_multiply_sites(s->L, tot);
  double theta = 0;
  for (long z = 0; z < M; z++) theta += tot[z];
  free(tot);
  o->theta = theta;
  o->phib = theta / exp(o->lnGN);
}

/* SF_Homopol2ndO.cpp:152-207 (full sweep over the direction components:
 * six in 3-D, the lumped five in 2-D and three in 1-D, each weighted by its
 * multiplicity in the composition) */
static void mol_compute_phi2(orc_sys* s, omol* o) {
  const int nc = orc_ncomp2(s->L);
  long M = s->M, M6 = nc * s->M;
  int N = o->N;
  double S = s->stiffness;
  double* Gi = s->Gi;
  double* Ginv = s->Ginv;
  for (int k = 0; k < nc; k++) memcpy(Gi + k * M, s->swf + o->seq[0] * M, sizeof(double) * M);
  for (int r = 2; r <= N; r++)
    orc_propagate2(s->L, Gi + (long)(r - 2) * M6, s->swf + o->seq[r - 1] * M, Gi + (long)(r - 1) * M6, S, 0);
  for (int k = 0; k < nc; k++) memcpy(Ginv + k * M, s->swf + o->seq[N - 1] * M, sizeof(double) * M);
  for (int q = N; q >= 1; q--) {
    int t = o->seq[q - 1];
    if (q < N) {
      memcpy(s->Gtmp, Ginv, sizeof(double) * M6);
      orc_propagate2(s->L, s->Gtmp, s->swf + t * M, Ginv, S, 1);
    }
    for (int k = 0; k < nc; k++) {
      const double w = orc_wcomp2(s->L, k);
      const double* A = Ginv + k * M;
      const double* B = Gi + (long)(q - 1) * M6 + k * M;
      for (long z = 0; z < M; z++) o->phi[t * M + z] += w * A[z] * B[z];
    }
  }
  o->lnGN = orc_lngn2(s->L, Ginv);
  for (int t = 0; t < s->ns; t++) {
    int used = 0;
    for (int k = 0; k < N; k++) used |= o->seq[k] == t;
    if (used) correct_dc(s, o->phi + t * M, s->swf + t * M);
  }
  double c;
  if (o->freedom == SFX_RANGE_RESTRICTED) {
    range_normalize(s, o, 1.0 / 6.0);
    return;
  }
  if (o->freedom == SFX_RESTRICTED) {
    o->lnC = log(o->theta) - o->lnGN - log(1.0 * N);
    o->phib = o->theta / exp(o->lnGN);
    c = o->theta / N / (6.0 * exp(o->lnGN));
  } else {
    o->lnC = log(o->phib / N);
    c = o->phib / (6.0 * N);
  }
  for (int t = 0; t < s->ns; t++) scale(s, o->phi + t * M, c);
}

static void mol_compute_phi(orc_sys* s, omol* o) {
  long M = s->M;
  int N = o->N;
  double* G;
  if (s->order == 2 && o->kind != 0) {
    memset(o->phi, 0, sizeof(double) * s->ns * M);
    mol_compute_phi2(s, o);
    return;
  }
  memset(o->phi, 0, sizeof(double) * s->ns * M);
  if (o->kind == 0) { /* SF_Monomer1stO::Matrix1 */
    int t = o->seq[0];
    G = s->swf + t * M;
    memcpy(o->phi + t * M, G, sizeof(double) * M);
    o->lnGN = orc_lngn(s->L, G);
    if (o->freedom == SFX_RANGE_RESTRICTED) {
      range_normalize(s, o, 1.0);
      return;
    }
    if (o->freedom == SFX_RESTRICTED) {
      o->lnC = log(o->theta) - o->lnGN;
      scale(s, o->phi + t * M, o->theta / exp(o->lnGN));
      o->phib = o->theta / exp(o->lnGN);
    } else {
      o->lnC = log(o->phib);
      scale(s, o->phi + t * M, o->phib);
    }
    return;
  }
  int half = o->kind == 1 || o->kind == 3;
  double* Gi = s->Gi;
  double* Ginv = s->Ginv;
  memcpy(Gi, s->swf + o->seq[0] * M, sizeof(double) * M);
  if (half) {
    /* SF_Homopol1stO.cpp:569-596, SF_Copol1stO.cpp:1650-1682 */
    for (int r = 2; r <= N / 2; r++)
      orc_propagate(s->L, Gi + (long)(r - 2) * M, s->swf + o->seq[r - 1] * M, Gi + (long)(r - 1) * M);
    int r = N / 2;
    memcpy(Ginv, Gi + (long)(r - 1) * M, sizeof(double) * M);
    if (N % 2 == 1) {
      int t = o->seq[N / 2];
      memcpy(s->Gtmp, Ginv, sizeof(double) * M);
      orc_propagate(s->L, s->Gtmp, s->swf + t * M, Ginv);
      connect(s, o->phi + t * M, Ginv, Ginv);
      scale(s, o->phi + t * M, 0.5);
    }
    for (int q = (N + 3) / 2; q <= N; q++) {
      int qinv = N - q + 1;
      int t = o->seq[q - 1];
      memcpy(s->Gtmp, Ginv, sizeof(double) * M);
      orc_propagate(s->L, s->Gtmp, s->swf + t * M, Ginv);
      connect(s, o->phi + t * M, Ginv, Gi + (long)(qinv - 1) * M);
    }
  } else {
    /* SF_Copol1stO.cpp:1204-1232 */
    for (int r = 2; r <= N; r++)
      orc_propagate(s->L, Gi + (long)(r - 2) * M, s->swf + o->seq[r - 1] * M, Gi + (long)(r - 1) * M);
    memcpy(Ginv, s->swf + o->seq[N - 1] * M, sizeof(double) * M);
    connect(s, o->phi + o->seq[N - 1] * M, Ginv, Gi + (long)(N - 1) * M);
    for (int q = N - 1; q >= 1; q--) {
      int t = o->seq[q - 1];
      memcpy(s->Gtmp, Ginv, sizeof(double) * M);
      orc_propagate(s->L, s->Gtmp, s->swf + t * M, Ginv);
      connect(s, o->phi + t * M, Ginv, Gi + (long)(q - 1) * M);
    }
  }
  /* the homopolymer corrects before ln G_N, the copolymer after; same result */
  o->lnGN = orc_lngn(s->L, Ginv);
  for (int t = 0; t < s->ns; t++) {
    int used = 0;
    for (int k = 0; k < N; k++) used |= o->seq[k] == t;
    if (used) correct_dc(s, o->phi + t * M, s->swf + t * M);
  }
  double f = half ? 2.0 : 1.0;
  double c;
  if (o->freedom == SFX_RANGE_RESTRICTED) {
    range_normalize(s, o, f);
    return;
  }
  if (o->freedom == SFX_RESTRICTED) {
    o->lnC = log(o->theta) - o->lnGN - log(1.0 * N);
    o->phib = o->theta / exp(o->lnGN);
    c = f * o->theta / N / exp(o->lnGN);
  } else {
    o->lnC = log(o->phib / N);
    c = f * o->phib / N;
  }
  for (int t = 0; t < s->ns; t++) scale(s, o->phi + t * M, c);
}

int orc_sys_compute_phi(orc_sys* s, const double* x) {
  long M = s->M;
  unpack_x(s, x);
  update_swf(s);
  /* SF_MolList1stO.cpp:280-292 */
  double phibSolv = 1;
  int solv = -1;
  for (int m = 0; m < s->nm; m++) {
    if (s->mols[m].freedom == SFX_SOLVENT) {
      solv = m;
      continue;
    }
    mol_compute_phi(s, &s->mols[m]);
    phibSolv -= s->mols[m].phib;
  }
  if (solv >= 0) {
    s->mols[solv].phib = phibSolv;
    mol_compute_phi(s, &s->mols[solv]);
  }
  /* state totals; frozen states carry their range profile */
  memset(s->phi, 0, sizeof(double) * s->ns * M);
  for (int t = 0; t < s->ns; t++) {
    if (s->frozen[t]) {
      memcpy(s->phi + t * M, s->fprof + t * M, sizeof(double) * M);
      continue;
    }
    for (int m = 0; m < s->nm; m++)
      for (long z = 0; z < M; z++) s->phi[t * M + z] += s->mols[m].phi[t * M + z];
  }
  /* segment bulk fractions */
  for (int t = 0; t < s->ns; t++) {
    if (s->form == SFX_FORM_MELT) {
      s->phib_state[t] = s->melt_phib[t];
      continue;
    }
    double pb = 0;
    for (int m = 0; m < s->nm; m++) {
      int cnt = 0;
      for (int k = 0; k < s->mols[m].N; k++) cnt += s->mols[m].seq[k] == t;
      pb += s->mols[m].phib * cnt / s->mols[m].N;
    }
    s->phib_state[t] = s->frozen[t] ? 0 : pb;
  }
  for (int m = 0; m < s->nm; m++)
    if (!isfinite(s->mols[m].lnGN)) return SFX_E_NONFINITE;
  return 0;
}

static void put_sides(orc_sys* s) {
  long M = s->M;
  for (int t = 0; t < s->ns; t++) orc_side(s->L, s->phi + t * M, s->rho + t * M);
  for (long z = 0; z < M; z++) {
    double pt = 0;
    for (int t = 0; t < s->ns; t++) pt += s->phi[t * M + z];
    s->phitot[z] = pt;
  }
}

static double chem_int(const orc_sys* s, int t, long z) {
  double c = 0;
  for (int j = 0; j < s->ns; j++) {
    double x = s->chi[t * s->ns + j];
    if (x != 0) c += x * s->rho[j * s->M + z];
  }
  return c;
}

static double chem_int_bulk(const orc_sys* s, int t) {
  if (s->frozen[t]) return 0;
  double c = 0;
  for (int j = 0; j < s->ns; j++) {
    double x = s->chi[t * s->ns + j];
    if (x != 0) c += x * s->phib_state[j];
  }
  return c;
}

int orc_sys_residual(orc_sys* s, const double* x, double* f) {
  long M = s->M;
  int rc = orc_sys_compute_phi(s, x);
  put_sides(s);
  int nv = s->n_itvar;
  double up[64];
  if (s->form == SFX_FORM_MELT) memset(f, 0, sizeof(double) * nv * M);
  for (long k = 0; k < s->n_sites; k++) {
    long z = s->sites[k];
    double av = 0;
    for (int i = 0; i < nv; i++) {
      int t = s->itvar_rep[i];
      up[i] = -log(s->swf[t * M + z]);
      up[i] -= chem_int(s, t, z) / s->phitot[z];
      up[i] += chem_int_bulk(s, t);
      av += up[i];
    }
    av /= nv;
    for (int i = 0; i < nv; i++) {
      double v = -1.0 + 1.0 / s->phitot[z] + up[i] - av;
      if (s->form == SFX_FORM_MELT)
        f[i * M + z] = v;
      else
        f[k * nv + i] = v;
    }
  }
  return rc;
}

void orc_sys_get_phi(const orc_sys* s, int t, double* out) { memcpy(out, s->phi + t * s->M, sizeof(double) * s->M); }
void orc_sys_get_swf(const orc_sys* s, int t, double* out) { memcpy(out, s->swf + t * s->M, sizeof(double) * s->M); }
double orc_sys_lngn(const orc_sys* s, int m) { return s->mols[m].lnGN; }
double orc_sys_phibulk(const orc_sys* s, int m) { return s->mols[m].phib; }

/* Free energy of the state of the last residual evaluation. */
double orc_sys_free_energy(orc_sys* s) {
  long M = s->M;
  int ns = s->ns;
  if (s->form == SFX_FORM_MELT) {
    /* MPS.cpp:253-261 generalised to n states: sum_{i<j} chi_ij <phi_i>.phi_j
     * - sum_i phi_i.u_i + sum_mol (theta/N) ln(theta/G_N) over the interior. */
    double F = 0;
    for (int i = 0; i < ns; i++)
      for (int j = i + 1; j < ns; j++) {
        double c = s->chi[i * ns + j];
        if (c == 0) continue;
        double d = 0;
        for (long z = 0; z < M; z++)
          if (orc_lat_is_interior(s->L, z)) d += s->rho[i * M + z] * s->phi[j * M + z];
        F += c * d;
      }
    for (int t = 0; t < ns; t++) {
      if (s->frozen[t]) continue;
      int iv = s->state_itvar[t];
      double d = 0;
      for (long z = 0; z < M; z++)
        if (orc_lat_is_interior(s->L, z)) d += s->phi[t * M + z] * s->u[iv * M + z];
      F -= d;
    }
    for (int m = 0; m < s->nm; m++) {
      omol* o = &s->mols[m];
      double n = o->theta / o->N;
      F += n * log(n * o->N / exp(o->lnGN));
    }
    return F;
  }
  /* SF_MolList1stO.cpp:815-940 with no charges and no internal states */
  double* ent = calloc(M, sizeof(double));
  double* con = calloc(M, sizeof(double));
  for (int m = 0; m < s->nm; m++) {
    omol* o = &s->mols[m];
    double N = o->N;
    for (long z = 0; z < M; z++) {
      double pm = 0;
      for (int t = 0; t < ns; t++) pm += o->phi[t * M + z];
      ent[z] -= (pm / N) * (log(N) + o->lnC);
    }
  }
  for (int m = 0; m < s->nm; m++) {
    omol* o = &s->mols[m];
    for (int t = 0; t < ns; t++) {
      int used = 0;
      for (int k = 0; k < o->N; k++) used |= o->seq[k] == t;
      if (!used) continue;
      for (long z = 0; z < M; z++) {
        double G = s->swf[t * M + z];
        if (G > 0) ent[z] -= o->phi[t * M + z] * log(G);
      }
    }
  }
  orc_subtract_boundaries(s->L, ent);
  for (int m = 0; m < s->nm; m++) {
    omol* o = &s->mols[m];
    for (long z = 0; z < M; z++) {
      double pm = 0;
      for (int t = 0; t < ns; t++) pm += o->phi[t * M + z];
      con[z] -= pm * o->chemIntRef;
    }
  }
  for (int t = 0; t < ns; t++)
    for (long z = 0; z < M; z++) con[z] += 0.5 * s->phi[t * M + z] * chem_int(s, t, z);
  orc_subtract_boundaries(s->L, con);
  double* fe = calloc(M, sizeof(double));
  for (long z = 0; z < M; z++) fe[z] = -ent[z] + con[z];
  orc_multiply_sites(s->L, fe);
  double F = 0;
  for (long z = 0; z < M; z++) F += fe[z];
  free(ent);
  free(con);
  free(fe);
  return F;
}

/* Chemical potential of molecule m at the state of the last evaluation
 * (SF_MolList1stO.cpp:389-452): ln phi_b + 1 - N sum_m' phi_b,m'/N_m'
 * + N (sum_t phi_ref,t ChemIntBulk_t - chemIntRef) - N/2 sum_t phi_b,t ChemIntBulk_t,
 * with phi_ref,t the fraction of the chain's segments in state t. */
double orc_sys_chemical_potential(const orc_sys* s, int m) {
  const omol* o = &s->mols[m];
  double N = o->N;
  double v = log(o->phib) + 1.0;
  for (int k = 0; k < s->nm; k++) v -= N * (s->mols[k].phib / s->mols[k].N);
  double d = 0;
  for (int t = 0; t < s->ns; t++) {
    int cnt = 0;
    for (int k = 0; k < o->N; k++) cnt += o->seq[k] == t;
    if (cnt) d += ((double)cnt / o->N) * chem_int_bulk(s, t);
  }
  v += N * (d - o->chemIntRef);
  d = 0;
  for (int t = 0; t < s->ns; t++) d += s->phib_state[t] * chem_int_bulk(s, t);
  v -= 0.5 * N * d;
  return v;
}

/* Grand potential of the last evaluation (SF_MolList1stO.cpp:637-647): the
 * profile of :1001-1100 after SubtractBoundaries, weighted by lattice sites. */
double orc_sys_grand_potential(orc_sys* s) {
  long M = s->M;
  int ns = s->ns;
  double* w = calloc(M, sizeof(double));
  for (int m = 0; m < s->nm; m++) {
    const omol* o = &s->mols[m];
    for (long z = 0; z < M; z++) {
      double pm = 0;
      for (int t = 0; t < ns; t++) pm += o->phi[t * M + z];
      if (pm > 0) w[z] -= (pm - o->phib) / o->N;
    }
  }
  for (int t = 0; t < ns; t++) {
    if (s->frozen[t]) continue;
    double cb = chem_int_bulk(s, t);
    for (long z = 0; z < M; z++) {
      double G = s->swf[t * M + z], up = 0;
      if (G > 0) up = -log(G) - chem_int(s, t, z) + cb;
      w[z] -= up * s->phi[t * M + z];
    }
  }
  for (int i = 0; i < ns; i++) {
    if (s->frozen[i]) continue;
    for (int j = 0; j < ns; j++) {
      if (s->frozen[j]) continue;
      double c = s->chi[i * ns + j];
      if (c == 0) continue;
      for (long z = 0; z < M; z++) {
        w[z] -= 0.5 * c * s->phi[i * M + z] * s->rho[j * M + z];
        if (s->phi[i * M + z] > 0) w[z] += 0.5 * c * s->phib_state[i] * s->phib_state[j];
      }
    }
  }
  orc_subtract_boundaries(s->L, w);
  orc_multiply_sites(s->L, w);
  double G = 0;
  for (long z = 0; z < M; z++) G += w[z];
  free(w);
  return G;
}

/* DIIS of MPS.cpp:228-284 applied to the general residual. */
int orc_sys_solve_diis(orc_sys* s, double* x, double tol, int max_iter, int m, double eta,
                       int* iterations, double* final_norm) {
  long iv = orc_sys_nvar(s);
  double* x0 = calloc(iv, sizeof(double));
  double* xR = calloc(iv * m, sizeof(double));
  double* x_x0 = calloc(iv * m, sizeof(double));
  double* Aij = calloc(m * m, sizeof(double));
  double* Apij = calloc(m * m, sizeof(double));
  double* Ci = calloc(m, sizeof(double));
  double* g = calloc(iv, sizeof(double));
  int it = 0, k_diis = 1, k = 0;
  memcpy(x0, x, iv * sizeof(double)); /* the start (0 in MPS.cpp:263) */
  orc_sys_residual(s, x, g);
  for (long i = 0; i < iv; i++) x[i] += -eta * g[i];
  for (long i = 0; i < iv; i++) x_x0[i] = x[i] - x0[i];
  memcpy(xR, x, iv * sizeof(double));
  double err = 0;
  for (long i = 0; i < iv; i++) err += g[i] * g[i];
  err = sqrt(err);
  while (err > tol && it < max_iter) {
    it++;
    memcpy(x0, x, iv * sizeof(double));
    orc_sys_residual(s, x, g);
    k = it % m;
    k_diis++;
    for (long i = 0; i < iv; i++) x[i] += -eta * g[i];
    memcpy(xR + k * iv, x, iv * sizeof(double));
    for (long i = 0; i < iv; i++) x_x0[k * iv + i] = x[i] - x0[i];
    if (k_diis > m) {
      k_diis = m;
      for (int i = 1; i < m; i++)
        for (int j = 1; j < m; j++) Aij[m * (i - 1) + j - 1] = Aij[m * i + j];
    }
    for (int i = 0; i < k_diis; i++) {
      int posi = k - k_diis + 1 + i;
      if (posi < 0) posi += m;
      double d = 0;
      for (long q = 0; q < iv; q++) d += x_x0[posi * iv + q] * x_x0[k * iv + q];
      Aij[i + m * (k_diis - 1)] = Aij[k_diis - 1 + m * i] = d;
    }
    for (int i = 0; i < k_diis; i++)
      for (int j = 0; j < k_diis; j++) Apij[j + k_diis * i] = Aij[j + m * i];
    orc_diis_coeffs(Apij, k_diis, Ci);
    double normC = 0;
    for (int i = 0; i < k_diis; i++) normC += Ci[i];
    for (int i = 0; i < k_diis; i++) Ci[i] /= normC;
    memset(x, 0, iv * sizeof(double));
    for (int i = 0; i < k_diis; i++) {
      int posi = k - k_diis + 1 + i;
      if (posi < 0) posi += m;
      for (long q = 0; q < iv; q++) x[q] += Ci[i] * xR[posi * iv + q];
    }
    err = 0;
    for (long i = 0; i < iv; i++) err += g[i] * g[i];
    err = sqrt(err);
  }
  if (iterations) *iterations = it;
  if (final_norm) *final_norm = err;
  free(x0);
  free(xR);
  free(x_x0);
  free(Aij);
  free(Apij);
  free(Ci);
  free(g);
  return err <= tol ? 0 : SFX_E_NOTCONVERGED;
}
