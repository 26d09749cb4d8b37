// sfx_adjoint.hip -- exact Jacobian products of the SCF residual.
//
// Tangent (forward-mode) propagators: with g_t = exp(-u_t) and a direction v,
//   G(1) = g,             dG(1) = -v g
//   G(s) = g A[G(s-1)],   dG(s) = g (A[dG(s-1)] - v A[G(s-1)])
// and the same for the backward propagator, where A is the lattice stencil of
// the reference's PropagateG. The composition phi = pref * sum G Ginv / g gives
//   dphi = phi (dpref/pref + v) + pref * sum (dG Ginv + G dGinv) / g,
// dpref/pref = -dlnG_N for restricted chains (theta / (N G_N)), 0 for free
// chains, dphi_b,solvent / phi_b,solvent for the solvent closure. The residual
// of SF_Solve.cpp:554-590 is then linearised site by site. Every product is
// exact to rounding; there is no finite-difference step anywhere.
//
// The transpose product J^T w runs the same recursions backwards (adjoint
// propagators with the transposed stencil, applied as a scatter).
#include <cmath>
#include <type_traits>

#include "sfx_device.h"
#include "sfx_order2.h"
#include "sfx_block3.h"
#include "sfx_solve.h"

namespace sfx {

constexpr int kTB = 256;

template <typename T>
static __global__ void k_zero(T* p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = T(0);
}
template <typename T>
static void zero(T* p, long long n, hipStream_t st) {
  k_zero<T><<<grid1(n, kTB), kTB, 0, st>>>(p, n);
}

struct Tangent {
  long long M = 0;
  int maxstore = 0;
  DevBuf<double> dstore, dpp, dacc, dphi_m, dphi, drho, dphitot, dmol, v64, v32, partial2;
  // transposed products (allocated by the first VJP): per-state adjoints of
  // the contact averages, of phi and of the per-state direction, the total's
  // adjoint, and the per-state sweep direction (FP64 and FP32)
  DevBuf<double> rbar, phibar, vbar, ptbar, xdir64, xdir32, vpart;
  bool vjp_ready = false;
  void init_vjp(System& S) {
    const size_t nsM = (size_t)S.ns * M;
    rbar.alloc(nsM);
    phibar.alloc(nsM);
    vbar.alloc(nsM);
    ptbar.alloc((size_t)M);
    xdir64.alloc(nsM);
    xdir32.alloc(nsM);  // half used
    vpart.alloc((size_t)SFX_MAXM * S.npart);
    vjp_ready = true;
  }
  DevBuf<double> dacc2, dpp2;   // the backward chain's, in a middle-out sweep
  void init(System& S) {
    M = S.lat.d.M;
    maxstore = S.store_slots;
    const size_t comps = S.lat.d.order == 2 ? 6 : 1;
    dstore.alloc((size_t)maxstore * comps * M);  // the same slots as the propagator store
    dpp.alloc(2 * comps * (size_t)M);
    dacc.alloc((size_t)S.nm * S.ns * M);
    dphi_m.alloc((size_t)S.nm * S.ns * M);
    dphi.alloc((size_t)S.ns * M);
    drho.alloc((size_t)S.ns * M);
    dphitot.alloc((size_t)M);
    dmol.alloc((size_t)S.nm * 4);
    v64.alloc((size_t)S.ns * M);
    v32.alloc((size_t)S.ns * M);  // per-state FP32 directions (half used)
    if (S.use_block3) {
      dacc2.alloc((size_t)S.nm * S.ns * M);
      dpp2.alloc(2 * (size_t)M);
      zero<double>(dpp2.p, 2 * (long long)M, S.stream);  // halo slots are never written
    }
    partial2.alloc((size_t)S.npart);
  }
};


static __global__ void k_to_f32(const double* a, float* b, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

template <typename T>
__device__ __forceinline__ T vdir(const DevLattice& L, const T* v, int x, int y, int z) {
  return v[sidx(L, x, y, z)];
}

// rank 1: G = g, dG = -v g
template <typename T>
struct TInit {
  const T* g;
  const T* v;
  T* G;
  T* dG;
  RankScale sc;
};

// rank 1 of one chain, or of two (blockIdx.y) when a middle-out sweep seeds both
template <typename T>
__global__ void k_tinit2(DevLattice L, TInit<T> c0, TInit<T> c1) {
  const TInit<T>& c = blockIdx.y ? c1 : c0;
  const T* __restrict__ g = c.g;
  const T* __restrict__ v = c.v;
  T* __restrict__ G = c.G;
  T* __restrict__ dG = c.dG;
  const RankScale& sc = c.sc;
  long long q = site_ordinal();
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    T gv = g[r];
    G[r] = gv;
    dG[r] = -v[r] * gv;
    m = (float)gv;
  }
  if (sc.maxv) {
    __shared__ float sm[kTB / 64];
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, sm[i]);
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = 0.0;
    }
  }
}

template <typename T>
static void tinit(const DevLattice& L, const T* g, const T* v, T* G, T* dG, const RankScale& sc, hipStream_t st) {
  const TInit<T> c{g, v, G, dG, sc};
  k_tinit2<T><<<dim3(grid1(L.V, kTB), 1), kTB, 0, st>>>(L, c, c);
}
template <typename T>
static void tinit2(const DevLattice& L, const TInit<T>& c0, const TInit<T>& c1, hipStream_t st) {
  k_tinit2<T><<<dim3(grid1(L.V, kTB), 2), kTB, 0, st>>>(L, c0, c1);
}

// G' = g A[G] s, dG' = g (A[dG] - v A[G]) s; optional composition of
// acc += G' Gp c, dacc += (dG' Gp + G' dGp) c
template <typename T, int KIND, bool SCALED, bool CONNECT>
__global__ void __launch_bounds__(kTB) k_tprop(DevLattice L, const T* __restrict__ Gin, const T* __restrict__ dGin,
                                               const T* __restrict__ g, const T* __restrict__ v,
                                               T* __restrict__ Gout, T* __restrict__ dGout,
                                               const T* __restrict__ Gp, const T* __restrict__ dGp,
                                               T* __restrict__ dacc, RankScale sc) {
  long long q = site_ordinal();
  const bool in = q < L.V;
  // this launch's loads first: the step scale below waits on the previous
  // launch's maxima while they are in flight
  long long r = 0;
  T a = T(0), da = T(0), gv = T(0), vr = T(0), gpr = T(0), dgpr = T(0), dac = T(0);
  if (in) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    r = sidx(L, x, y, z);
    a = stencil<T, KIND>(L, Gin, x, y, z);
    da = stencil<T, KIND>(L, dGin, x, y, z);
    gv = g[r];
    vr = v[r];
    if constexpr (CONNECT) {
      gpr = Gp[r];
      dgpr = dGp[r];
      dac = dacc[r];
    }
  }
  T scale = T(1), cfac = T(1);
  double lncum = 0;
  if constexpr (SCALED) {
    const double l0 = sc.lnv[sc.step - 1];
    double lp = 0.0, rf = 0.0;
    if constexpr (CONNECT) {
      lp = sc.lnpair[sc.pair];
      rf = *sc.ref;
    }
    float mprev, unused;
    block_step_max2(sc.maxv, sc.step - 1, nullptr, 0, mprev, unused);
    if (!(mprev > 0.f)) mprev = 1.f;
    lncum = l0 + log((double)mprev);
    scale = (T)(1.f / mprev);
    if constexpr (CONNECT) cfac = (T)exp(lncum + lp - rf);
  }
  float m = 0.f;
  if (in) {
    T Go = gv * a * scale;
    T dGo = gv * (da - vr * a) * scale;
    Gout[r] = Go;
    dGout[r] = dGo;
    if constexpr (CONNECT) dacc[r] = dac + (dGo * gpr + Go * dgpr) * cfac;
    m = (float)Go;
  }
  if constexpr (SCALED) {
    __shared__ float sm[kTB / 64];
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, sm[i]);
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = lncum;
    }
  }
}

// dacc += (dA B + A dB) * half * c at interior sites
template <typename T, bool SCALED>
__global__ void k_tconnect(DevLattice L, const T* __restrict__ A, const T* __restrict__ dA, const T* __restrict__ B,
                           const T* __restrict__ dB, T* __restrict__ dacc, RankScale sc, double half) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  T c = T(half);
  if constexpr (SCALED) c = (T)(half * exp(sc.lnv[sc.step] + sc.lnpair[sc.pair] - *sc.ref));
  int x, y, z;
  interior_coords(L, q, x, y, z);
  long long r = sidx(L, x, y, z);
  dacc[r] += (dA[r] * B[r] + A[r] * dB[r]) * c;
}

// two weighted sums at once: sum w G and sum w dG (per block)
template <typename T>
__global__ void k_wsum2(DevLattice L, const T* __restrict__ G, const T* __restrict__ dG, double* __restrict__ p1,
                        double* __restrict__ p2) {
  long long q = site_ordinal();
  double a = 0, b = 0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    double w = lngn_weight(L, x, y, z);
    a = (double)G[r] * w;
    b = (double)dG[r] * w;
  }
  __shared__ double sa[kTB / 64], sb[kTB / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0, B = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      A += sa[i];
      B += sb[i];
    }
    p1[blockIdx.x] = A;
    p2[blockIdx.x] = B;
  }
}

// monomer: sum w g and sum w (-v g)
__global__ void k_wsum_mono(DevLattice L, const double* __restrict__ g, const double* __restrict__ v,
                            double* __restrict__ p1, double* __restrict__ p2) {
  long long q = site_ordinal();
  double a = 0, b = 0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    double w = lngn_weight(L, x, y, z);
    a = g[r] * w;
    b = -v[r] * g[r] * w;
  }
  __shared__ double sa[kTB / 64], sb[kTB / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0, B = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      A += sa[i];
      B += sb[i];
    }
    p1[blockIdx.x] = A;
    p2[blockIdx.x] = B;
  }
}

// tangent molecule scalars: dmol[0] = dlnGN, [1] = dpref/pref, [2] = dphib
__global__ void k_tmol_finalize(MolParams mp, const double* p1, const double* p2, int n, const double* molscal,
                                double* dmol, int nmol) {
  __shared__ double r1[256 / 64], r2[256 / 64];
  double s1 = 0, s2 = 0;
  if (p1 && p2)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      s1 += p1[i];
      s2 += p2[i];
    }
  for (int o = 32; o > 0; o >>= 1) {
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    r1[threadIdx.x >> 6] = s1;
    r2[threadIdx.x >> 6] = s2;
  }
  __syncthreads();
  if (threadIdx.x || blockIdx.x) return;
  s1 = s2 = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
    s1 += r1[i];
    s2 += r2[i];
  }
  double* d = dmol + mp.index * 4;
  double dln = (p1 && n > 0) ? s2 / s1 : 0.0;
  const double* ms = molscal + mp.index * MS_COUNT;
  double phib = ms[MS_PHIB];
  // the prefactor of the tangent sweep's scaling (its own reference log-scale)
  const double f = mp.kind == 2 ? 2.0 : mp.kind == 3 ? 1.0 / 6.0 : 1.0;
  if (mp.freedom == SFX_RESTRICTED)
    d[3] = f * exp(log(mp.theta) - ms[MS_LNGN] - (mp.kind == 0 ? 0.0 : log((double)mp.N)) + ms[MS_REF]);
  else
    d[3] = f * (mp.kind == 0 ? phib : phib / mp.N) * exp(ms[MS_REF]);
  d[0] = dln;
  if (mp.freedom == SFX_RESTRICTED) {
    d[1] = -dln;
    d[2] = -phib * dln;
  } else if (mp.freedom == SFX_SOLVENT) {
    double dpb = 0;
    for (int m = 0; m < nmol; m++)
      if (m != mp.index) dpb -= dmol[m * 4 + 2];
    d[2] = dpb;
    d[1] = phib != 0 ? dpb / phib : 0.0;
  } else {
    d[1] = 0;
    d[2] = 0;
  }
}

// dphi_{m,t} at interior sites
template <typename T>
__global__ void k_tmol_phi(DevLattice L, unsigned uses, int ns, const double* __restrict__ phi_m,
                           const T* __restrict__ dacc, const T* __restrict__ dacc2, const double* __restrict__ g,
                           const double* __restrict__ v, const double* __restrict__ dm, int monomer,
                           double* __restrict__ dphi_m) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r0 = sidx(L, x, y, z);
  const double rel = dm[1];
  for (int t = 0; t < ns; t++) {  // every state of the molecule (bit t of uses), arrays [state][M]
    if (!((uses >> t) & 1u)) continue;
    const long long r = (long long)t * L.M + r0;
    const double ph = phi_m[r];
    const double gv = g[r];
    double out;
    if (monomer)
      out = ph * (rel - v[r]);
    else
      out = gv > 0 ? ph * (rel + v[r]) + dm[3] * ((double)dacc[r] + (dacc2 ? (double)dacc2[r] : 0.0)) / gv : 0.0;
    dphi_m[r] = out;
  }
}

// tangent totals (frozen states have no tangent); halos from their sources
__global__ void k_tphi_total(DevLattice L, int nstates, int nmol, const double* __restrict__ dphi_m,
                             const int* __restrict__ mol_uses, const int* __restrict__ frozen,
                             double* __restrict__ dphi, double* __restrict__ dphitot) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  int sx = remap(x, L.nx, L.b1[0], L.b2[0]);
  int sy = L.grad >= 2 ? remap(y, L.ny, L.b1[1], L.b2[1]) : y;
  int sz = L.grad >= 3 ? remap(z, L.nz, L.b1[2], L.b2[2]) : z;
  long long s = sidx(L, sx, sy, sz);
  bool surf_halo = !is_interior(L, sx, sy, sz);
  double tot = 0;
  for (int t = 0; t < nstates; t++) {
    double v = 0;
    if (!frozen[t] && !surf_halo)
      for (int m = 0; m < nmol; m++)
        if (mol_uses[m * nstates + t]) v += dphi_m[((long long)m * nstates + t) * L.M + s];
    dphi[(long long)t * L.M + i] = v;
    tot += v;
  }
  dphitot[i] = tot;
}

__global__ void k_tsides(DevLattice L, int nstates, const double* __restrict__ dphi, double* __restrict__ drho) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  long long r = sidx(L, x, y, z);
  for (int t = 0; t < nstates; t++) drho[(long long)t * L.M + r] = side_at(L, dphi + (long long)t * L.M, x, y, z);
}

// linearised residual (SF_Solve.cpp:554-590)
__global__ void k_jres(DevLattice L, SysParams sp, const unsigned char* __restrict__ itmask,
                       const double* __restrict__ v, const double* __restrict__ rho,
                       const double* __restrict__ drho, const double* __restrict__ phitot,
                       const double* __restrict__ dphitot, const double* __restrict__ molscal,
                       const double* __restrict__ dmol, double* __restrict__ jv) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  const int nv = sp.n_itvar;
  if (!itmask[i]) {
    for (int k = 0; k < nv; k++) jv[(long long)k * L.M + i] = 0.0;
    return;
  }
  double pt = phitot[i], dpt = dphitot[i];
  double du[SFX_MAXS];
  double av = 0;
  for (int k = 0; k < nv; k++) {
    int t = sp.itvar_rep[k];
    double ci = 0, dci = 0, dcb = 0;
    for (int j = 0; j < sp.ns; j++) {
      double c = sp.chi[t * SFX_MAXS + j];
      if (c == 0) continue;
      ci += c * rho[(long long)j * L.M + i];
      dci += c * drho[(long long)j * L.M + i];
      if (!sp.melt && !sp.frozen[j]) {
        double dpb = 0;
        for (int m = 0; m < sp.nm; m++) dpb += dmol[m * 4 + 2] * sp.mol_frac[m * SFX_MAXS + j];
        dcb += c * dpb;
      }
    }
    double d = v[(long long)k * L.M + i] - dci / pt + ci * dpt / (pt * pt) + dcb;
    du[k] = d;
    av += d;
  }
  av /= nv;
  for (int k = 0; k < nv; k++) jv[(long long)k * L.M + i] = -dpt / (pt * pt) + du[k] - av;
}

// per-state direction v_t = v_{itvar(t)} (zero for frozen states), in T
// (out32 optional: the same directions rounded to FP32)
__global__ void k_state_dir(long long M, int ns, const int* st_itvar, const double* v, double* out,
                            float* out32) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  for (int t = 0; t < ns; t++) {
    int iv = st_itvar[t];
    const double d = iv >= 0 ? v[(long long)iv * M + i] : 0.0;
    out[(long long)t * M + i] = d;
    if (out32) out32[(long long)t * M + i] = (float)d;
  }
}

// ---------------------------------------------------------------------------
template <typename T, bool SCALED, bool CONNECT>
static void tprop_k(const DevLattice& L, const T* Gin, const T* dGin, const T* g, const T* v, T* Gout, T* dGout,
                    const T* Gp, const T* dGp, T* dacc, const RankScale& sc, hipStream_t st) {
  unsigned nb = grid1(L.V, kTB);
#define TP(K) k_tprop<T, K, SCALED, CONNECT><<<nb, kTB, 0, st>>>(L, Gin, dGin, g, v, Gout, dGout, Gp, dGp, dacc, sc)
  switch (L.kind) {
    case ST_1D: TP(ST_1D); break;
    case ST_2D: TP(ST_2D); break;
    case ST_3D_SC: TP(ST_3D_SC); break;
    case ST_3D_FCC: TP(ST_3D_FCC); break;
    case ST_3D_HEX: TP(ST_3D_HEX); break;
  }
#undef TP
}
template <typename T>
static void tprop(const DevLattice& L, const T* Gin, const T* dGin, const T* g, const T* v, T* Gout, T* dGout,
                  const T* Gp, const T* dGp, T* dacc, const RankScale& sc, hipStream_t st) {
  bool scaled = sc.maxv != nullptr;
  if (dacc) {
    if (scaled) tprop_k<T, true, true>(L, Gin, dGin, g, v, Gout, dGout, Gp, dGp, dacc, sc, st);
    else tprop_k<T, false, true>(L, Gin, dGin, g, v, Gout, dGout, Gp, dGp, dacc, sc, st);
  } else {
    if (scaled) tprop_k<T, true, false>(L, Gin, dGin, g, v, Gout, dGout, Gp, dGp, dacc, sc, st);
    else tprop_k<T, false, false>(L, Gin, dGin, g, v, Gout, dGout, Gp, dGp, dacc, sc, st);
  }
}

static __global__ void k_set_ref2(double* scal, const double* lnv, int idx, double factor) {
  if (threadIdx.x == 0 && blockIdx.x == 0) scal[MS_REF] = lnv ? factor * lnv[idx] : 0.0;
}

// Tangent sweep of molecule m with per-state directions vt (type T, ns*M).
// The base propagators are recomputed alongside in the same precision; the
// base phi_m comes from the last compute_phi at the same x.
// FP32 tangent sweep of a first-order chain on the 3-D simple-cubic lattice,
// several ranks per launch (sfx_block3.hip); the structure of tsweep below.
static void tsweep_blocked(System& S, Tangent& W, int m, const float* vt, const float* g32) {
  const DevLattice& L = S.lat.d;
  const long long Mn = L.M;
  Molecule& mo = S.mols[m];
  hipStream_t st = S.stream;
  double* scal = S.d_molscal + (size_t)m * MS_COUNT;
  double* lnf = S.d_lnf + (size_t)m * (S.maxN + 2);
  double* lnb = S.d_lnb + (size_t)m * (S.maxN + 2);
  unsigned* maxf = S.d_maxf + (size_t)m * (S.maxN + 2) * kStepWords;
  unsigned* maxb = S.d_maxb + (size_t)m * (S.maxN + 2) * kStepWords;
  float* store = (float*)S.d_store;
  float* dstore = (float*)W.dstore.p;
  float* pp[2] = {(float*)S.d_pp, (float*)S.d_pp + Mn};
  float* dpp[2] = {(float*)W.dpp.p, (float*)W.dpp.p + Mn};
  float* dacc = (float*)W.dacc.p;
  auto daccp = [&](int t) { return dacc + ((size_t)m * S.ns + t) * Mn; };
  auto gp = [&](int t) { return g32 + (size_t)t * Mn; };
  auto vp = [&](int t) { return vt + (size_t)t * Mn; };
  auto slot = [&](int i) { return store + (size_t)i * Mn; };
  auto dslot = [&](int i) { return dstore + (size_t)i * Mn; };
  const int N = mo.N;
  const ChainPlan& plan = mo.plan;
  {
    ZeroList zl;
    for (int t = 0; t < S.ns; t++)
      if (mo.uses[t]) {
        zero_list_add(zl, daccp(t), Mn, st);
        if (S.middle_out(mo)) zero_list_add(zl, (float*)W.dacc2.p + ((size_t)m * S.ns + t) * Mn, Mn, st);
      }
    zero_list_add(zl, maxf, (long long)(S.maxN + 2) * kStepWords, st);
    zero_list_add(zl, maxb, (long long)(S.maxN + 2) * kStepWords, st);
    launch_zero_list(zl, st);
  }
  auto fsc = [&](int step, int pair) { return RankScale{maxf, lnf, lnf, scal + MS_REF, step, pair}; };
  auto bsc = [&](int step, int pair) { return RankScale{maxb, lnb, lnf, scal + MS_REF, step, pair}; };
  auto state_of_rank = [&](int s) { return mo.seq[s - 1]; };
  auto nosplit = [](int) { return false; };
  // the arguments of ranks sa..sb of a chain from (in, din); with an
  // accumulator each rank is composed with the ranks in slots pslots[k],
  // whose log scale is lnpair[pidx[k]]
  auto args = [&](int sa, int sb, const float* in, const float* din, int t, unsigned* maxv, double* lnv,
                  const double* lnpair, float* const* outs, float* const* douts, const int* pslots, const int* pidx,
                  float* acc_t) {
    B3Args<float> a{};
    a.in = in;
    a.din = din;
    a.g = gp(t);
    a.v = vp(t);
    a.nr = sb - sa + 1;
    for (int k = 0; k < a.nr; k++) {
      a.out[k] = outs[k];
      a.dout[k] = douts[k];
      if (acc_t) {
        a.pair[k] = slot(pslots[k]);
        a.dpair[k] = dslot(pslots[k]);
        a.pidx[k] = pidx[k];
      }
    }
    a.acc = acc_t;
    a.maxv = maxv;
    a.lnv = lnv;
    a.lnpair = lnpair;
    a.ref = scal + MS_REF;
    a.gmax = S.d_gmax + (size_t)t * kStepWords;
    a.s0 = sa - 1;
    return a;
  };
  auto go = [&](const B3Launch<float>& la, bool connect) {
    int nr = 0;
    for (int c = 0; c < la.nchains; c++) nr += la.a[c].nr;
    const auto ev = S.prof.events(connect ? 3 : 2, SFX_FP32, nr);
    launch_block3<float>(L, la, true, true, connect, st, ev.first, ev.second);
  };
  auto launch = [&](int sa, int sb, int in_slot_or_pp, bool in_is_pp, int t, unsigned* maxv, double* lnv,
                    float* const* outs, float* const* douts, const int* pslots, const int* pidx, bool connect) {
    B3Launch<float> la{};
    la.a[0] = args(sa, sb, in_is_pp ? pp[in_slot_or_pp] : slot(in_slot_or_pp),
                   in_is_pp ? dpp[in_slot_or_pp] : dslot(in_slot_or_pp), t, maxv, lnv, lnf, outs, douts, pslots,
                   pidx, connect ? daccp(t) : nullptr);
    la.nchains = 1;
    go(la, connect);
  };
  if (S.middle_out(mo)) {
    // the schedule of System::sweep_blocked's middle-out branch, with tangents
    const int h = N / 2;
    float* dacc2 = (float*)W.dacc2.p;
    auto dacc2p = [&](int t) { return dacc2 + ((size_t)m * S.ns + t) * Mn; };
    float* qq[2] = {(float*)W.dpp2.p, (float*)W.dpp2.p + Mn};  // backward running rank
    float* dqq[2] = {(float*)W.dpp2.p + 2 * Mn, (float*)W.dpp2.p + 3 * Mn};
    // (the second accumulators were cleared with the first above)
    auto rank_of = [&](int k) { return N - k + 1; };
    auto bstate = [&](int k) { return mo.seq[rank_of(k) - 1]; };
    const int tN = mo.seq[N - 1];
    tinit2<float>(L, TInit<float>{gp(mo.seq[0]), vp(mo.seq[0]), slot(0), dslot(0), fsc(1, 0)},
                  TInit<float>{gp(tN), vp(tN), slot(N - 1), dslot(N - 1), bsc(1, N)}, st);
    const auto F1 = chain_launches(2, h, 2, state_of_rank, kB3TanRanks);
    const auto B1 = chain_launches(2, N - h, 2, bstate, kB3TanRanks);
    for (size_t i = 0; i < std::max(F1.size(), B1.size()); i++) {
      B3Launch<float> la{};
      if (i < F1.size()) {
        const int sa = F1[i].first, sb = F1[i].second;
        float* outs[kB3Ranks] = {};
        float* douts[kB3Ranks] = {};
        for (int s = sa; s <= sb; s++) {
          outs[s - sa] = slot(s - 1);
          douts[s - sa] = dslot(s - 1);
        }
        la.a[la.nchains++] = args(sa, sb, slot(sa - 2), dslot(sa - 2), mo.seq[sa - 1], maxf, lnf, nullptr, outs,
                                  douts, nullptr, nullptr, nullptr);
      }
      if (i < B1.size()) {
        const int ka = B1[i].first, kb = B1[i].second;
        float* outs[kB3Ranks] = {};
        float* douts[kB3Ranks] = {};
        for (int k = ka; k <= kb; k++) {
          outs[k - ka] = slot(rank_of(k) - 1);
          douts[k - ka] = dslot(rank_of(k) - 1);
        }
        const int in = rank_of(ka - 1) - 1;
        la.a[la.nchains++] = args(ka, kb, slot(in), dslot(in), bstate(ka), maxb, lnb, nullptr, outs, douts, nullptr,
                                  nullptr, nullptr);
      }
      go(la, false);
    }
    launch_set_ref_pair(scal, lnf, h, lnb, N - h, st);
    const auto F2 = chain_launches(h + 1, N, h + 1, state_of_rank, kB3TanRanks);
    const auto B2 = chain_launches(N - h + 1, N, N - h + 1, bstate, kB3TanRanks);
    int fc = 0, bc = 0;
    for (size_t i = 0; i < std::max(F2.size(), B2.size()); i++) {
      B3Launch<float> la{};
      if (i < F2.size()) {
        const int sa = F2[i].first, sb = F2[i].second;
        float* outs[kB3Ranks] = {};
        float* douts[kB3Ranks] = {};
        int pslots[kB3Ranks] = {}, pidx[kB3Ranks] = {};
        for (int s = sa; s <= sb; s++) {
          pslots[s - sa] = s - 1;  // backward rank s
          pidx[s - sa] = N - s + 1;
        }
        outs[sb - sa] = pp[fc ^ 1];
        douts[sb - sa] = dpp[fc ^ 1];
        const int t = mo.seq[sa - 1];
        la.a[la.nchains++] = args(sa, sb, i == 0 ? slot(h - 1) : pp[fc], i == 0 ? dslot(h - 1) : dpp[fc], t, maxf, lnf,
                                  lnb, outs, douts, pslots, pidx, daccp(t));
        fc ^= 1;
      }
      if (i < B2.size()) {
        const int ka = B2[i].first, kb = B2[i].second;
        float* outs[kB3Ranks] = {};
        float* douts[kB3Ranks] = {};
        int pslots[kB3Ranks] = {}, pidx[kB3Ranks] = {};
        for (int k = ka; k <= kb; k++) {
          pslots[k - ka] = rank_of(k) - 1;  // forward rank
          pidx[k - ka] = rank_of(k);
        }
        outs[kb - ka] = qq[bc ^ 1];
        douts[kb - ka] = dqq[bc ^ 1];
        const int t = bstate(ka);
        la.a[la.nchains++] = args(ka, kb, i == 0 ? slot(h) : qq[bc], i == 0 ? dslot(h) : dqq[bc], t, maxb, lnb, lnf,
                                  outs, douts, pslots, pidx, dacc2p(t));
        bc ^= 1;
      }
      go(la, true);
    }
    unsigned nb = grid1(L.V, kTB);
    k_wsum2<float><<<nb, kTB, 0, st>>>(L, pp[fc], dpp[fc], S.d_partial, W.partial2.p);  // forward rank N
    k_tmol_finalize<<<1, 256, 0, st>>>(mo.params(m), S.d_partial, W.partial2.p, (int)nb, S.d_molscal, W.dmol.p, S.nm);
    return;
  }
  tinit<float>(L, gp(mo.seq[0]), vp(mo.seq[0]), slot(0), dslot(0), fsc(1, 0), st);
  {
    int prev = 0, parity = 0;
    chain_blocks(2, plan.n, 2, state_of_rank, nosplit, [&](int sa, int sb) {
      float* outs[kB3Ranks] = {};
      float* douts[kB3Ranks] = {};
      int last = -1;
      for (int s = sa; s <= sb; s++) {
        int d = -1;
        if (!plan.c)
          d = s - 1;
        else if ((s - 1) % plan.c == 0)
          d = (s - 1) / plan.c;
        else if (s == sb)
          d = plan.seg(parity);
        outs[s - sa] = d >= 0 ? slot(d) : nullptr;
        douts[s - sa] = d >= 0 ? dslot(d) : nullptr;
        if (s == sb) last = d;
      }
      launch(sa, sb, prev, false, mo.seq[sa - 1], maxf, lnf, outs, douts, nullptr, nullptr, false);
      if (plan.c && last == plan.seg(parity)) parity ^= 1;
      prev = last;
    }, kB3TanRanks);
  }
  int resident[2] = {-1, -1};
  auto get = [&](int s) -> int {
    if (!plan.c) return s - 1;
    const int j = (s - 1) / plan.c, start = 1 + j * plan.c, set = j & 1;
    if (s == start) return j;
    if (resident[set] != j) {
      const int end = std::min(start + plan.c - 1, plan.n);
      int prev = j;
      chain_blocks(start + 1, end, 2, state_of_rank, nosplit, [&](int sa, int sb) {
        float* outs[kB3Ranks] = {};
        float* douts[kB3Ranks] = {};
        for (int s2 = sa; s2 <= sb; s2++) {
          outs[s2 - sa] = slot(plan.seg_in(set, s2 - start - 1));
          douts[s2 - sa] = dslot(plan.seg_in(set, s2 - start - 1));
        }
        launch(sa, sb, prev, false, mo.seq[sa - 1], maxf, lnf, outs, douts, nullptr, nullptr, false);
        prev = plan.seg_in(set, sb - start - 1);
      }, kB3TanRanks);
      resident[set] = j;
    }
    return plan.seg_in(set, s - start - 1);
  };
  int cur = 0;
  const float* fin = nullptr;
  const float* dfin = nullptr;
  if (mo.kind == 1) {
    k_set_ref2<<<1, 1, 0, st>>>(scal, lnf, N, 1.0);
    const int tN = mo.seq[N - 1];
    tinit<float>(L, gp(tN), vp(tN), pp[0], dpp[0], bsc(1, N), st);
    const int iN = get(N);
    k_tconnect<float, true><<<grid1(L.V, kTB), kTB, 0, st>>>(L, pp[0], dpp[0], slot(iN), dslot(iN), daccp(tN),
                                                             bsc(1, N), 1.0);
    auto rank_of = [&](int k) { return N - k + 1; };
    chain_blocks(
        2, N, 2, [&](int k) { return mo.seq[rank_of(k) - 1]; },
        nosplit,
        [&](int ka, int kb) {
          float* outs[kB3Ranks] = {};
          float* douts[kB3Ranks] = {};
          int pslots[kB3Ranks] = {}, pidx[kB3Ranks] = {};
          for (int k = ka; k <= kb; k++) {
            pslots[k - ka] = get(rank_of(k));
            pidx[k - ka] = rank_of(k);
          }
          outs[kb - ka] = pp[cur ^ 1];
          douts[kb - ka] = dpp[cur ^ 1];
          launch(ka, kb, cur, true, mo.seq[rank_of(ka) - 1], maxb, lnb, outs, douts, pslots, pidx, true);
          cur ^= 1;
        }, kB3TanRanks);
  } else {
    const int h = N / 2;
    k_set_ref2<<<1, 1, 0, st>>>(scal, lnf, h, 2.0);
    const int ih = get(h);
    launch_copy_interior<float>(L, slot(ih), pp[0], st);
    launch_copy_interior<float>(L, dslot(ih), dpp[0], st);
    if (N % 2 == 1) {
      const int s = h + 1, t = mo.seq[s - 1];
      tprop<float>(L, pp[0], dpp[0], gp(t), vp(t), pp[1], dpp[1], nullptr, nullptr, nullptr, fsc(s, 0), st);
      cur = 1;
      k_tconnect<float, true><<<grid1(L.V, kTB), kTB, 0, st>>>(L, pp[cur], dpp[cur], pp[cur], dpp[cur], daccp(t),
                                                               fsc(s, s), 0.5);
    }
    auto pair_of = [&](int s) { return N - s + 1; };
    chain_blocks(
        (N + 3) / 2, N, 2, state_of_rank, nosplit,
        [&](int sa, int sb) {
          float* outs[kB3Ranks] = {};
          float* douts[kB3Ranks] = {};
          int pslots[kB3Ranks] = {}, pidx[kB3Ranks] = {};
          for (int s = sa; s <= sb; s++) {
            pslots[s - sa] = get(pair_of(s));
            pidx[s - sa] = pair_of(s);
          }
          outs[sb - sa] = pp[cur ^ 1];
          douts[sb - sa] = dpp[cur ^ 1];
          launch(sa, sb, cur, true, mo.seq[sa - 1], maxf, lnf, outs, douts, pslots, pidx, true);
          cur ^= 1;
        }, kB3TanRanks);
  }
  fin = pp[cur];
  dfin = dpp[cur];
  unsigned nb = grid1(L.V, kTB);
  k_wsum2<float><<<nb, kTB, 0, st>>>(L, fin, dfin, S.d_partial, W.partial2.p);
  k_tmol_finalize<<<1, 256, 0, st>>>(mo.params(m), S.d_partial, W.partial2.p, (int)nb, S.d_molscal, W.dmol.p, S.nm);
}

template <typename T>
static void tsweep(System& S, Tangent& W, int m, bool scaled, const T* vt, const T* g32or64, const double* v64) {
  if constexpr (std::is_same<T, float>::value) {
    const int kind = S.mols[m].kind;
    if (scaled && S.use_block3 && (kind == 1 || kind == 2)) {
      tsweep_blocked(S, W, m, vt, g32or64);
      return;
    }
  }
  const DevLattice& L = S.lat.d;
  const long long Mn = L.M;
  Molecule& mo = S.mols[m];
  hipStream_t st = S.stream;
  double* scal = S.d_molscal + (size_t)m * MS_COUNT;
  double* lnf = S.d_lnf + (size_t)m * (S.maxN + 2);
  double* lnb = S.d_lnb + (size_t)m * (S.maxN + 2);
  unsigned* maxf = S.d_maxf + (size_t)m * (S.maxN + 2) * kStepWords;
  unsigned* maxb = S.d_maxb + (size_t)m * (S.maxN + 2) * kStepWords;
  T* store = (T*)S.d_store;
  T* dstore = (T*)W.dstore.p;
  T* pp[2] = {(T*)S.d_pp, (T*)S.d_pp + Mn};
  T* dpp[2] = {(T*)W.dpp.p, (T*)W.dpp.p + Mn};
  T* dacc = (T*)W.dacc.p;
  auto daccp = [&](int t) { return dacc + ((size_t)m * S.ns + t) * Mn; };
  auto gp = [&](int t) { return g32or64 + (size_t)t * Mn; };
  auto vp = [&](int t) { return vt + (size_t)t * Mn; };
  const int N = mo.N;
  const int P = sizeof(T) == 4 ? SFX_FP32 : SFX_FP64;
  for (int t = 0; t < S.ns; t++)
    if (mo.uses[t]) zero<T>(daccp(t), Mn, st);
  RankScale none{nullptr, nullptr, nullptr, nullptr, 0, 0};
  if (scaled) {
    zero<unsigned>(maxf, (S.maxN + 2) * kStepWords, st);
    zero<unsigned>(maxb, (S.maxN + 2) * kStepWords, st);
  }
  auto fsc = [&](int step, int pair) { return scaled ? RankScale{maxf, lnf, lnf, scal + MS_REF, step, pair} : none; };
  auto bsc = [&](int step, int pair) { return scaled ? RankScale{maxb, lnb, lnf, scal + MS_REF, step, pair} : none; };
  const T* fin = nullptr;
  const T* dfin = nullptr;
  double* dm = W.dmol.p;
  const ChainPlan& plan = mo.plan;
  if (mo.kind == 3) {
    const long long M6 = 6 * Mn;
    T* pp6[2] = {(T*)S.d_pp, (T*)S.d_pp + M6};
    T* dpp6[2] = {(T*)W.dpp.p, (T*)W.dpp.p + M6};
    auto slot6 = [&](int i) { return store + (size_t)i * M6; };
    auto dslot6 = [&](int i) { return dstore + (size_t)i * M6; };
    const double Sst = S.stiffness;
    // forward ranks in the W = T G form (tangent dW = T dG); the backward pass
    // recovers G_f(s), dG_f(s) from W_f(s-1), dW_f(s-1)
    launch_init2f<T>(L, true, gp(mo.seq[0]), vp(mo.seq[0]), slot6(0), dslot6(0), Sst, fsc(1, 0), st);
    auto step = [&](int s, int src, int dst) {
      S.prof.begin(2, P, st);
      launch_prop2f<T>(L, true, slot6(src), dslot6(src), gp(mo.seq[s - 1]), vp(mo.seq[s - 1]), slot6(dst),
                       dslot6(dst), Sst, fsc(s, 0), st);
      S.prof.end(st);
    };
    chain_forward(plan, step);
    ChainReader<decltype(step)> rd{plan, step};
    k_set_ref2<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, N, 1.0);
    auto bsc2 = [&](int step, int pair) {
      RankScale r = bsc(step, pair);
      if (scaled) r.pmaxv = maxf;
      return r;
    };
    int tN = mo.seq[N - 1];
    const int iN = rd.get(N - 1);
    launch_init2b<T>(L, true, gp(tN), vp(tN), pp6[0], dpp6[0], slot6(iN), dslot6(iN), daccp(tN), bsc2(1, N), st);
    for (int s = N - 1; s >= 1; s--) {
      int k = N - s + 1, t = mo.seq[s - 1];
      const T* wp = nullptr;
      const T* dwp = nullptr;
      if (s >= 2) {
        const int is = rd.get(s - 1);
        wp = slot6(is);
        dwp = dslot6(is);
      }
      S.prof.begin(3, P, st);
      launch_prop2b<T>(L, true, pp6[(k - 2) & 1], dpp6[(k - 2) & 1], gp(t), vp(t), pp6[(k - 1) & 1],
                       dpp6[(k - 1) & 1], wp, dwp, daccp(t), Sst, bsc2(k, s), st);
      S.prof.end(st);
    }
    int nb = launch_wsum6<T>(L, pp6[(N - 1) & 1], dpp6[(N - 1) & 1], S.d_partial, W.partial2.p, st);
    k_tmol_finalize<<<1, 256, 0, st>>>(mo.params(m), S.d_partial, W.partial2.p, nb, S.d_molscal, dm, S.nm);
  } else if (mo.kind == 0) {
    int t = mo.seq[0];
    unsigned nb = grid1(L.V, kTB);
    k_wsum_mono<<<nb, kTB, 0, st>>>(L, S.d_swf + (size_t)t * Mn, v64 + (size_t)t * Mn, S.d_partial, W.partial2.p);
    k_tmol_finalize<<<1, 256, 0, st>>>(mo.params(m), S.d_partial, W.partial2.p, (int)nb, S.d_molscal, dm, S.nm);
  } else {
    auto slot = [&](int i) { return store + (size_t)i * Mn; };
    auto dslot = [&](int i) { return dstore + (size_t)i * Mn; };
    tinit<T>(L, gp(mo.seq[0]), vp(mo.seq[0]), slot(0), dslot(0), fsc(1, 0), st);
    auto step = [&](int s, int src, int dst) {
      S.prof.begin(2, P, st);
      tprop<T>(L, slot(src), dslot(src), gp(mo.seq[s - 1]), vp(mo.seq[s - 1]), slot(dst), dslot(dst), nullptr,
               nullptr, nullptr, fsc(s, 0), st);
      S.prof.end(st);
    };
    chain_forward(plan, step);  // ranks 2..N (full sweep) or 2..N/2 (half store)
    ChainReader<decltype(step)> rd{plan, step};
    if (mo.kind == 1) {
      k_set_ref2<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, N, 1.0);
      int tN = mo.seq[N - 1];
      tinit<T>(L, gp(tN), vp(tN), pp[0], dpp[0], bsc(1, N), st);
      const int iN = rd.get(N);
      if (scaled)
        k_tconnect<T, true><<<grid1(L.V, kTB), kTB, 0, st>>>(L, pp[0], dpp[0], slot(iN), dslot(iN), daccp(tN),
                                                             bsc(1, N), 1.0);
      else
        k_tconnect<T, false><<<grid1(L.V, kTB), kTB, 0, st>>>(L, pp[0], dpp[0], slot(iN), dslot(iN), daccp(tN),
                                                              bsc(1, N), 1.0);
      for (int s = N - 1; s >= 1; s--) {
        int k = N - s + 1;
        int t = mo.seq[s - 1];
        const int is = rd.get(s);
        S.prof.begin(3, P, st);
        tprop<T>(L, pp[(k - 2) & 1], dpp[(k - 2) & 1], gp(t), vp(t), pp[(k - 1) & 1], dpp[(k - 1) & 1], slot(is),
                 dslot(is), daccp(t), bsc(k, s), st);
        S.prof.end(st);
      }
      fin = pp[(N - 1) & 1];
      dfin = dpp[(N - 1) & 1];
    } else {
      const int h = N / 2;
      k_set_ref2<<<1, 1, 0, st>>>(scal, scaled ? lnf : nullptr, h, 2.0);
      const int ih = rd.get(h);
      launch_copy_interior<T>(L, slot(ih), pp[0], st);
      launch_copy_interior<T>(L, dslot(ih), dpp[0], st);
      int cur = 0;
      if (N % 2 == 1) {
        int s = h + 1, t = mo.seq[s - 1];
        tprop<T>(L, pp[cur], dpp[cur], gp(t), vp(t), pp[cur ^ 1], dpp[cur ^ 1], nullptr, nullptr, nullptr, fsc(s, 0), st);
        cur ^= 1;
        RankScale c = fsc(s, s);
        if (scaled)
          k_tconnect<T, true><<<grid1(L.V, kTB), kTB, 0, st>>>(L, pp[cur], dpp[cur], pp[cur], dpp[cur], daccp(t), c, 0.5);
        else
          k_tconnect<T, false><<<grid1(L.V, kTB), kTB, 0, st>>>(L, pp[cur], dpp[cur], pp[cur], dpp[cur], daccp(t), c, 0.5);
      }
      for (int s = (N + 3) / 2; s <= N; s++) {
        int sinv = N - s + 1, t = mo.seq[s - 1];
        const int is = rd.get(sinv);
        S.prof.begin(3, P, st);
        tprop<T>(L, pp[cur], dpp[cur], gp(t), vp(t), pp[cur ^ 1], dpp[cur ^ 1], slot(is), dslot(is), daccp(t),
                 fsc(s, sinv), st);
        S.prof.end(st);
        cur ^= 1;
      }
      fin = pp[cur];
      dfin = dpp[cur];
    }
    unsigned nb = grid1(L.V, kTB);
    k_wsum2<T><<<nb, kTB, 0, st>>>(L, fin, dfin, S.d_partial, W.partial2.p);
    k_tmol_finalize<<<1, 256, 0, st>>>(mo.params(m), S.d_partial, W.partial2.p, (int)nb, S.d_molscal, dm, S.nm);
  }
}

template <typename T>
static void tmol_phi(System& S, Tangent& W, int m, const double* v64state) {
  const DevLattice& L = S.lat.d;
  const long long Mn = L.M;
  Molecule& mo = S.mols[m];
  // the FP32 blocked tangent sweep of a middle-out chain leaves half of the
  // composition in the second accumulator
  const bool two = sizeof(T) == 4 && S.middle_out(mo);
  const size_t off = (size_t)m * S.ns * Mn;
  k_tmol_phi<T><<<grid1(L.V, kTB), kTB, 0, S.stream>>>(
      L, mo.uses_mask(), S.ns, S.d_phi_m + off, (const T*)W.dacc.p + off, two ? (const T*)W.dacc2.p + off : nullptr,
      S.d_swf, v64state, W.dmol.p + m * 4, mo.kind == 0, W.dphi_m.p + off);
}

void tangent_free(Tangent* t) { delete t; }

// J v at the base state of the last compute_phi / residual at S.d_u.
void jvp_at_base(System& S, const double* v, double* jv, int precision) {
  SFX_REQUIRE(!S.has_range(), SFX_E_UNSUPPORTED,
              "Jacobian products of range_restricted molecules are not built; solve with DIIS or Newton");
  if (!S.tan) {
    S.tan = new Tangent();
    S.tan->init(S);
  }
  Tangent& W = *S.tan;
  const DevLattice& L = S.lat.d;
  const long long Mn = L.M;
  hipStream_t st = S.stream;
  double* v64 = W.v64.p;
  bool fp32 = precision == SFX_FP32;
  float* v32 = (float*)W.v32.p;
  k_state_dir<<<grid1(Mn, kTB), kTB, 0, st>>>(Mn, S.ns, S.d_st_itvar, v, v64, fp32 ? v32 : nullptr);
  if (fp32 && S.use_block3 && !S.gmax_valid) {  // the base's FP32 SWF maxima, unless its compute_phi made them
    launch_gmax(L, S.d_swf32, S.ns, S.d_gmax, st);
    S.gmax_valid = true;
  }
  for (int pass = 0; pass < 2; pass++)
    for (int m = 0; m < S.nm; m++) {
      bool solv = S.mols[m].freedom == SFX_SOLVENT;
      if ((pass == 0) == solv) continue;
      if (fp32) {
        tsweep<float>(S, W, m, true, v32, S.d_swf32, v64);
        tmol_phi<float>(S, W, m, v64);
      } else {
        tsweep<double>(S, W, m, false, v64, S.d_swf, v64);
        tmol_phi<double>(S, W, m, v64);
      }
    }
  k_tphi_total<<<grid1(Mn, kTB), kTB, 0, st>>>(L, S.ns, S.nm, W.dphi_m.p, S.d_mol_uses, S.d_frozen, W.dphi.p,
                                               W.dphitot.p);
  k_tsides<<<grid1(L.V, kTB), kTB, 0, st>>>(L, S.ns, W.dphi.p, W.drho.p);
  k_jres<<<grid1(Mn, kTB), kTB, 0, st>>>(L, S.sp, S.d_itmask, v, S.d_rho, W.drho.p, S.d_phitot, W.dphitot.p,
                                         S.d_molscal, W.dmol.p, jv);
  SFX_HIP(hipGetLastError());
}

void jvp_internal(System& S, const double* v, double* jv, int precision) {
  S.compute_phi(precision);
  jvp_at_base(S, v, jv, precision);
}

// ---------------------------------------------------------------------------
// Transposed products J^T w. The composition of a chain is the derivative of
// its single-chain partition function, phi_m,t(r) = -(pref / D(r)) dQ/du_t(r),
// wherever the backward sweep is the adjoint of the forward sweep in the
// site-volume inner product D (D = the ComputeLnGN weight of a site: 1, 1/2
// on a mirror2 boundary layer, L(z) on curved lattices). The tangent sweep's
// map T: v -> dacc / g then satisfies T^T(y) = D T(y / D) (the Hessian of Q
// is symmetric), so the transposed sweep is the tangent sweep run with the
// direction y / D. The contact average is symmetric in the same inner product
// (Lat1DFlat.cpp:510-524, Lat1DCylinder.cpp:160-171, Lat3D.cpp:1729-1784), so
// its transpose is D <. / D>. Everything else of the residual
// (SF_Solve.cpp:554-590), the normalisations and the solvent closure is
// transposed term by term, in the reverse order of the forward product.

// true where the backward sweep is the forward sweep's adjoint in D. Not so:
// the 2-D cylinder's second-order sweep (its outward step is not split, as in
// the reference); a curved lattice with a mirror2 side; a 1-D flat lattice
// with a layer adjustment other than 1; a mirror on a HEX lattice (its
// neighbour set is not symmetric under reflection, so the mirrored stencil is
// not); a second-order 1-D or 2-D lattice with a mirror2 lower bound (the
// reference copies the components there without reversing them).
bool vjp_supported(const System& S) {
  const DevLattice& L = S.lat.d;
  if (L.curved) {
    for (int k = 0; k < L.grad; k++)
      if (L.mirror2_1[k] || L.mirror2_2[k]) return false;
    if (L.grad == 2 && L.order == 2) return false;
  }
  if (L.grad == 1 && !L.curved && L.ladj != 1.0) return false;
  if (L.kind == ST_3D_HEX)
    for (int k = 0; k < 3; k++)
      if (!L.periodic[k]) return false;
  if (L.order == 2 && L.grad < 3)
    for (int k = 0; k < L.grad; k++)
      if (L.mirror2_1[k]) return false;
  return true;
}

// per site: w-hat = w - mean_k w (the transposed mean subtraction); the direct
// part of J^T w (w-hat); the adjoint of every state's contact average; the
// adjoint of phi_total; and per molecule the workgroup's part of the adjoint of
// its bulk fraction through ChemIntBulk
__global__ void k_vjp_res(DevLattice L, SysParams sp, const unsigned char* __restrict__ itmask,
                          const double* __restrict__ w, const double* __restrict__ rho,
                          const double* __restrict__ phitot, double* __restrict__ jtw, double* __restrict__ rbar,
                          double* __restrict__ ptbar, double* __restrict__ part) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nv = sp.n_itvar;
  double cb[SFX_MAXM];
  for (int m = 0; m < SFX_MAXM; m++) cb[m] = 0.0;
  if (i < L.M) {
    if (!itmask[i]) {
      for (int k = 0; k < nv; k++) jtw[(long long)k * L.M + i] = 0.0;
      for (int j = 0; j < sp.ns; j++) rbar[(long long)j * L.M + i] = 0.0;
      ptbar[i] = 0.0;
    } else {
      const double pt = phitot[i];
      double wsum = 0.0;
      for (int k = 0; k < nv; k++) wsum += w[(long long)k * L.M + i];
      const double mean = wsum / nv;
      double pbar = -wsum / (pt * pt);
      for (int j = 0; j < sp.ns; j++) rbar[(long long)j * L.M + i] = 0.0;
      for (int k = 0; k < nv; k++) {
        const double wh = w[(long long)k * L.M + i] - mean;
        jtw[(long long)k * L.M + i] = wh;
        const int t = sp.itvar_rep[k];
        double ci = 0.0;
        for (int j = 0; j < sp.ns; j++) {
          const double c = sp.chi[t * SFX_MAXS + j];
          if (c == 0) continue;
          ci += c * rho[(long long)j * L.M + i];
          rbar[(long long)j * L.M + i] -= wh * c / pt;
          if (!sp.melt && !sp.frozen[j])
            for (int m = 0; m < sp.nm; m++) cb[m] += wh * c * sp.mol_frac[m * SFX_MAXS + j];
        }
        pbar += wh * ci / (pt * pt);
      }
      ptbar[i] = pbar;
    }
  }
  __shared__ double red[SFX_MAXM][kTB / 64];
  for (int m = 0; m < sp.nm; m++) {
    double v = cb[m];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[m][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)sp.nm) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += red[threadIdx.x][k];
    part[(long long)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// q_t = rbar_t / D on the interior, halos from their sources (0 behind a
// surface), frozen states 0: the input of the transposed contact average
__global__ void k_vjp_q(DevLattice L, int ns, const int* __restrict__ frozen, const double* __restrict__ rbar,
                        double* __restrict__ q) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  const int sx = remap(x, L.nx, L.b1[0], L.b2[0]);
  const int sy = L.grad >= 2 ? remap(y, L.ny, L.b1[1], L.b2[1]) : y;
  const int sz = L.grad >= 3 ? remap(z, L.nz, L.b1[2], L.b2[2]) : z;
  const bool surf = !is_interior(L, sx, sy, sz);
  const long long s = sidx(L, sx, sy, sz);
  const double d = surf ? 1.0 : lngn_weight(L, sx, sy, sz);
  for (int t = 0; t < ns; t++)
    q[(long long)t * L.M + i] = (frozen[t] || surf) ? 0.0 : rbar[(long long)t * L.M + s] / d;
}

// phibar_t = D <q_t> + ptbar on the interior (non-frozen states), 0 elsewhere
__global__ void k_vjp_phibar(DevLattice L, int ns, const int* __restrict__ frozen, const double* __restrict__ q,
                             const double* __restrict__ ptbar, double* __restrict__ phibar) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  const bool in = is_interior(L, x, y, z);
  const double d = in ? lngn_weight(L, x, y, z) : 0.0;
  for (int t = 0; t < ns; t++)
    phibar[(long long)t * L.M + i] =
        (in && !frozen[t]) ? d * side_at(L, q + (long long)t * L.M, x, y, z) + ptbar[i] : 0.0;
}

// per-block sums over the interior of sum_{t in uses} a_t b_t (with D when weighted)
__global__ void k_vjp_dot(DevLattice L, unsigned uses, int ns, const double* __restrict__ a,
                          const double* __restrict__ b, int weighted, double* __restrict__ part) {
  const long long q = site_ordinal();
  double v = 0.0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    const long long r = sidx(L, x, y, z);
    for (int t = 0; t < ns; t++)
      if ((uses >> t) & 1u) v += a[(long long)t * L.M + r] * (b ? b[(long long)t * L.M + r] : 1.0);
    if (weighted) v *= lngn_weight(L, x, y, z);
  }
  __shared__ double red[kTB / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += red[k];
    part[blockIdx.x] = s;
  }
}

// vbar_t += a phi_m,t phibar_t + c D phi_m,t on the interior (states in uses)
__global__ void k_vjp_direct(DevLattice L, unsigned uses, int ns, const double* __restrict__ phi_m,
                             const double* __restrict__ phibar, double a, double c, double* __restrict__ vbar) {
  const long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r0 = sidx(L, x, y, z);
  const double d = lngn_weight(L, x, y, z);
  for (int t = 0; t < ns; t++) {
    if (!((uses >> t) & 1u)) continue;
    const long long r = (long long)t * L.M + r0;
    vbar[r] += phi_m[r] * (a * phibar[r] + c * d);
  }
}

// the sweep direction x_t = phibar_t / D on the interior for t in uses, else 0
__global__ void k_vjp_dir(DevLattice L, unsigned uses, int ns, const double* __restrict__ phibar,
                          double* __restrict__ x64, float* __restrict__ x32) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  const bool in = is_interior(L, x, y, z);
  const double d = in ? lngn_weight(L, x, y, z) : 1.0;
  for (int t = 0; t < ns; t++) {
    const double v = (in && ((uses >> t) & 1u)) ? phibar[(long long)t * L.M + i] / d : 0.0;
    x64[(long long)t * L.M + i] = v;
    x32[(long long)t * L.M + i] = (float)v;
  }
}

// vbar_t += D dm[3] (dacc_t + dacc2_t) / g_t: the transposed sweep's share
template <typename T>
__global__ void k_vjp_sweep_term(DevLattice L, unsigned uses, int ns, const T* __restrict__ dacc,
                                 const T* __restrict__ dacc2, const double* __restrict__ g,
                                 const double* __restrict__ dm, double* __restrict__ vbar) {
  const long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r0 = sidx(L, x, y, z);
  const double d = lngn_weight(L, x, y, z);
  for (int t = 0; t < ns; t++) {
    if (!((uses >> t) & 1u)) continue;
    const long long r = (long long)t * L.M + r0;
    const double gv = g[r];
    if (gv > 0) vbar[r] += d * dm[3] * ((double)dacc[r] + (dacc2 ? (double)dacc2[r] : 0.0)) / gv;
  }
}

// J^T w = w-hat + sum over the states of each iteration variable of vbar, on
// the iteration sites (the transpose of k_state_dir)
__global__ void k_vjp_final(long long M, int nv, int ns, const int* __restrict__ st_itvar,
                            const unsigned char* __restrict__ itmask, const double* __restrict__ vbar,
                            double* __restrict__ jtw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  if (!itmask[i]) return;  // k_vjp_res zeroed these
  for (int t = 0; t < ns; t++) {
    const int k = st_itvar[t];
    if (k >= 0) jtw[(long long)k * M + i] += vbar[(long long)t * M + i];
  }
}

static double host_sum(const double* part, int n, hipStream_t st) {
  std::vector<double> h(n);
  SFX_HIP(hipMemcpyAsync(h.data(), part, n * sizeof(double), hipMemcpyDeviceToHost, st));
  SFX_HIP(hipStreamSynchronize(st));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

// J^T w at the base state of the last compute_phi at S.d_u
void vjp_at_base(System& S, const double* w, double* jtw, int precision) {
  SFX_REQUIRE(!S.has_range(), SFX_E_UNSUPPORTED,
              "Jacobian products of range_restricted molecules are not built; solve with DIIS or Newton");
  SFX_REQUIRE(vjp_supported(S), SFX_E_UNSUPPORTED,
              "transposed products need a backward sweep that is the forward sweep's adjoint: not on a curved "
              "lattice with a mirror2 side, the second-order 2-D cylinder, a layer adjustment other than 1, a "
              "mirror on a HEX lattice or a second-order 1-D/2-D mirror2 lower bound");
  if (!S.tan) {
    S.tan = new Tangent();
    S.tan->init(S);
  }
  Tangent& W = *S.tan;
  if (!W.vjp_ready) W.init_vjp(S);
  const DevLattice& L = S.lat.d;
  const long long Mn = L.M;
  hipStream_t st = S.stream;
  const bool fp32 = precision == SFX_FP32;
  const unsigned nbM = grid1(Mn, kTB), nbV = grid1(L.V, kTB);
  // residual: the direct part, the contact averages' and the total's adjoints,
  // and each molecule's bulk-fraction adjoint through ChemIntBulk
  k_vjp_res<<<nbM, kTB, 0, st>>>(L, S.sp, S.d_itmask, w, S.d_rho, S.d_phitot, jtw, W.rbar.p, W.ptbar.p, W.vpart.p);
  std::vector<double> pbbar(S.nm, 0.0);
  for (int m = 0; m < S.nm; m++) pbbar[m] = host_sum(W.vpart.p + (size_t)m * nbM, (int)nbM, st);
  // the contact averages and phi_total, transposed
  k_vjp_q<<<nbM, kTB, 0, st>>>(L, S.ns, S.d_frozen, W.rbar.p, W.xdir64.p);
  k_vjp_phibar<<<nbM, kTB, 0, st>>>(L, S.ns, S.d_frozen, W.xdir64.p, W.ptbar.p, W.phibar.p);
  SFX_HIP(hipMemsetAsync(W.vbar.p, 0, (size_t)S.ns * Mn * sizeof(double), st));
  std::vector<double> molscal((size_t)S.nm * MS_COUNT);
  SFX_HIP(hipMemcpyAsync(molscal.data(), S.d_molscal, molscal.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SFX_HIP(hipStreamSynchronize(st));
  if (fp32 && S.use_block3 && !S.gmax_valid) {
    launch_gmax(L, S.d_swf32, S.ns, S.d_gmax, st);
    S.gmax_valid = true;
  }
  // molecules in the reverse order of the forward product: the solvent first
  // (its bulk fraction closes on the others'), then the rest
  for (int pass = 0; pass < 2; pass++)
    for (int m = 0; m < S.nm; m++) {
      const Molecule& mo = S.mols[m];
      const bool solv = mo.freedom == SFX_SOLVENT;
      if ((pass == 0) != solv) continue;
      const unsigned uses = mo.uses_mask();
      const double* phi_m = S.d_phi_m + (size_t)m * S.ns * Mn;
      k_vjp_dot<<<nbV, kTB, 0, st>>>(L, uses, S.ns, phi_m, W.phibar.p, 0, W.vpart.p);
      const double relbar = host_sum(W.vpart.p, (int)nbV, st);
      const double phib = molscal[(size_t)m * MS_COUNT + MS_PHIB];
      const bool monomer = mo.kind == 0;
      double lnbar = 0.0;
      if (solv) {
        // rel = dphib / phib, dphib = -sum_{m' != solvent} dphib_m'
        const double dpb = (phib != 0 ? relbar / phib : 0.0) + pbbar[m];
        for (int o = 0; o < S.nm; o++)
          if (o != m) pbbar[o] -= dpb;
      } else if (mo.freedom == SFX_RESTRICTED) {
        // rel = -dlnGN, dphib = -phib dlnGN
        lnbar = -relbar - phib * pbbar[m];
      }
      // dlnGN = -(N / theta) sum_t,r D phi_m,t v_t, theta = sum D phi_m
      double coef = 0.0;
      if (lnbar != 0.0) {
        k_vjp_dot<<<nbV, kTB, 0, st>>>(L, uses, S.ns, phi_m, nullptr, 1, W.vpart.p);
        const double theta = host_sum(W.vpart.p, (int)nbV, st);
        coef = -lnbar * mo.N / theta;
      }
      k_vjp_direct<<<nbV, kTB, 0, st>>>(L, uses, S.ns, phi_m, W.phibar.p, monomer ? -1.0 : 1.0, coef, W.vbar.p);
      if (monomer) continue;
      // the transposed sweep: the tangent sweep with direction phibar / D
      k_vjp_dir<<<nbM, kTB, 0, st>>>(L, uses, S.ns, W.phibar.p, W.xdir64.p, (float*)W.xdir32.p);
      const size_t off = (size_t)m * S.ns * Mn;
      if (fp32) {
        tsweep<float>(S, W, m, true, (const float*)W.xdir32.p, S.d_swf32, W.xdir64.p);
        const bool two = S.middle_out(mo);
        k_vjp_sweep_term<float><<<nbV, kTB, 0, st>>>(L, uses, S.ns, (const float*)W.dacc.p + off,
                                                     two ? (const float*)W.dacc2.p + off : nullptr, S.d_swf,
                                                     W.dmol.p + m * 4, W.vbar.p);
      } else {
        tsweep<double>(S, W, m, false, W.xdir64.p, S.d_swf, W.xdir64.p);
        k_vjp_sweep_term<double><<<nbV, kTB, 0, st>>>(L, uses, S.ns, (const double*)W.dacc.p + off, nullptr,
                                                      S.d_swf, W.dmol.p + m * 4, W.vbar.p);
      }
    }
  k_vjp_final<<<nbM, kTB, 0, st>>>(Mn, S.n_itvar, S.ns, S.d_st_itvar, S.d_itmask, W.vbar.p, jtw);
  SFX_HIP(hipGetLastError());
}

void vjp_internal(System& S, const double* w, double* jtw, int precision) {
  S.compute_phi(precision);
  vjp_at_base(S, w, jtw, precision);
}

// ABI-layout vector (host or device) -> internal [itvar][M] (zero off sites)
void load_vec(System& S, const double* x, int where, double* dst) {
  // reuse load_x through a temporary swap of d_u
  double* save = S.d_u;
  S.d_u = dst;
  try {
    S.load_x(x, where);
  } catch (...) {
    S.d_u = save;
    throw;
  }
  S.d_u = save;
}

}  // namespace sfx
