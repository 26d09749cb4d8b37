// sfx_kernels.hip -- gfx950 kernels of the SCF hot path.
//
// One thread per interior lattice site. Propagator arrays keep the padded
// reference layout but only interior sites are ever written or read (neighbour
// coordinates are remapped, see sfx_device.h), so halo slots cost memory only.
#include "sfx_device.h"
#include <algorithm>

#include "sfx_kernels.h"

namespace sfx {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// block-level reductions (wave64)
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ float block_max(float v) {
  __shared__ float sm[kBlock / 64];
  v = wave_max(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (w == 0) {
    v = lane < (int)(blockDim.x >> 6) ? sm[lane] : 0.f;
    v = wave_max(v);
  }
  return v;
}
// deterministic: fixed tree inside the block
__device__ double block_sum(double v) {
  __shared__ double sm[kBlock / 64];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sm[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sm[i];
  __syncthreads();
  return r;  // valid in thread 0
}

// ---------------------------------------------------------------------------
// SWF: g_t(r) = exp(-u_{itvar(t)}(src(r))) at iteration sites, 0 elsewhere,
// halos from their source site (SF_Solve.cpp:794-852 + SetBoundaries).
__global__ void k_swf(DevLattice L, const double* __restrict__ u, const unsigned char* __restrict__ itmask,
                      int nstates, const int* __restrict__ st_itvar, double* __restrict__ swf,
                      float* __restrict__ swf32, unsigned* __restrict__ zero, int nzero) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nzero) zero[i] = 0u;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  int sx = remap(x, L.nx, L.b1[0], L.b2[0]);
  int sy = L.grad >= 2 ? remap(y, L.ny, L.b1[1], L.b2[1]) : y;
  int sz = L.grad >= 3 ? remap(z, L.nz, L.b1[2], L.b2[2]) : z;
  long long s = sidx(L, sx, sy, sz);
  bool it = itmask[s];
  for (int t = 0; t < nstates; t++) {
    int iv = st_itvar[t];
    double g = (iv >= 0 && it) ? exp(-u[(long long)iv * L.M + s]) : 0.0;
    swf[(long long)t * L.M + i] = g;
    if (swf32) swf32[(long long)t * L.M + i] = (float)g;
  }
}

// ---------------------------------------------------------------------------
// rank 1: G(1) = g (interior), optional running max for the FP32 log-scale
// (blockIdx.y picks the chain when one launch seeds two)
template <typename T>
__global__ void k_init_rank(DevLattice L, InitRank<T> c0, InitRank<T> c1) {
  const InitRank<T>& c = blockIdx.y ? c1 : c0;
  const T* __restrict__ g = c.g;
  T* __restrict__ out = c.out;
  const RankScale& sc = c.sc;
  long long q = site_ordinal();
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    T v = g[r];
    out[r] = v;
    m = (float)v;
  }
  if (sc.maxv) {
    m = block_max(m);
    if (threadIdx.x == 0) {
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = 0.0;
    }
  }
}

// scale factor of step `step`: out values are multiplied by 1/max(previous step)
__device__ __forceinline__ float step_scale(const RankScale& sc, double& lncum) {
  const double l0 = sc.lnv[sc.step - 1];  // issued before the wait on the maxima
  float mprev, unused;
  block_step_max2(sc.maxv, sc.step - 1, nullptr, 0, mprev, unused);
  if (!(mprev > 0.f)) mprev = 1.f;  // all-zero propagator: keep it unscaled
  lncum = l0 + log((double)mprev);
  return 1.f / mprev;
}

// Gout = g * stencil(Gin) [ * scale ]; optional composition acc += Gout * Gpair * cfac
template <typename T, typename A, int KIND, bool SCALED, bool CONNECT>
__global__ void __launch_bounds__(kBlock) k_prop(DevLattice L, const T* __restrict__ Gin,
                                                 const T* __restrict__ g, T* __restrict__ Gout,
                                                 const T* __restrict__ Gpair, A* __restrict__ acc,
                                                 RankScale sc) {
  long long q = site_ordinal();
  const bool in = q < L.V;
  // this launch's loads first: the step scale below waits on the previous
  // launch's maxima while they are in flight
  long long r = 0;
  T v = T(0), gpv = T(0);
  A a0 = A(0);
  if (in) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    r = sidx(L, x, y, z);
    v = g[r] * stencil<T, KIND>(L, Gin, x, y, z);
    if constexpr (CONNECT) {
      gpv = Gpair[r];
      a0 = acc[r];
    }
  }
  T scale = T(1);
  double lncum = 0;
  T cfac = T(1);
  if constexpr (SCALED) {
    double lp = 0.0, rf = 0.0;
    if constexpr (CONNECT) {
      lp = sc.lnpair[sc.pair];
      rf = *sc.ref;
    }
    scale = (T)step_scale(sc, lncum);
    if constexpr (CONNECT) cfac = (T)exp(lncum + lp - rf);
  }
  float m = 0.f;
  if (in) {
    if constexpr (SCALED) v *= scale;
    Gout[r] = v;
    if constexpr (CONNECT) acc[r] = a0 + (A)(v * gpv * cfac);
    m = (float)v;
  }
  if constexpr (SCALED) {
    m = block_max(m);
    if (threadIdx.x == 0) {
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = lncum;
    }
  }
}

// acc += A * B * cfac at interior sites (ConnectG, Lat1DFlat1stO.cpp:259-265)
template <typename T, typename A, bool SCALED>
__global__ void k_connect(DevLattice L, const T* __restrict__ Ga, const T* __restrict__ Gb,
                          A* __restrict__ acc, RankScale sc, double half) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  T cfac = T(half);
  if constexpr (SCALED) cfac = (T)(half * exp(sc.lnv[sc.step] + sc.lnpair[sc.pair] - *sc.ref));
  int x, y, z;
  interior_coords(L, q, x, y, z);
  long long r = sidx(L, x, y, z);
  acc[r] += (A)(Ga[r] * Gb[r] * cfac);
}

template <typename T>
__global__ void k_copy_interior(DevLattice L, const T* __restrict__ a, T* __restrict__ b) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  long long r = sidx(L, x, y, z);
  b[r] = a[r];
}

// weighted interior sum of G (ComputeLnGN before the log), per-block partials
template <typename T>
__global__ void k_wsum(DevLattice L, const T* __restrict__ G, double* __restrict__ partial) {
  long long q = site_ordinal();
  double v = 0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    v = (double)G[sidx(L, x, y, z)] * lngn_weight(L, x, y, z);
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// ---------------------------------------------------------------------------
// Molecule scalars after a sweep (one thread). Normalisation per freedom:
// SF_Monomer1stO.cpp:110-130, SF_Homopol1stO.cpp:599-619, SF_Copol1stO.cpp:1242-1275.
__global__ void k_mol_finalize(MolParams mp, const double* __restrict__ partial, int nparts,
                               double* __restrict__ scal, const double* __restrict__ lnfN,
                               const double* __restrict__ allscal, int nmol) {
  // the partial sums, reduced by the whole block in a fixed order
  __shared__ double red[256 / 64];
  double s = 0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
  double lnsum = mp.kind == 3 ? log(s / 6.0) : log(s);  // second order: Lat3D2ndO.cpp:1024
  double lnscale = lnfN ? *lnfN : 0.0;  // FP32: log-scale of the final rank
  double lnGN = lnsum + lnscale;
  double ref = scal[MS_REF];
  double N = mp.N;
  double phib = mp.phib, lnC, lnpref;
  if (mp.freedom == SFX_SOLVENT) {
    phib = 1.0;
    for (int m = 0; m < nmol; m++)
      if (m != mp.index) phib -= allscal[m * MS_COUNT + MS_PHIB];
  }
  // half-stored chains (SymMatrix1 factor 2); second order divides by 6 (Lat3D2ndO.cpp:1037)
  double f = mp.kind == 2 ? 2.0 : mp.kind == 3 ? 1.0 / 6.0 : 1.0;
  double pref;
  if (mp.freedom == SFX_RANGE_RESTRICTED) {
    // the profile before the range condition fixes its constant (k_range_finalize)
    lnC = 0.0;
    phib = 0.0;
    pref = f * exp(ref);
  } else if (mp.freedom == SFX_RESTRICTED) {
    lnC = log(mp.theta) - lnGN - (mp.kind == 0 ? 0.0 : log(N));
    phib = mp.theta / exp(lnGN);
    lnpref = lnC;
    pref = f * exp(lnpref + ref);
  } else {
    // phi = (phi_b / N) G G_inv / g: phi_b multiplies directly, since the
    // solvent closure can make it negative (NormPhiFree, Lat1DFlat1stO.cpp:317-331)
    const double C = mp.kind == 0 ? phib : phib / N;
    lnC = log(C);  // NaN for phi_b <= 0, as ln of the reference's lnCb; used by the free energy only
    pref = f * C * exp(ref);
  }
  scal[MS_LNGN] = lnGN;
  scal[MS_LNC] = lnC;
  scal[MS_PHIB] = phib;
  scal[MS_PREF] = pref;
}

// phi_{m,t} = pref_m * acc_{m,t} / g_t (g>0, else 0); monomers phi = pref * g.
template <typename A>
__global__ void k_mol_phi(DevLattice L, const A* __restrict__ acc, const A* __restrict__ acc2,
                          const double* __restrict__ g, const double* __restrict__ scal, int monomer,
                          double* __restrict__ phi_m) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  long long r = sidx(L, x, y, z);
  double pref = scal[MS_PREF];
  double gv = g[r];
  double v;
  if (monomer)
    v = pref * gv;
  else
    v = gv > 0 ? (((double)acc[r] + (acc2 ? (double)acc2[r] : 0.0)) / gv) * pref : 0.0;
  phi_m[r] = v;
}

__global__ void k_set_ref_pair(double* scal, const double* lna, int ia, const double* lnb, int ib) {
  if (threadIdx.x == 0 && blockIdx.x == 0) scal[MS_REF] = lna ? lna[ia] + lnb[ib] : 0.0;
}

void launch_set_ref_pair(double* scal, const double* lna, int ia, const double* lnb, int ib, hipStream_t st) {
  k_set_ref_pair<<<1, 1, 0, st>>>(scal, lna, ia, lnb, ib);
}

// state totals over molecules + frozen profiles, halos from their sources;
// phitot on interior sites.
__device__ void bulk_fractions(const SysParams& sp, const double* __restrict__ molscal, double* __restrict__ phib_state,
                               double* __restrict__ cbulk);

// state totals; the first thread also derives the bulk fractions and ChemIntBulk
// from the molecules' phi_b (no dependence on the sites)
__global__ void k_phi_total(DevLattice L, int nstates, int nmol, const double* __restrict__ phi_m,
                            const int* __restrict__ mol_uses, const double* __restrict__ fprof,
                            const int* __restrict__ frozen, double* __restrict__ phi,
                            double* __restrict__ phitot, SysParams sp, const double* __restrict__ molscal,
                            double* __restrict__ phib_state, double* __restrict__ cbulk) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && phib_state) bulk_fractions(sp, molscal, phib_state, cbulk);
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
    if (frozen[t]) {
      v = fprof[(long long)t * L.M + s];  // PutSides applies SetBoundaries to it as well
    } else if (!surf_halo) {
      for (int m = 0; m < nmol; m++)
        if (mol_uses[m * nstates + t]) v += phi_m[((long long)m * nstates + t) * L.M + s];
    }
    phi[(long long)t * L.M + i] = v;
    tot += v;
  }
  phitot[i] = tot;
}

// ---------------------------------------------------------------------------
__global__ void k_sides(DevLattice L, int nstates, const double* __restrict__ phi, double* __restrict__ rho) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  bool inside = is_interior(L, x, y, z);
  bool surf = false;
  if (!inside) {
    int c[3] = {x, y, z}, n[3] = {L.nx, L.ny, L.nz};
    for (int k = 0; k < L.grad; k++)
      if ((c[k] == 0 && L.surf1[k]) || (c[k] == n[k] + 1 && L.surf2[k])) surf = true;
  }
  for (int t = 0; t < nstates; t++) {
    const double* P = phi + (long long)t * L.M;
    rho[(long long)t * L.M + i] = (inside || surf) ? side_at(L, P, x, y, z) : 0.0;
  }
}

// ---------------------------------------------------------------------------
// segment bulk fractions and ChemIntBulk (SF_MolStructure.cpp:605-612,
// SF_SegmentList.cpp:610-623); one thread
__device__ void bulk_fractions(const SysParams& sp, const double* __restrict__ molscal, double* __restrict__ phib_state,
                               double* __restrict__ cbulk) {
  for (int t = 0; t < sp.ns; t++) {
    double pb = 0;
    if (sp.melt)
      pb = sp.melt_phib[t];
    else if (!sp.frozen[t])
      for (int m = 0; m < sp.nm; m++) pb += molscal[m * MS_COUNT + MS_PHIB] * sp.mol_frac[m * SFX_MAXS + t];
    phib_state[t] = pb;
  }
  for (int t = 0; t < sp.ns; t++) {
    double c = 0;
    if (!sp.frozen[t])
      for (int j = 0; j < sp.ns; j++) {
        double x = sp.chi[t * SFX_MAXS + j];
        if (x != 0) c += x * phib_state[j];
      }
    cbulk[t] = c;
  }
}

// f = -1 + 1/phitot + u'_i - <u'> per iteration site (SF_Solve.cpp:554-590),
// written in the internal [itvar][M] layout (zero off iteration sites).
__global__ void k_residual(DevLattice L, SysParams sp, const unsigned char* __restrict__ itmask,
                           const double* __restrict__ swf, const double* __restrict__ rho,
                           const double* __restrict__ phitot, const double* __restrict__ cbulk,
                           double* __restrict__ f) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  const int nv = sp.n_itvar;
  if (!itmask[i]) {
    for (int k = 0; k < nv; k++) f[(long long)k * L.M + i] = 0.0;
    return;
  }
  double up[SFX_MAXS];
  double pt = phitot[i];
  double av = 0;
  for (int k = 0; k < nv; k++) {
    int t = sp.itvar_rep[k];
    double ci = 0;
    for (int j = 0; j < sp.ns; j++) {
      double x = sp.chi[t * SFX_MAXS + j];
      if (x != 0) ci += x * rho[(long long)j * L.M + i];
    }
    double v = -log(swf[(long long)t * L.M + i]);
    v -= ci / pt;
    v += cbulk[t];
    up[k] = v;
    av += v;
  }
  av /= nv;
  for (int k = 0; k < nv; k++) f[(long long)k * L.M + i] = -1.0 + 1.0 / pt + up[k] - av;
}

// per-site free-energy density, block partials (SF_MolList1stO.cpp:815-940)
__global__ void k_free_energy(DevLattice L, SysParams sp, const double* __restrict__ phi,
                              const double* __restrict__ phi_m, const double* __restrict__ rho,
                              const double* __restrict__ swf, const double* __restrict__ u,
                              const double* __restrict__ molscal, double* __restrict__ partial) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (i < L.M) {
    int x, y, z;
    site_coords(L, i, x, y, z);
    const int ns = sp.ns;
    if (sp.melt) {
      if (is_interior(L, x, y, z)) {
        for (int a = 0; a < ns; a++)
          for (int b = a + 1; b < ns; b++) {
            double c = sp.chi[a * SFX_MAXS + b];
            if (c != 0) v += c * rho[(long long)a * L.M + i] * phi[(long long)b * L.M + i];
          }
        for (int t = 0; t < ns; t++)
          if (!sp.frozen[t]) v -= phi[(long long)t * L.M + i] * u[(long long)sp.st_itvar[t] * L.M + i];
      }
    } else {
      double w = fe_weight(L, x, y, z);
      if (w != 0) {
        int sx = remap(x, L.nx, L.b1[0], L.b2[0]);
        int sy = L.grad >= 2 ? remap(y, L.ny, L.b1[1], L.b2[1]) : y;
        int sz = L.grad >= 3 ? remap(z, L.nz, L.b1[2], L.b2[2]) : z;
        long long s = sidx(L, sx, sy, sz);
        bool liquid_site = is_interior(L, sx, sy, sz);
        double ent = 0, con = 0;
        for (int m = 0; m < sp.nm; m++) {
          double pm = 0;
          if (liquid_site)
            for (int t = 0; t < ns; t++)
              if (sp.mol_frac[m * SFX_MAXS + t] > 0) pm += phi_m[((long long)m * ns + t) * L.M + s];
          double N = sp.mol_N[m];
          ent -= (pm / N) * (log(N) + molscal[m * MS_COUNT + MS_LNC]);
          con -= pm * sp.mol_ref[m];
          if (liquid_site)
            for (int t = 0; t < ns; t++) {
              if (sp.mol_frac[m * SFX_MAXS + t] <= 0) continue;
              double G = swf[(long long)t * L.M + s];
              if (G > 0) ent -= phi_m[((long long)m * ns + t) * L.M + s] * log(G);
            }
        }
        for (int t = 0; t < ns; t++) {
          double ci = 0;
          for (int j = 0; j < ns; j++) {
            double c = sp.chi[t * SFX_MAXS + j];
            if (c != 0) ci += c * rho[(long long)j * L.M + i];
          }
          con += 0.5 * phi[(long long)t * L.M + i] * ci;
        }
        v = (-ent + con) * w;
      }
    }
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// Grand potential (SF form): the profile of SF_MolList1stO.cpp:1001-1100
// without charges, grafted chains or external potentials, weighted as the free
// energy (SubtractBoundaries, MultiplyWithLatticeSites), summed per block.
//   omega = -sum_mol [phi_mol > 0] (phi_mol - phi_b,mol) / N
//           - sum_t phi_t u'_t,  u'_t = -ln g_t - ChemInt_t + ChemIntBulk_t (g_t > 0)
//           - 1/2 sum_tj chi_tj phi_t <phi_j> + 1/2 sum_tj [phi_t > 0] chi_tj phi_b,t phi_b,j
// (states t, j not frozen in the last two sums).
__global__ void k_grand_potential(DevLattice L, SysParams sp, const double* __restrict__ phi,
                                  const double* __restrict__ phi_m, const double* __restrict__ rho,
                                  const double* __restrict__ swf, const double* __restrict__ molscal,
                                  const double* __restrict__ phib_state, const double* __restrict__ cbulk,
                                  double* __restrict__ partial) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (i < L.M) {
    int x, y, z;
    site_coords(L, i, x, y, z);
    const double w = fe_weight(L, x, y, z);
    if (w != 0) {
      const int ns = sp.ns;
      int sx = remap(x, L.nx, L.b1[0], L.b2[0]);
      int sy = L.grad >= 2 ? remap(y, L.ny, L.b1[1], L.b2[1]) : y;
      int sz = L.grad >= 3 ? remap(z, L.nz, L.b1[2], L.b2[2]) : z;
      const long long s = sidx(L, sx, sy, sz);
      const bool liquid_site = is_interior(L, sx, sy, sz);
      double om = 0;
      for (int m = 0; m < sp.nm; m++) {
        double pm = 0;
        if (liquid_site)
          for (int t = 0; t < ns; t++)
            if (sp.mol_frac[m * SFX_MAXS + t] > 0) pm += phi_m[((long long)m * ns + t) * L.M + s];
        if (pm > 0) om -= (pm - molscal[m * MS_COUNT + MS_PHIB]) / sp.mol_N[m];
      }
      for (int t = 0; t < ns; t++) {
        if (sp.frozen[t]) continue;
        const double G = swf[(long long)t * L.M + s];
        double up = 0;
        if (G > 0) {
          double ci = 0;
          for (int j = 0; j < ns; j++) {
            const double c = sp.chi[t * SFX_MAXS + j];
            if (c != 0) ci += c * rho[(long long)j * L.M + i];
          }
          up = -log(G) - ci + cbulk[t];
        }
        om -= up * phi[(long long)t * L.M + i];
      }
      for (int t = 0; t < ns; t++) {
        if (sp.frozen[t]) continue;
        const double pt = phi[(long long)t * L.M + i];
        for (int j = 0; j < ns; j++) {
          if (sp.frozen[j]) continue;
          const double c = sp.chi[t * SFX_MAXS + j];
          if (c == 0) continue;
          om -= 0.5 * c * pt * rho[(long long)j * L.M + i];
          if (pt > 0) om += 0.5 * c * phib_state[t] * phib_state[j];
        }
      }
      v = om * w;
    }
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < n; i++) s += partial[i];
    *out = s;
  }
}

// ---------------------------------------------------------------------------
// host launchers
template <typename T, typename A, bool SCALED, bool CONNECT>
static void launch_prop_k(const DevLattice& L, const T* Gin, const T* g, T* Gout, const T* Gpair, A* acc,
                          const RankScale& sc, hipStream_t st) {
  unsigned nb = grid1(L.V, kBlock);
  switch (L.kind) {
    case ST_1D: k_prop<T, A, ST_1D, SCALED, CONNECT><<<nb, kBlock, 0, st>>>(L, Gin, g, Gout, Gpair, acc, sc); break;
    case ST_2D: k_prop<T, A, ST_2D, SCALED, CONNECT><<<nb, kBlock, 0, st>>>(L, Gin, g, Gout, Gpair, acc, sc); break;
    case ST_3D_SC: k_prop<T, A, ST_3D_SC, SCALED, CONNECT><<<nb, kBlock, 0, st>>>(L, Gin, g, Gout, Gpair, acc, sc); break;
    case ST_3D_FCC: k_prop<T, A, ST_3D_FCC, SCALED, CONNECT><<<nb, kBlock, 0, st>>>(L, Gin, g, Gout, Gpair, acc, sc); break;
    case ST_3D_HEX: k_prop<T, A, ST_3D_HEX, SCALED, CONNECT><<<nb, kBlock, 0, st>>>(L, Gin, g, Gout, Gpair, acc, sc); break;
  }
}

template <typename T, typename A>
void launch_prop(const DevLattice& L, const T* Gin, const T* g, T* Gout, const T* Gpair, A* acc,
                 const RankScale& sc, hipStream_t st) {
  bool scaled = sc.maxv != nullptr;
  if (acc) {
    if (scaled) launch_prop_k<T, A, true, true>(L, Gin, g, Gout, Gpair, acc, sc, st);
    else launch_prop_k<T, A, false, true>(L, Gin, g, Gout, Gpair, acc, sc, st);
  } else {
    if (scaled) launch_prop_k<T, A, true, false>(L, Gin, g, Gout, Gpair, acc, sc, st);
    else launch_prop_k<T, A, false, false>(L, Gin, g, Gout, Gpair, acc, sc, st);
  }
}

template <typename T>
void launch_init_rank(const DevLattice& L, const T* g, T* out, const RankScale& sc, hipStream_t st) {
  const InitRank<T> c{g, out, sc};
  k_init_rank<T><<<dim3(grid1(L.V, kBlock), 1), kBlock, 0, st>>>(L, c, c);
}

template <typename T>
void launch_init_rank2(const DevLattice& L, const InitRank<T>& c0, const InitRank<T>& c1, hipStream_t st) {
  k_init_rank<T><<<dim3(grid1(L.V, kBlock), 2), kBlock, 0, st>>>(L, c0, c1);
}

template <typename T, typename A>
void launch_connect(const DevLattice& L, const T* a, const T* b, A* acc, const RankScale& sc, double half,
                    hipStream_t st) {
  if (sc.maxv)
    k_connect<T, A, true><<<grid1(L.V, kBlock), kBlock, 0, st>>>(L, a, b, acc, sc, half);
  else
    k_connect<T, A, false><<<grid1(L.V, kBlock), kBlock, 0, st>>>(L, a, b, acc, sc, half);
}

template <typename T>
void launch_copy_interior(const DevLattice& L, const T* a, T* b, hipStream_t st) {
  k_copy_interior<T><<<grid1(L.V, kBlock), kBlock, 0, st>>>(L, a, b);
}

template <typename T>
int launch_wsum(const DevLattice& L, const T* G, double* partial, hipStream_t st) {
  unsigned nb = grid1(L.V, kBlock);
  k_wsum<T><<<nb, kBlock, 0, st>>>(L, G, partial);
  return (int)nb;
}

void launch_mol_finalize(const MolParams& mp, const double* partial, int nparts, double* scal,
                         const double* lnfN, const double* allscal, int nmol, hipStream_t st) {
  k_mol_finalize<<<1, 256, 0, st>>>(mp, partial, nparts, scal, lnfN, allscal, nmol);
}

template <typename A>
void launch_mol_phi(const DevLattice& L, const A* acc, const double* g, const double* scal, int monomer,
                    double* phi_m, hipStream_t st, const A* acc2) {
  k_mol_phi<A><<<grid1(L.V, kBlock), kBlock, 0, st>>>(L, acc, acc2, g, scal, monomer, phi_m);
}

// every state t set in `uses` (one bit per state): phi_m[t] = pref (acc[t] + acc2[t]) / g[t],
// arrays [state][M] from the given bases
template <typename A>
__global__ void k_mol_phi_states(DevLattice L, unsigned uses, int ns, const A* __restrict__ acc,
                                 const A* __restrict__ acc2, const double* __restrict__ g,
                                 const double* __restrict__ scal, double* __restrict__ phi_m) {
  long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r = sidx(L, x, y, z);
  const double pref = scal[MS_PREF];
  for (int t = 0; t < ns; t++) {
    if (!((uses >> t) & 1u)) continue;
    const long long i = (long long)t * L.M + r;
    const double gv = g[i];
    phi_m[i] = gv > 0 ? (((double)acc[i] + (acc2 ? (double)acc2[i] : 0.0)) / gv) * pref : 0.0;
  }
}

template <typename A>
void launch_mol_phi_states(const DevLattice& L, unsigned uses, int ns, const A* acc, const A* acc2, const double* g,
                           const double* scal, double* phi_m, hipStream_t st) {
  k_mol_phi_states<A><<<grid1(L.V, kBlock), kBlock, 0, st>>>(L, uses, ns, acc, acc2, g, scal, phi_m);
}

void launch_swf(const DevLattice& L, const double* u, const unsigned char* itmask, int nstates,
                const int* st_itvar, double* swf, float* swf32, hipStream_t st, unsigned* zero, int nzero) {
  if (!zero) nzero = 0;
  k_swf<<<grid1(std::max<long long>(L.M, nzero), kBlock), kBlock, 0, st>>>(L, u, itmask, nstates, st_itvar, swf,
                                                                           swf32, zero, nzero);
}

void launch_phi_total(const DevLattice& L, int nstates, int nmol, const double* phi_m, const int* mol_uses,
                      const double* fprof, const int* frozen, double* phi, double* phitot, hipStream_t st,
                      const SysParams* sp, const double* molscal, double* phib_state, double* cbulk) {
  SysParams none{};
  k_phi_total<<<grid1(L.M, kBlock), kBlock, 0, st>>>(L, nstates, nmol, phi_m, mol_uses, fprof, frozen, phi,
                                                     phitot, sp ? *sp : none, molscal, sp ? phib_state : nullptr,
                                                     cbulk);
}

void launch_sides(const DevLattice& L, int nstates, const double* phi, double* rho, hipStream_t st) {
  k_sides<<<grid1(L.M, kBlock), kBlock, 0, st>>>(L, nstates, phi, rho);
}


void launch_residual(const DevLattice& L, const SysParams& sp, const unsigned char* itmask, const double* swf,
                     const double* rho, const double* phitot, const double* cbulk, double* f, hipStream_t st) {
  k_residual<<<grid1(L.M, kBlock), kBlock, 0, st>>>(L, sp, itmask, swf, rho, phitot, cbulk, f);
}

int launch_free_energy(const DevLattice& L, const SysParams& sp, const double* phi, const double* phi_m,
                       const double* rho, const double* swf, const double* u, const double* molscal,
                       double* partial, hipStream_t st) {
  unsigned nb = grid1(L.M, kBlock);
  k_free_energy<<<nb, kBlock, 0, st>>>(L, sp, phi, phi_m, rho, swf, u, molscal, partial);
  return (int)nb;
}

__global__ void k_zero_list(ZeroList z, long long total) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long off = i;
    int k = 0;
    while (k < z.n - 1 && off >= z.words[k]) {
      off -= z.words[k];
      k++;
    }
    ((unsigned*)z.p[k])[off] = 0u;
  }
}

void launch_zero_list(const ZeroList& z, hipStream_t st) {
  long long total = 0;
  for (int k = 0; k < z.n; k++) total += z.words[k];
  if (total <= 0) return;
  const unsigned nb = (unsigned)std::min<long long>(grid1(total, kBlock), 4096);
  k_zero_list<<<nb, kBlock, 0, st>>>(z, total);
}

void zero_list_add(ZeroList& z, void* p, long long words, hipStream_t st) {
  if (words <= 0) return;
  if (z.n == 8) {
    launch_zero_list(z, st);
    z.n = 0;
  }
  z.p[z.n] = p;
  z.words[z.n] = words;
  z.n++;
}

int launch_grand_potential(const DevLattice& L, const SysParams& sp, const double* phi, const double* phi_m,
                           const double* rho, const double* swf, const double* molscal, const double* phib_state,
                           const double* cbulk, double* partial, hipStream_t st) {
  const unsigned nb = grid1(L.M, kBlock);
  k_grand_potential<<<nb, kBlock, 0, st>>>(L, sp, phi, phi_m, rho, swf, molscal, phib_state, cbulk, partial);
  return (int)nb;
}

void launch_sum_partials(const double* partial, int n, double* out, hipStream_t st) {
  k_sum_partials<<<1, 64, 0, st>>>(partial, n, out);
}

// explicit instantiations
template void launch_prop<double, double>(const DevLattice&, const double*, const double*, double*, const double*,
                                          double*, const RankScale&, hipStream_t);
template void launch_prop<float, float>(const DevLattice&, const float*, const float*, float*, const float*, float*,
                                        const RankScale&, hipStream_t);
template void launch_init_rank<double>(const DevLattice&, const double*, double*, const RankScale&, hipStream_t);
template void launch_init_rank<float>(const DevLattice&, const float*, float*, const RankScale&, hipStream_t);
template void launch_init_rank2<double>(const DevLattice&, const InitRank<double>&, const InitRank<double>&,
                                        hipStream_t);
template void launch_init_rank2<float>(const DevLattice&, const InitRank<float>&, const InitRank<float>&,
                                       hipStream_t);
template void launch_connect<double, double>(const DevLattice&, const double*, const double*, double*,
                                             const RankScale&, double, hipStream_t);
template void launch_connect<float, float>(const DevLattice&, const float*, const float*, float*, const RankScale&,
                                           double, hipStream_t);
template void launch_copy_interior<double>(const DevLattice&, const double*, double*, hipStream_t);
template void launch_copy_interior<float>(const DevLattice&, const float*, float*, hipStream_t);
__global__ void k_range_sum(DevLattice L, RangeBox box, unsigned uses, int ns, const double* __restrict__ phi_m,
                            double* __restrict__ partial) {
  const long long q = site_ordinal();
  double v = 0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    const int c[3] = {x, y, z};
    bool in = true;
    for (int a = 0; a < L.grad; a++) in = in && c[a] >= box.lo[a] && c[a] <= box.hi[a];
    if (in) {
      const long long r = sidx(L, x, y, z);
      for (int t = 0; t < ns; t++)
        if (uses >> t & 1u) v += phi_m[(size_t)t * L.M + r];
    }
  }
  v = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void k_range_finalize(const double* __restrict__ partial, int nparts, double phi_range, int N,
                                 double* __restrict__ scal) {
  __shared__ double red[256 / 64];
  double s = 0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
  const double C = phi_range / s;
  scal[MS_RANGE_C] = C;
  scal[MS_PREF] *= C;
  scal[MS_LNC] = log(C);
  // theta = sum_z L phi = C N G_N (every rank composes to G_N), so phi_b = theta / G_N = N C
  scal[MS_PHIB] = (double)N * C;
}

__global__ void k_range_scale(DevLattice L, unsigned uses, int ns, const double* __restrict__ scal,
                              double* __restrict__ phi_m) {
  const long long q = site_ordinal();
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r = sidx(L, x, y, z);
  const double C = scal[MS_RANGE_C];
  for (int t = 0; t < ns; t++)
    if (uses >> t & 1u) phi_m[(size_t)t * L.M + r] *= C;
}

int launch_range_sum(const DevLattice& L, RangeBox box, unsigned uses, int ns, const double* phi_m, double* partial,
                     hipStream_t st) {
  const unsigned nb = grid1(L.V, kBlock);
  k_range_sum<<<nb, kBlock, 0, st>>>(L, box, uses, ns, phi_m, partial);
  return (int)nb;
}

void launch_range_scale(const DevLattice& L, unsigned uses, int ns, const double* partial, int nparts,
                        double phi_range, int N, double* scal, double* phi_m, hipStream_t st) {
  k_range_finalize<<<1, 256, 0, st>>>(partial, nparts, phi_range, N, scal);
  k_range_scale<<<grid1(L.V, kBlock), kBlock, 0, st>>>(L, uses, ns, scal, phi_m);
}

template int launch_wsum<double>(const DevLattice&, const double*, double*, hipStream_t);
template int launch_wsum<float>(const DevLattice&, const float*, double*, hipStream_t);
template void launch_mol_phi<double>(const DevLattice&, const double*, const double*, const double*, int, double*,
                                     hipStream_t, const double*);
template void launch_mol_phi<float>(const DevLattice&, const float*, const double*, const double*, int, double*,
                                    hipStream_t, const float*);
template void launch_mol_phi_states<double>(const DevLattice&, unsigned, int, const double*, const double*,
                                            const double*, const double*, double*, hipStream_t);
template void launch_mol_phi_states<float>(const DevLattice&, unsigned, int, const float*, const float*,
                                           const double*, const double*, double*, hipStream_t);

}  // namespace sfx
