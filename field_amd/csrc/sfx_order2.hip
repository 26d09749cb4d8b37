// sfx_order2.hip -- second-order (semiflexible) propagators on standard
// lattices (Lat3D2ndO.cpp; Lat2D2ndO.cpp, Lat2DCyl2ndO.cpp, Lat1D2ndO.cpp,
// Lat1DCyl2ndO.cpp and Lat1DSphere2ndO.cpp as the same six components with
// the missing axes homogeneous), base and tangent.
//
// A propagator has six direction components stored [k][M] per rank; the bond
// into a site points along e_k, e = {+x, +y, +z, -z, -y, -x}, so 5-k is the
// reversal of k. With T_kk = S, T_k,5-k = 0 and T_kj = (1-S)/4 otherwise:
//   forward  G^k_s(r) = g(r) sum_j T_kj G^j_{s-1}(r - e_k)   (PropagateF, :64-152)
//   backward G^k(r)   = g(r) sum_j T_jk G^j(r + e_j)          (PropagateB, :597-678)
// A mirror bound swaps the +/- components of its axis (SetBx1(P,Q), Lat3D.cpp:
// 1276-1290); a periodic bound copies plainly. A neighbour read that leaves
// the box is remapped to its source site and, across a mirror, takes the
// reversed component.
//
// Curved lattices (Lat1DCyl2ndO.cpp, Lat1DSphere2ndO.cpp, Lat2DCyl2ndO.cpp)
// split every radial step: with u(x) = 6 lambda_+1(x) and d(x) = 6
// lambda_-1(x) (:22-25 of each), the part u (outward) or d (inward) of the
// step reaches the neighbouring layer and the rest stays in the layer with
// the radial components reversed (LReflect/UReflect, Lat1DCyl2ndO.cpp:52-62):
//   forward, k = +x: T-row k of G(x-1), times d(x), plus T-row -x of G(x),
//                    times 1 - d(x); k = -x likewise with G(x+1) and u(x);
//   backward, z_+x = u G^+x(x+1) + (1-u) G^-x(x), z_-x = d G^-x(x-1) + (1-d) G^+x(x).
// The 2-D cylinder's backward sweep reads the plain outward neighbour, z_+x =
// G^+x(x+1), although it computes the split (Lat2DCyl2ndO.cpp:223-227).
#include "sfx_device.h"
#include "sfx_order2.h"

namespace sfx {

constexpr int kB2 = 256;

__device__ __forceinline__ float wmax2(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// neighbour (x,y,z) + d along axis a (d = +-1): returns the site index and
// whether the halo read reverses the axis' components, or -1 across a surface
// bound (the frozen surface layer carries no propagator). An axis beyond the
// lattice's gradients is homogeneous: the neighbour is the site itself (the
// 1-D and 2-D lattices lump those directions, Lat1D2ndO.cpp:32-62,
// Lat2D2ndO.cpp:31-92).
__device__ __forceinline__ long long nb2(const DevLattice& L, int x, int y, int z, int a, int d, bool& flip) {
  int c[3] = {x, y, z};
  int n[3] = {L.nx, L.ny, L.nz};
  flip = false;
  if (a >= L.grad) return sidx(L, x, y, z);
  int v = c[a] + d;
  if (v == 0) {
    if (L.surf1[a]) return -1;
    v = L.b1[a];
    flip = L.flip1[a];
  } else if (v == n[a] + 1) {
    if (L.surf2[a]) return -1;
    v = L.b2[a];
    flip = L.flip2[a];
  }
  c[a] = v;
  return sidx(L, c[0], c[1], c[2]);
}

__constant__ const int kAxis[6] = {0, 1, 2, 2, 1, 0};
__constant__ const int kSign[6] = {1, 1, 1, -1, -1, -1};

template <typename T>
__device__ __forceinline__ void transfer_rows(T S, T Lw, const T in[6], T y[6]) {
  // y_k = sum_j T_kj in_j = S in_k + L (sum - in_k - in_{5-k})
  T sum = in[0] + in[1] + in[2] + in[3] + in[4] + in[5];
#pragma unroll
  for (int k = 0; k < 6; k++) y[k] = S * in[k] + Lw * (sum - in[k] - in[5 - k]);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// y = T in on the matrix cores, FP32: v_mfma_f32_4x4x1_16b_f32 runs 16
// independent 4x4x1 products, block b on lanes 4b..4b+3. Lane 4b+j of a block
// is site j of the block: per component c it gives B[0][j] = in[c] of its own
// site and A[i][0] = T[i][c] for the row i = j it stands for, and the product
// D[i][j] comes back in lane 4b+j, register i. Two accumulators hold rows 0-3
// and rows 4-5 (rows 6-7 of the second are zero), so after six steps every
// lane holds the six outputs of its own site: no data moves between lanes.
// MFMA reads all 64 lanes, so every lane of the wave must call it (lanes
// without a site pass zeros).
__device__ __forceinline__ void transfer_mfma(float S, float Lw, const float in[6], float y[6]) {
  const int i = threadIdx.x & 3;
  f32x4 lo = {0.f, 0.f, 0.f, 0.f}, hi = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 6; c++) {
    const float tlo = i == c ? S : (i == 5 - c ? 0.f : Lw);
    const int ih = i + 4;
    const float thi = ih > 5 ? 0.f : (ih == c ? S : (ih == 5 - c ? 0.f : Lw));
    lo = __builtin_amdgcn_mfma_f32_4x4x1f32(tlo, in[c], lo, 0, 0, 0);
    hi = __builtin_amdgcn_mfma_f32_4x4x1f32(thi, in[c], hi, 0, 0, 0);
  }
  y[0] = lo[0];
  y[1] = lo[1];
  y[2] = lo[2];
  y[3] = lo[3];
  y[4] = hi[0];
  y[5] = hi[1];
}

// the transfer of one site, on the matrix cores for FP32 when MFMA is set;
// wave-uniform control flow is required in the MFMA case
template <typename T, bool MFMA>
__device__ __forceinline__ void transfer(T S, T Lw, const T in[6], T y[6]) {
  if constexpr (MFMA && sizeof(T) == 4)
    transfer_mfma(S, Lw, in, y);
  else
    transfer_rows(S, Lw, in, y);
}

// the six neighbours r + e_d of a site (e = {+x, +y, +z, -z, -y, -x}) with the
// component reversal of a mirror halo, or -1 across a surface (nb2)
__device__ __forceinline__ void neighbours(const DevLattice& L, int x, int y, int z, long long nbr[6],
                                           bool fl[6]) {
#pragma unroll
  for (int d = 0; d < 6; d++) nbr[d] = nb2(L, x, y, z, kAxis[d], kSign[d], fl[d]);
}

// T-row k of a rank at the neighbour r - e_k = r + e_{5-k}, read from the
// rank's stored W = T G form: W^k there, W^{5-k} across a mirror, 0 across a
// surface; a curved radial step mixes in W^{5-k} at r itself
template <typename T, bool CURVED>
__device__ __forceinline__ T fwd_row(const T* __restrict__ W, long long M, const long long nbr[6], const bool fl[6],
                                     long long r, int k, T up, T dn) {
  const long long nbi = nbr[5 - k];
  T a = nbi < 0 ? T(0) : W[(long long)(fl[5 - k] ? 5 - k : k) * M + nbi];
  if constexpr (CURVED) {
    if (k == 0 || k == 5) {
      const T f = k == 0 ? dn : up;
      a = f * a + (T(1) - f) * W[(long long)(5 - k) * M + r];
    }
  }
  return a;
}

template <typename T, bool CURVED>
__device__ __forceinline__ void curved_fractions(const DevLattice& L, int x, T& up, T& dn) {
  up = T(1);
  dn = T(1);
  if constexpr (CURVED) {
    up = (T)(6.0 * L.lp[x]);
    dn = (T)(6.0 * L.lm[x]);
  }
}

// the forward partner G_f(s) at site r (and its tangent), recovered from the
// stored W_f(s-1) exactly as the forward step produced it: g(r) times the
// forward step's scale times the gathered row; rank 1 (Wp null) is g itself.
// pair_rows gathers the rows (the loads), pair_scale applies g and the scale,
// so a kernel can issue the loads before it waits on the scale.
template <typename T, bool TANGENT, bool CURVED>
__device__ __forceinline__ void pair_rows(const DevLattice& L, const T* __restrict__ Wp, const T* __restrict__ dWp,
                                          long long M, int x, const long long nbr[6], const bool fl[6], long long r,
                                          T pa[6], T dpa[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) pa[k] = dpa[k] = T(0);
  if (!Wp) return;
  T up, dn;
  curved_fractions<T, CURVED>(L, x, up, dn);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    pa[k] = fwd_row<T, CURVED>(Wp, M, nbr, fl, r, k, up, dn);
    if constexpr (TANGENT) dpa[k] = fwd_row<T, CURVED>(dWp, M, nbr, fl, r, k, up, dn);
  }
}

template <typename T, bool TANGENT>
__device__ __forceinline__ void pair_scale(bool rank1, T graw, T vv, T pscale, const T pa[6], const T dpa[6], T gp[6],
                                           T dgp[6]) {
  if (rank1) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      gp[k] = graw;
      dgp[k] = -vv * graw;
    }
    return;
  }
  const T gvp = graw * pscale;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    gp[k] = gvp * pa[k];
    if constexpr (TANGENT) dgp[k] = gvp * (dpa[k] - vv * pa[k]);
  }
}

template <typename T, bool TANGENT, bool CURVED>
__device__ __forceinline__ void pair_rank(const DevLattice& L, const T* __restrict__ Wp, const T* __restrict__ dWp,
                                          T graw, T vv, long long M, int x, const long long nbr[6],
                                          const bool fl[6], long long r, T pscale, T gp[6], T dgp[6]) {
  T pa[6], dpa[6];
  pair_rows<T, TANGENT, CURVED>(L, Wp, dWp, M, x, nbr, fl, r, pa, dpa);
  pair_scale<T, TANGENT>(!Wp, graw, vv, pscale, pa, dpa, gp, dgp);
}

// forward: G^k_s(r) = g(r) (T G_{s-1})^k(r - e_k) from the stored W_{s-1};
// the rank is stored as W_s = T G_s (tangent: dW_s = T dG_s)
template <typename T, bool SCALED, bool TANGENT, bool CURVED, bool MFMA>
__global__ void __launch_bounds__(kB2) k_prop2f(DevLattice L, const T* __restrict__ Win, const T* __restrict__ dWin,
                                                const T* __restrict__ g, const T* __restrict__ v,
                                                T* __restrict__ Wout, T* __restrict__ dWout, T S, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  const bool in = q < L.V;
  // this launch's loads first: the step scale below waits on the previous
  // launch's maxima while they are in flight
  long long r = 0;
  T graw = T(0), vv = T(0);
  T a[6], da[6];
#pragma unroll
  for (int k = 0; k < 6; k++) a[k] = da[k] = T(0);
  if (in) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    r = sidx(L, x, y, z);
    graw = g[r];
    if constexpr (TANGENT) vv = v[r];
    T up, dn;
    curved_fractions<T, CURVED>(L, x, up, dn);
    long long nbr[6];
    bool fl[6];
    neighbours(L, x, y, z, nbr, fl);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      a[k] = fwd_row<T, CURVED>(Win, M, nbr, fl, r, k, up, dn);
      if constexpr (TANGENT) da[k] = fwd_row<T, CURVED>(dWin, M, nbr, fl, r, k, up, dn);
    }
  }
  T scale = T(1);
  double lncum = 0;
  if constexpr (SCALED) {
    const double l0 = sc.lnv[sc.step - 1];
    float mp, unused;
    block_step_max2(sc.maxv, sc.step - 1, nullptr, 0, mp, unused);
    if (!(mp > 0.f)) mp = 1.f;
    lncum = l0 + log((double)mp);
    scale = (T)(1.f / mp);
  }
  const T Lw = (T(1) - S) / T(4);
  float m = 0.f;
  // lanes without a site carry zeros through the transfer (it may run on the
  // matrix cores, which need the whole wave)
  const T gv = graw * scale;
  T G[6], dG[6], w[6], dw[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    G[k] = gv * a[k];
    m = fmaxf(m, (float)G[k]);
    dG[k] = TANGENT ? gv * (da[k] - vv * a[k]) : T(0);
  }
  transfer<T, MFMA>(S, Lw, G, w);
  if constexpr (TANGENT) transfer<T, MFMA>(S, Lw, dG, dw);
  if (in) {
#pragma unroll
    for (int k = 0; k < 6; k++) Wout[(long long)k * M + r] = w[k];
    if constexpr (TANGENT) {
#pragma unroll
      for (int k = 0; k < 6; k++) dWout[(long long)k * M + r] = dw[k];
    }
  }
  if constexpr (SCALED) {
    __shared__ float sm[kB2 / 64];
    m = wmax2(m);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, sm[i]);
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = lncum;
    }
  }
}

// backward: out = g T z with z_j = G^j(r + e_j) (reversed component across a
// mirror), then acc += c sum_k out^k G_f^k (tangent: dacc) with the forward
// partner recovered from W_f(s-1)
template <typename T, bool SCALED, bool TANGENT, bool CURVED, bool MFMA>
__global__ void __launch_bounds__(kB2) k_prop2b(DevLattice L, const T* __restrict__ Gin, const T* __restrict__ dGin,
                                                const T* __restrict__ g, const T* __restrict__ v,
                                                T* __restrict__ Gout, T* __restrict__ dGout,
                                                const T* __restrict__ Wp, const T* __restrict__ dWp,
                                                T* __restrict__ acc, T S, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  const bool in = q < L.V;
  // this launch's loads first (the chain's previous rank, g, v, the partner's
  // rows and acc): the scales below wait on the previous launch's maxima while
  // they are in flight
  const T Lw = (T(1) - S) / T(4);
  long long r = 0;
  T o[6], d[6], pa[6], dpa[6], zz[6], dz[6];
  T graw = T(0), vv = T(0), acc0 = T(0);
#pragma unroll
  for (int k = 0; k < 6; k++) o[k] = d[k] = zz[k] = dz[k] = T(0);
  if (in) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    r = sidx(L, x, y, z);
    long long nbr[6];
    bool fl[6];
    neighbours(L, x, y, z, nbr, fl);
#pragma unroll
    for (int j = 0; j < 6; j++) {
      const long long nbi = nbr[j];
      const int jj = fl[j] ? 5 - j : j;
      zz[j] = nbi < 0 ? T(0) : Gin[(long long)jj * M + nbi];
      if constexpr (TANGENT) dz[j] = nbi < 0 ? T(0) : dGin[(long long)jj * M + nbi];
    }
    if constexpr (CURVED) {
      const T up = L.grad == 2 ? T(1) : (T)(6.0 * L.lp[x]);  // 2-D: plain outward read
      const T dn = (T)(6.0 * L.lm[x]);
      const T g0 = Gin[r], g5 = Gin[5 * M + r];
      zz[0] = up * zz[0] + (T(1) - up) * g5;
      zz[5] = dn * zz[5] + (T(1) - dn) * g0;
      if constexpr (TANGENT) {
        const T d0 = dGin[r], d5 = dGin[5 * M + r];
        dz[0] = up * dz[0] + (T(1) - up) * d5;
        dz[5] = dn * dz[5] + (T(1) - dn) * d0;
      }
    }
    graw = g[r];
    if constexpr (TANGENT) vv = v[r];
    pair_rows<T, TANGENT, CURVED>(L, Wp, dWp, M, x, nbr, fl, r, pa, dpa);
    acc0 = acc[r];
  }
  // outside the per-site branch: lanes without a site transfer zeros (the
  // matrix-core form needs the whole wave)
  transfer<T, MFMA>(S, Lw, zz, o);
  if constexpr (TANGENT) transfer<T, MFMA>(S, Lw, dz, d);
  T scale = T(1), cfac = T(1), pscale = T(1);
  double lncum = 0;
  if constexpr (SCALED) {
    const double l0 = sc.lnv[sc.step - 1], lp = sc.lnpair[sc.pair], rf = *sc.ref;
    float mp, mpp;
    block_step_max2(sc.maxv, sc.step - 1, Wp ? sc.pmaxv : nullptr, sc.pair - 1, mp, mpp);
    if (!(mp > 0.f)) mp = 1.f;
    if (!(mpp > 0.f)) mpp = 1.f;
    lncum = l0 + log((double)mp);
    scale = (T)(1.f / mp);
    pscale = (T)(1.f / mpp);
    cfac = (T)exp(lncum + lp - rf);
  }
  float m = 0.f;
  if (in) {
    const T gv = graw * scale;
    T gp[6], dgp[6];
    pair_scale<T, TANGENT>(!Wp, graw, vv, pscale, pa, dpa, gp, dgp);
    T a = T(0), da = T(0);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      T ok = gv * o[k];
      Gout[(long long)k * M + r] = ok;
      m = fmaxf(m, (float)ok);
      a += ok * gp[k];
      if constexpr (TANGENT) {
        T dk = gv * (d[k] - vv * o[k]);
        dGout[(long long)k * M + r] = dk;
        da += dk * gp[k] + ok * dgp[k];
      }
    }
    if constexpr (TANGENT)
      acc[r] = acc0 + da * cfac;
    else
      acc[r] = acc0 + a * cfac;
  }
  if constexpr (SCALED) {
    __shared__ float sm[kB2 / 64];
    m = wmax2(m);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, sm[i]);
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = lncum;
    }
  }
}

// rank 1 of the forward chain: all six components of G are g
// (SF_Homopol2ndO.cpp:167; tangent -v g), stored in the W = T G form
template <typename T, bool TANGENT>
__global__ void k_init2f(DevLattice L, const T* __restrict__ g, const T* __restrict__ v, T* __restrict__ W,
                         T* __restrict__ dW, T S, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  const T Lw = (T(1) - S) / T(4);
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    const T gv = g[r];
    T G[6], w[6];
#pragma unroll
    for (int k = 0; k < 6; k++) G[k] = gv;
    transfer_rows(S, Lw, G, w);
#pragma unroll
    for (int k = 0; k < 6; k++) W[(long long)k * M + r] = w[k];
    if constexpr (TANGENT) {
      const T vv = v[r];
#pragma unroll
      for (int k = 0; k < 6; k++) G[k] = -vv * gv;
      transfer_rows(S, Lw, G, w);
#pragma unroll
      for (int k = 0; k < 6; k++) dW[(long long)k * M + r] = w[k];
    }
    m = (float)gv;
  }
  if (sc.maxv) {
    __shared__ float sm[kB2 / 64];
    m = wmax2(m);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, sm[i]);
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = 0.0;
    }
  }
}

// rank N of the backward chain: all six components are g (SF_Homopol2ndO.cpp:176;
// tangent -v g), composed with the forward partner recovered from Wp = W_f(N-1)
template <typename T, bool SCALED, bool TANGENT, bool CURVED>
__global__ void k_init2b(DevLattice L, const T* __restrict__ g, const T* __restrict__ v, T* __restrict__ G,
                         T* __restrict__ dG, const T* __restrict__ Wp, const T* __restrict__ dWp,
                         T* __restrict__ acc, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  T pscale = T(1);
  if constexpr (SCALED) {
    float mpp, unused;
    block_step_max2(sc.pmaxv, sc.pair - 1, nullptr, 0, mpp, unused);
    if (!(mpp > 0.f)) mpp = 1.f;
    pscale = (T)(1.f / mpp);
  }
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    const T gv = g[r];
    const T vv = TANGENT ? v[r] : T(0);
    long long nbr[6];
    bool fl[6];
    neighbours(L, x, y, z, nbr, fl);
    T gp[6], dgp[6];
    pair_rank<T, TANGENT, CURVED>(L, Wp, dWp, gv, vv, M, x, nbr, fl, r, pscale, gp, dgp);
    T a = T(0);
    for (int k = 0; k < 6; k++) {
      G[(long long)k * M + r] = gv;
      if constexpr (TANGENT) {
        dG[(long long)k * M + r] = -vv * gv;
        a += -vv * gv * gp[k] + gv * dgp[k];
      } else {
        a += gv * gp[k];
      }
    }
    T c = T(1);
    if constexpr (SCALED) c = (T)exp(sc.lnpair[sc.pair] - *sc.ref);
    acc[r] += a * c;
    m = (float)gv;
  }
  if constexpr (SCALED) {
    __shared__ float sm[kB2 / 64];
    m = wmax2(m);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, sm[i]);
      put_step_max(sc.maxv, sc.step, m);
      if (blockIdx.x == 0) sc.lnv[sc.step] = 0.0;
    }
  }
}

// weighted sums over the six components (ln G_N before /6 and the log)
template <typename T>
__global__ void k_wsum6(DevLattice L, const T* __restrict__ G, const T* __restrict__ dG, double* __restrict__ p1,
                        double* __restrict__ p2) {
  long long q = site_ordinal();
  double a = 0, b = 0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    // flat 1-D sums every layer plainly (Lat1D2ndO.cpp ComputeLnGN); 2-D and 3-D
    // halve mirror2 edges (Lat2D2ndO.cpp:387-439, Lat3D2ndO.cpp:996-1025);
    // curved lattices weight by the layer volume (Lat1DCyl2ndO.cpp:319-330,
    // Lat2DCyl2ndO.cpp:410-462)
    double w = L.grad == 1 && !L.curved ? 1.0 : lngn_weight(L, x, y, z);
    for (int k = 0; k < 6; k++) {
      a += (double)G[(long long)k * L.M + r] * w;
      if (dG) b += (double)dG[(long long)k * L.M + r] * w;
    }
  }
  __shared__ double sa[kB2 / 64], sb[kB2 / 64];
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
    if (p2) p2[blockIdx.x] = B;
  }
}

// ---------------------------------------------------------------------------
#define SFX_O2_DISPATCH(K, ...)                 \
  do {                                          \
    if (sc_on) {                                \
      if (tangent) {                            \
        if (L.curved) K(true, true, true);      \
        else K(true, true, false);              \
      } else {                                  \
        if (L.curved) K(true, false, true);     \
        else K(true, false, false);             \
      }                                         \
    } else {                                    \
      if (tangent) {                            \
        if (L.curved) K(false, true, true);     \
        else K(false, true, false);             \
      } else {                                  \
        if (L.curved) K(false, false, true);    \
        else K(false, false, false);            \
      }                                         \
    }                                           \
  } while (0)

// the matrix-core transfer runs for FP32 sweeps when the lattice asks for it
// (sfx_options.transfer_kernel: 1 always, 2 in tangent sweeps only, where it
// measured faster than the rows at 128^3; the base sweeps measured the other
// way); FP64 keeps the structured rows
static bool use_mfma(const DevLattice& L, bool tangent, size_t word) {
  return word == 4 && (L.transfer_mfma == 1 || (L.transfer_mfma == 2 && tangent));
}
template <typename T>
void launch_prop2f(const DevLattice& L, bool tangent, const T* Win, const T* dWin, const T* g, const T* v, T* Wout,
                   T* dWout, double S, const RankScale& sc, hipStream_t st) {
  const unsigned nb = grid1(L.V, kB2);
  const bool sc_on = sc.maxv != nullptr;
  const T s = (T)S;
  if (use_mfma(L, tangent, sizeof(T))) {
#define K(SC, TG, CV) k_prop2f<T, SC, TG, CV, true><<<nb, kB2, 0, st>>>(L, Win, dWin, g, v, Wout, dWout, s, sc)
    SFX_O2_DISPATCH(K);
#undef K
  } else {
#define K(SC, TG, CV) k_prop2f<T, SC, TG, CV, false><<<nb, kB2, 0, st>>>(L, Win, dWin, g, v, Wout, dWout, s, sc)
    SFX_O2_DISPATCH(K);
#undef K
  }
}

template <typename T>
void launch_prop2b(const DevLattice& L, bool tangent, const T* Gin, const T* dGin, const T* g, const T* v, T* Gout,
                   T* dGout, const T* Wp, const T* dWp, T* acc, double S, const RankScale& sc, hipStream_t st) {
  const unsigned nb = grid1(L.V, kB2);
  const bool sc_on = sc.maxv != nullptr;
  const T s = (T)S;
  if (use_mfma(L, tangent, sizeof(T))) {
#define K(SC, TG, CV) \
  k_prop2b<T, SC, TG, CV, true><<<nb, kB2, 0, st>>>(L, Gin, dGin, g, v, Gout, dGout, Wp, dWp, acc, s, sc)
    SFX_O2_DISPATCH(K);
#undef K
  } else {
#define K(SC, TG, CV) \
  k_prop2b<T, SC, TG, CV, false><<<nb, kB2, 0, st>>>(L, Gin, dGin, g, v, Gout, dGout, Wp, dWp, acc, s, sc)
    SFX_O2_DISPATCH(K);
#undef K
  }
}

template <typename T>
void launch_init2f(const DevLattice& L, bool tangent, const T* g, const T* v, T* W, T* dW, double S,
                   const RankScale& sc, hipStream_t st) {
  const unsigned nb = grid1(L.V, kB2);
  if (tangent)
    k_init2f<T, true><<<nb, kB2, 0, st>>>(L, g, v, W, dW, (T)S, sc);
  else
    k_init2f<T, false><<<nb, kB2, 0, st>>>(L, g, v, W, dW, (T)S, sc);
}

template <typename T>
void launch_init2b(const DevLattice& L, bool tangent, const T* g, const T* v, T* G, T* dG, const T* Wp,
                   const T* dWp, T* acc, const RankScale& sc, hipStream_t st) {
  const unsigned nb = grid1(L.V, kB2);
  const bool sc_on = sc.maxv != nullptr;
#define K(SC, TG, CV) k_init2b<T, SC, TG, CV><<<nb, kB2, 0, st>>>(L, g, v, G, dG, Wp, dWp, acc, sc)
  SFX_O2_DISPATCH(K);
#undef K
}
#undef SFX_O2_DISPATCH

template <typename T>
int launch_wsum6(const DevLattice& L, const T* G, const T* dG, double* p1, double* p2, hipStream_t st) {
  unsigned nb = grid1(L.V, kB2);
  k_wsum6<T><<<nb, kB2, 0, st>>>(L, G, dG, p1, p2);
  return (int)nb;
}

template void launch_prop2f<double>(const DevLattice&, bool, const double*, const double*, const double*,
                                    const double*, double*, double*, double, const RankScale&, hipStream_t);
template void launch_prop2f<float>(const DevLattice&, bool, const float*, const float*, const float*, const float*,
                                   float*, float*, double, const RankScale&, hipStream_t);
template void launch_prop2b<double>(const DevLattice&, bool, const double*, const double*, const double*,
                                    const double*, double*, double*, const double*, const double*, double*, double,
                                    const RankScale&, hipStream_t);
template void launch_prop2b<float>(const DevLattice&, bool, const float*, const float*, const float*, const float*,
                                   float*, float*, const float*, const float*, float*, double, const RankScale&,
                                   hipStream_t);
template void launch_init2f<double>(const DevLattice&, bool, const double*, const double*, double*, double*, double,
                                    const RankScale&, hipStream_t);
template void launch_init2f<float>(const DevLattice&, bool, const float*, const float*, float*, float*, double,
                                   const RankScale&, hipStream_t);
template void launch_init2b<double>(const DevLattice&, bool, const double*, const double*, double*, double*,
                                    const double*, const double*, double*, const RankScale&, hipStream_t);
template void launch_init2b<float>(const DevLattice&, bool, const float*, const float*, float*, float*, const float*,
                                   const float*, float*, const RankScale&, hipStream_t);
template int launch_wsum6<double>(const DevLattice&, const double*, const double*, double*, double*, hipStream_t);
template int launch_wsum6<float>(const DevLattice&, const float*, const float*, double*, double*, hipStream_t);

}  // namespace sfx
