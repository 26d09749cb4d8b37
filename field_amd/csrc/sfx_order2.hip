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

// forward: out^k(r) = g(r) y^k(r - e_k), y = T in (y of the reversed component across a mirror)
template <typename T, bool SCALED, bool TANGENT, bool CURVED>
__global__ void __launch_bounds__(kB2) k_prop2f(DevLattice L, const T* __restrict__ Gin, const T* __restrict__ dGin,
                                                const T* __restrict__ g, const T* __restrict__ v,
                                                T* __restrict__ Gout, T* __restrict__ dGout, T S, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  T scale = T(1);
  double lncum = 0;
  if constexpr (SCALED) {
    float mp = get_step_max(sc.maxv, sc.step - 1);
    if (!(mp > 0.f)) mp = 1.f;
    lncum = sc.lnv[sc.step - 1] + log((double)mp);
    scale = (T)(1.f / mp);
  }
  const T Lw = (T(1) - S) / T(4);
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    T gv = g[r] * scale;
    T vv = TANGENT ? v[r] : T(0);
    // curved: T-rows of the site itself (the part of a radial step that stays)
    T ys[6], dys[6];
    T up = T(1), dn = T(1);
    if constexpr (CURVED) {
      up = (T)(6.0 * L.lp[x]);
      dn = (T)(6.0 * L.lm[x]);
      T self[6];
#pragma unroll
      for (int j = 0; j < 6; j++) self[j] = Gin[(long long)j * M + r];
      transfer_rows(S, Lw, self, ys);
      if constexpr (TANGENT) {
#pragma unroll
        for (int j = 0; j < 6; j++) self[j] = dGin[(long long)j * M + r];
        transfer_rows(S, Lw, self, dys);
      }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
      bool flip;
      long long nbi = nb2(L, x, y, z, kAxis[k], -kSign[k], flip);
      T in[6], yk[6];
#pragma unroll
      for (int j = 0; j < 6; j++) in[j] = nbi < 0 ? T(0) : Gin[(long long)j * M + nbi];
      transfer_rows(S, Lw, in, yk);
      T a = flip ? yk[5 - k] : yk[k];
      const bool radial = CURVED && (k == 0 || k == 5);
      const T f = k == 0 ? dn : up;
      if (radial) a = f * a + (T(1) - f) * ys[5 - k];
      T o = gv * a;
      Gout[(long long)k * M + r] = o;
      m = fmaxf(m, (float)o);
      if constexpr (TANGENT) {
        T din[6], dyk[6];
#pragma unroll
        for (int j = 0; j < 6; j++) din[j] = nbi < 0 ? T(0) : dGin[(long long)j * M + nbi];
        transfer_rows(S, Lw, din, dyk);
        T da = flip ? dyk[5 - k] : dyk[k];
        if (radial) da = f * da + (T(1) - f) * dys[5 - k];
        dGout[(long long)k * M + r] = gv * (da - vv * a);
      }
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
// mirror), then acc += c sum_k out^k Gp^k (tangent: dacc)
template <typename T, bool SCALED, bool TANGENT, bool CURVED>
__global__ void __launch_bounds__(kB2) k_prop2b(DevLattice L, const T* __restrict__ Gin, const T* __restrict__ dGin,
                                                const T* __restrict__ g, const T* __restrict__ v,
                                                T* __restrict__ Gout, T* __restrict__ dGout,
                                                const T* __restrict__ Gp, const T* __restrict__ dGp,
                                                T* __restrict__ acc, T S, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  T scale = T(1), cfac = T(1);
  double lncum = 0;
  if constexpr (SCALED) {
    float mp = get_step_max(sc.maxv, sc.step - 1);
    if (!(mp > 0.f)) mp = 1.f;
    lncum = sc.lnv[sc.step - 1] + log((double)mp);
    scale = (T)(1.f / mp);
    cfac = (T)exp(lncum + sc.lnpair[sc.pair] - *sc.ref);
  }
  const T Lw = (T(1) - S) / T(4);
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    T zz[6], dz[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
      bool flip;
      long long nbi = nb2(L, x, y, z, kAxis[j], kSign[j], flip);
      int jj = flip ? 5 - j : j;
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
    T o[6], d[6];
    transfer_rows(S, Lw, zz, o);
    if constexpr (TANGENT) transfer_rows(S, Lw, dz, d);
    T gv = g[r] * scale;
    T vv = TANGENT ? v[r] : T(0);
    T a = T(0), da = T(0);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      T ok = gv * o[k];
      Gout[(long long)k * M + r] = ok;
      m = fmaxf(m, (float)ok);
      T gp = Gp[(long long)k * M + r];
      a += ok * gp;
      if constexpr (TANGENT) {
        T dk = gv * (d[k] - vv * o[k]);
        dGout[(long long)k * M + r] = dk;
        da += dk * gp + ok * dGp[(long long)k * M + r];
      }
    }
    if constexpr (TANGENT)
      acc[r] += da * cfac;
    else
      acc[r] += a * cfac;
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

// rank 1 / rank N: all six components equal g (SF_Homopol2ndO.cpp:167, 176);
// tangent -v g; optional composition with the stored rank
template <typename T, bool TANGENT>
__global__ void k_init2(DevLattice L, const T* __restrict__ g, const T* __restrict__ v, T* __restrict__ G,
                        T* __restrict__ dG, const T* __restrict__ Gp, const T* __restrict__ dGp,
                        T* __restrict__ acc, RankScale sc) {
  long long q = site_ordinal();
  const long long M = L.M;
  float m = 0.f;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    long long r = sidx(L, x, y, z);
    T gv = g[r];
    T vv = TANGENT ? v[r] : T(0);
    T a = T(0);
    for (int k = 0; k < 6; k++) {
      G[(long long)k * M + r] = gv;
      if constexpr (TANGENT) dG[(long long)k * M + r] = -vv * gv;
      if (Gp) {
        if constexpr (TANGENT)
          a += -vv * gv * Gp[(long long)k * M + r] + gv * dGp[(long long)k * M + r];
        else
          a += gv * Gp[(long long)k * M + r];
      }
    }
    if (acc) {
      T c = T(1);
      if (sc.maxv) c = (T)exp(sc.lnpair[sc.pair] - *sc.ref);
      acc[r] += a * c;
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
template <typename T>
void launch_prop2(const DevLattice& L, bool backward, bool tangent, const T* Gin, const T* dGin, const T* g,
                  const T* v, T* Gout, T* dGout, const T* Gp, const T* dGp, T* acc, double S, const RankScale& sc,
                  hipStream_t st) {
  unsigned nb = grid1(L.V, kB2);
  bool sc_on = sc.maxv != nullptr;
  T s = (T)S;
#define F2(SC, TG, CV) k_prop2f<T, SC, TG, CV><<<nb, kB2, 0, st>>>(L, Gin, dGin, g, v, Gout, dGout, s, sc)
#define B2(SC, TG, CV) \
  k_prop2b<T, SC, TG, CV><<<nb, kB2, 0, st>>>(L, Gin, dGin, g, v, Gout, dGout, Gp, dGp, acc, s, sc)
#define F(SC, TG) \
  do {                                   \
    if (L.curved) F2(SC, TG, true);      \
    else F2(SC, TG, false);              \
  } while (0)
#define B(SC, TG) \
  do {                                   \
    if (L.curved) B2(SC, TG, true);      \
    else B2(SC, TG, false);              \
  } while (0)
  if (!backward) {
    if (sc_on) {
      if (tangent) F(true, true); else F(true, false);
    } else {
      if (tangent) F(false, true); else F(false, false);
    }
  } else {
    if (sc_on) {
      if (tangent) B(true, true); else B(true, false);
    } else {
      if (tangent) B(false, true); else B(false, false);
    }
  }
#undef F
#undef B
#undef F2
#undef B2
}

template <typename T>
void launch_init2(const DevLattice& L, bool tangent, const T* g, const T* v, T* G, T* dG, const T* Gp,
                  const T* dGp, T* acc, const RankScale& sc, hipStream_t st) {
  unsigned nb = grid1(L.V, kB2);
  if (tangent)
    k_init2<T, true><<<nb, kB2, 0, st>>>(L, g, v, G, dG, Gp, dGp, acc, sc);
  else
    k_init2<T, false><<<nb, kB2, 0, st>>>(L, g, v, G, dG, Gp, dGp, acc, sc);
}

template <typename T>
int launch_wsum6(const DevLattice& L, const T* G, const T* dG, double* p1, double* p2, hipStream_t st) {
  unsigned nb = grid1(L.V, kB2);
  k_wsum6<T><<<nb, kB2, 0, st>>>(L, G, dG, p1, p2);
  return (int)nb;
}

template void launch_prop2<double>(const DevLattice&, bool, bool, const double*, const double*, const double*,
                                   const double*, double*, double*, const double*, const double*, double*, double,
                                   const RankScale&, hipStream_t);
template void launch_prop2<float>(const DevLattice&, bool, bool, const float*, const float*, const float*,
                                  const float*, float*, float*, const float*, const float*, float*, double,
                                  const RankScale&, hipStream_t);
template void launch_init2<double>(const DevLattice&, bool, const double*, const double*, double*, double*,
                                   const double*, const double*, double*, const RankScale&, hipStream_t);
template void launch_init2<float>(const DevLattice&, bool, const float*, const float*, float*, float*, const float*,
                                  const float*, float*, const RankScale&, hipStream_t);
template int launch_wsum6<double>(const DevLattice&, const double*, const double*, double*, double*, hipStream_t);
template int launch_wsum6<float>(const DevLattice&, const float*, const float*, double*, double*, hipStream_t);

}  // namespace sfx
