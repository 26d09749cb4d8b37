// sfx_block3.hip -- temporally blocked propagator ranks on the 3-D simple-cubic
// lattice: one launch advances a chain by up to four ranks.
//
// A workgroup owns an 8 x 8 x 16 output tile. It loads the previous rank and
// the state's SWF on the tile widened by the number of ranks nr in every
// direction (a lattice with bounds extends by periodic wrap or by reflection,
// mirror1 about the face and mirror2 about the first layer, which is what the
// per-rank halo copies of Lat3D.cpp:892-1562 amount to), then computes rank k
// on the tile widened by nr - k, ping-ponging between two LDS buffers:
//   G_k(r) = g(r) (1/6) sum_6 G_{k-1}(r + e)            (Lat3D1stO.cpp:53-260)
//   dG_k(r) = g(r) ((1/6) sum_6 dG_{k-1} - v(r) (1/6) sum_6 G_{k-1})
// Tile sites of every rank go to the store and/or the composition
// acc += G_k G_pair c_k (ConnectG, Lat3D1stO.cpp:299-312; tangent
// dacc += (dG_k G_pair + G_k dG_pair) c_k), accumulated in LDS and added to
// HBM once per launch. Each thread walks one (y, z) column along x and keeps
// the x-neighbours in registers; consecutive threads take consecutive z, so
// LDS accesses are conflict-free and HBM accesses coalesce along z.
//
// FP32 scaling: the input is divided by max(rank s0) and every rank by the
// lattice maximum of g, so no rank exceeds 1 and at most nr ranks decay
// between renormalisations; the logs of the divisors go to lnv.
#include "sfx_block3.h"
#include "sfx_device.h"

namespace sfx {

constexpr int kTX = 8, kTY = 8, kTZ = 16;
constexpr int kH = kB3Ranks;
constexpr int kRX = kTX + 2 * kH, kRY = kTY + 2 * kH, kRZ = kTZ + 2 * kH;
constexpr int kRYZ = kRY * kRZ;
constexpr int kR = kRX * kRYZ;  // region sites
constexpr int kT = kTX * kTY * kTZ;
constexpr int kB3 = 256;

bool block3_ok(const DevLattice& L) {
  return L.grad == 3 && L.kind == ST_3D_SC && L.order == 1 && L.nx > kH && L.ny > kH && L.nz > kH;
}

// lattice coordinate c -> interior coordinate. Within kH of the lattice one
// wrap or reflection is exact (every side has more than kH layers); farther
// out (a tile hanging over the lattice's far edge) the value never reaches a
// tile site inside the lattice, and the coordinate is only clamped into range.
__device__ __forceinline__ int extend(int c, int n, int periodic, int m2lo, int m2hi) {
  if (c < 1) c = periodic ? c + n : (m2lo ? 2 - c : 1 - c);
  else if (c > n) c = periodic ? c - n : (m2hi ? 2 * n - c : 2 * n + 1 - c);
  return c < 1 ? 1 : (c > n ? n : c);
}

template <typename T>
__device__ __forceinline__ T block_max_T(T m, float* red) {
  float v = (float)m;
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = red[0];
  for (int i = 1; i < kB3 / 64; i++) r = fmaxf(r, red[i]);
  return (T)r;
}

template <typename T, bool TANGENT, bool SCALED, bool CONNECT>
__global__ void __launch_bounds__(kB3) k_block3(DevLattice L, B3Args<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* G0 = (T*)smem;
  T* G1 = G0 + kR;
  T* gl = G1 + kR;
  T* D0 = TANGENT ? gl + kR : nullptr;
  T* D1 = TANGENT ? D0 + kR : nullptr;
  T* vl = TANGENT ? D1 + kR : nullptr;
  T* accT = (TANGENT ? vl : gl) + kR;
  __shared__ float red[kB3 / 64];

  const int t = threadIdx.x;
  const int nbz = (L.nz + kTZ - 1) / kTZ, nby = (L.ny + kTY - 1) / kTY;
  int b = blockIdx.x;
  const int bz = b % nbz;
  b /= nbz;
  const int by = b % nby;
  const int bx = b / nby;
  const int x0 = 1 + bx * kTX, y0 = 1 + by * kTY, z0 = 1 + bz * kTZ;  // tile origin (interior coords)
  const int nr = a.nr;

  // scale factors of this launch
  T in_scale = T(1), ginv = T(1);
  double ln_base = 0.0, lng = 0.0;
  if constexpr (SCALED) {
    float m0 = get_step_max(a.maxv, a.s0);
    if (!(m0 > 0.f)) m0 = 1.f;
    float gm = *a.gmax;
    if (!(gm > 0.f)) gm = 1.f;
    in_scale = (T)(1.f / m0);
    ginv = (T)(1.f / gm);
    ln_base = a.lnv[a.s0] + log((double)m0);
    lng = log((double)gm);
  }

  // load rank s0 (and its tangent), g (and v) on the tile widened by nr
  {
    const int lo = kH - nr, w = kTX + 2 * nr, wy = kTY + 2 * nr, wz = kTZ + 2 * nr;
    const int n = w * wy * wz;
    for (int i = t; i < n; i += kB3) {
      const int Z = i % wz, r = i / wz, Y = r % wy, X = r / wy;
      const int x = extend(x0 + lo + X - kH, L.nx, L.periodic[0], L.mirror2_1[0], L.mirror2_2[0]);
      const int y = extend(y0 + lo + Y - kH, L.ny, L.periodic[1], L.mirror2_1[1], L.mirror2_2[1]);
      const int z = extend(z0 + lo + Z - kH, L.nz, L.periodic[2], L.mirror2_1[2], L.mirror2_2[2]);
      const long long gi = sidx(L, x, y, z);
      const int li = (lo + X) * kRYZ + (lo + Y) * kRZ + (lo + Z);
      G0[li] = a.in[gi] * in_scale;
      gl[li] = a.g[gi];
      if constexpr (TANGENT) {
        D0[li] = a.din[gi] * in_scale;
        vl[li] = a.v[gi];
      }
    }
    if constexpr (CONNECT)
      for (int i = t; i < kT; i += kB3) accT[i] = T(0);
  }
  __syncthreads();

  T* src = G0;
  T* dst = G1;
  T* dsrc = D0;
  T* ddst = D1;
  T mloc = T(0);
  const T sixth = T(1.0 / 6.0);
  for (int k = 1; k <= nr; k++) {
    const int e = nr - k;
    const int lo = kH - e;
    const int bxn = kTX + 2 * e, byn = kTY + 2 * e, bzn = kTZ + 2 * e;
    const int ncol = byn * bzn;
    T* outk = a.out[k - 1];
    T* doutk = TANGENT ? a.dout[k - 1] : nullptr;
    const T* pk = CONNECT ? a.pair[k - 1] : nullptr;
    const T* dpk = (CONNECT && TANGENT) ? a.dpair[k - 1] : nullptr;
    T cfac = T(1);
    if constexpr (SCALED && CONNECT) cfac = (T)exp(ln_base + k * lng + a.lnpair[a.pidx[k - 1]] - *a.ref);
    for (int c = t; c < ncol; c += kB3) {
      const int Y = lo + c / bzn, Z = lo + c % bzn;
      const int base = Y * kRZ + Z;
      const bool yz_tile = Y >= kH && Y < kH + kTY && Z >= kH && Z < kH + kTZ && y0 + Y - kH <= L.ny &&
                           z0 + Z - kH <= L.nz;
      int idx = lo * kRYZ + base;
      T xm = src[idx - kRYZ], xc = src[idx];
      T dxm = T(0), dxc = T(0);
      if constexpr (TANGENT) {
        dxm = dsrc[idx - kRYZ];
        dxc = dsrc[idx];
      }
      for (int X = lo; X < lo + bxn; X++, idx += kRYZ) {
        const T xp = src[idx + kRYZ];
        const T s = (xm + xp + src[idx - kRZ] + src[idx + kRZ] + src[idx - 1] + src[idx + 1]) * sixth;
        const T gv = gl[idx] * ginv;
        const T val = gv * s;
        dst[idx] = val;
        T dval = T(0);
        if constexpr (TANGENT) {
          const T dxp = dsrc[idx + kRYZ];
          const T ds = (dxm + dxp + dsrc[idx - kRZ] + dsrc[idx + kRZ] + dsrc[idx - 1] + dsrc[idx + 1]) * sixth;
          dval = gv * (ds - vl[idx] * s);
          ddst[idx] = dval;
          dxm = dxc;
          dxc = dxp;
        }
        xm = xc;
        xc = xp;
        if (yz_tile && X >= kH && X < kH + kTX && x0 + X - kH <= L.nx) {
          const long long gi = sidx(L, x0 + X - kH, y0 + Y - kH, z0 + Z - kH);
          if (outk) outk[gi] = val;
          if constexpr (TANGENT) {
            if (doutk) doutk[gi] = dval;
          }
          if constexpr (CONNECT) {
            const int ti = ((X - kH) * kTY + (Y - kH)) * kTZ + (Z - kH);
            if constexpr (TANGENT)
              accT[ti] += (dval * pk[gi] + val * dpk[gi]) * cfac;
            else
              accT[ti] += val * pk[gi] * cfac;
          }
          if (k == nr) mloc = val > mloc ? val : mloc;
        }
      }
    }
    __syncthreads();
    T* tmp = src;
    src = dst;
    dst = tmp;
    if constexpr (TANGENT) {
      tmp = dsrc;
      dsrc = ddst;
      ddst = tmp;
    }
  }

  if constexpr (CONNECT) {
    for (int i = t; i < kT; i += kB3) {
      const int Z = i % kTZ, r = i / kTZ, Y = r % kTY, X = r / kTY;
      const int x = x0 + X, y = y0 + Y, z = z0 + Z;
      if (x <= L.nx && y <= L.ny && z <= L.nz) a.acc[sidx(L, x, y, z)] += accT[i];
    }
  }
  if constexpr (SCALED) {
    const T m = block_max_T<T>(mloc, red);
    if (t == 0) {
      put_step_max(a.maxv, a.s0 + nr, (float)m);
      if (blockIdx.x == 0)
        for (int k = 1; k <= nr; k++) a.lnv[a.s0 + k] = ln_base + k * lng;
    }
  }
}

template <typename T, bool TANGENT>
static size_t b3_lds() {
  return ((TANGENT ? 6 : 3) * (size_t)kR + kT) * sizeof(T);
}

template <typename T, bool TG, bool SC, bool CN>
static void b3_launch(const DevLattice& L, const B3Args<T>& a, unsigned nb, hipStream_t st) {
  static bool attr = false;  // dynamic LDS above the default limit needs the attribute once
  const size_t lds = b3_lds<T, TG>();
  if (!attr) {
    SFX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_block3<T, TG, SC, CN>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  k_block3<T, TG, SC, CN><<<nb, kB3, lds, st>>>(L, a);
}

template <typename T>
void launch_block3(const DevLattice& L, const B3Args<T>& a, bool tangent, bool scaled, bool connect,
                   hipStream_t st) {
  SFX_REQUIRE(a.nr >= 1 && a.nr <= kB3Ranks, SFX_E_INVALID, "blocked launch of 1..4 ranks");
  SFX_REQUIRE(!(tangent && sizeof(T) == 8), SFX_E_UNSUPPORTED, "FP64 tangent sweeps run rank by rank");
  const size_t lds_base = b3_lds<T, false>(), lds_tan = b3_lds<T, true>();
  SFX_REQUIRE((lds_base <= 160 * 1024 && (sizeof(T) == 8 || lds_tan <= 160 * 1024)), SFX_E_UNSUPPORTED,
              "blocked tile exceeds LDS");
  const unsigned nb = (unsigned)(((L.nx + kTX - 1) / kTX) * ((L.ny + kTY - 1) / kTY) * ((L.nz + kTZ - 1) / kTZ));
#define B3(TG, SC, CN) b3_launch<T, TG, SC, CN>(L, a, nb, st)
  if (tangent) {
    if (scaled) {
      if (connect) B3(true, true, true); else B3(true, true, false);
    } else {
      if (connect) B3(true, false, true); else B3(true, false, false);
    }
  } else {
    if (scaled) {
      if (connect) B3(false, true, true); else B3(false, true, false);
    } else {
      if (connect) B3(false, false, true); else B3(false, false, false);
    }
  }
#undef B3
  SFX_HIP(hipGetLastError());
}

// one block per state: max of the FP32 SWF over the interior
__global__ void k_gmax(DevLattice L, const float* __restrict__ swf32, float* __restrict__ gmax) {
  const float* g = swf32 + (long long)blockIdx.x * L.M;
  float m = 0.f;
  for (long long q = threadIdx.x; q < L.V; q += blockDim.x) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    m = fmaxf(m, g[sidx(L, x, y, z)]);
  }
  __shared__ float red[16];
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmaxf(m, red[i]);
    gmax[blockIdx.x] = m;
  }
}

void launch_gmax(const DevLattice& L, const float* swf32, int ns, float* gmax, hipStream_t st) {
  k_gmax<<<ns, 1024, 0, st>>>(L, swf32, gmax);
  SFX_HIP(hipGetLastError());
}

template void launch_block3<float>(const DevLattice&, const B3Args<float>&, bool, bool, bool, hipStream_t);
template void launch_block3<double>(const DevLattice&, const B3Args<double>&, bool, bool, bool, hipStream_t);

}  // namespace sfx
