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
// Every thread owns one tile site: it stores that site's value of every rank,
// composes acc += G_k G_pair c_k (ConnectG, Lat3D1stO.cpp:299-312; tangent
// dacc += (dG_k G_pair + G_k dG_pair) c_k) in a register with the partner
// values prefetched at launch start, and adds it to HBM once. In FP32 the
// widened ranks are computed four z sites per work item with float4 LDS
// accesses (consecutive items take consecutive 16-byte words of whole rows);
// in FP64 by (y, z) columns, several threads per column along x, consecutive
// threads on consecutive LDS words. LDS reads are conflict-free either way,
// and HBM accesses coalesce along z.
//
// FP32 scaling: the input is divided by max(rank s0) and every rank by the
// lattice maximum of g, so no rank exceeds 1 and at most nr ranks decay
// between renormalisations; the logs of the divisors go to lnv.
#include <hip/hip_ext.h>

#include <algorithm>

#include "sfx_block3.h"
#include "sfx_device.h"

namespace sfx {

constexpr int kTX = 8, kTY = 8, kTZ = 16;
constexpr int kH = kB3Ranks;  // largest widening: the lattice-side bound of block3_ok
// the tile widened by h (the most ranks a launch of the variant advances)
template <typename T, bool TANGENT>
struct B3Region {
  static constexpr int h = b3_max_ranks<T>(TANGENT);
  static constexpr int rx = kTX + 2 * h, ry = kTY + 2 * h, rz = kTZ + 2 * h;
  static constexpr int ryz = ry * rz;
  static constexpr int r = rx * ryz;  // region sites
};
constexpr int kT = kTX * kTY * kTZ;
constexpr int kB3 = 1024;  // threads per workgroup, one per tile site
static_assert(kB3 == kT, "a thread owns one tile site");

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
// FP32 tangent variants fit two workgroups per CU in LDS; the bound asks the
// compiler for the register budget (64 VGPRs) that lets them (the FP64 base
// variants, also two per CU, stay below it unasked)
__global__ void __launch_bounds__(kB3) __attribute__((amdgpu_waves_per_eu((TANGENT && sizeof(T) == 4) ? 8 : 1)))
k_block3(DevLattice L, B3Launch<T> la, int nb) {
  constexpr int kH = B3Region<T, TANGENT>::h;
  constexpr int kRZ = B3Region<T, TANGENT>::rz, kRYZ = B3Region<T, TANGENT>::ryz, kR = B3Region<T, TANGENT>::r;
  // chain of this workgroup and its tile number within that chain's grid
  const int chain = (int)blockIdx.x >= nb ? 1 : 0;
  const int bl = (int)blockIdx.x - chain * nb;
  const B3Args<T>& a = la.a[chain];
  extern __shared__ __align__(16) unsigned char smem[];
  T* G0 = (T*)smem;
  T* G1 = G0 + kR;
  T* gl = G1 + kR;
  T* D0 = TANGENT ? gl + kR : nullptr;
  T* D1 = TANGENT ? D0 + kR : nullptr;
  T* vl = TANGENT ? D1 + kR : nullptr;
  __shared__ float red[kB3 / 64];

  const int t = threadIdx.x;
  const int nbz = (L.nz + kTZ - 1) / kTZ, nby = (L.ny + kTY - 1) / kTY;
  // XCD-aware: each XCD takes a contiguous run of tiles (x-slabs), so the y / z
  // halos its tiles share stay in that XCD's L2
  int b = (int)xcd_tile(bl, nb);
  const int bz = b % nbz;
  b /= nbz;
  const int by = b % nby;
  const int bx = b / nby;
  const int x0 = 1 + bx * kTX, y0 = 1 + by * kTY, z0 = 1 + bz * kTZ;  // tile origin (interior coords)
  const int nr = a.nr;

  // the tile site this thread owns (kB3 == kT): z fastest, so HBM accesses coalesce
  const int oZ = t % kTZ, oY = (t / kTZ) % kTY, oX = t / (kTZ * kTY);
  const bool owner = x0 + oX <= L.nx && y0 + oY <= L.ny && z0 + oZ <= L.nz;
  const long long ogi = owner ? sidx(L, x0 + oX, y0 + oY, z0 + oZ) : 0;
  const int oidx = (kH + oX) * kRYZ + (kH + oY) * kRZ + (kH + oZ);

  // the logs the scale factors build on were written by earlier launches: load
  // them now, so their latency overlaps the region loads below
  const int lane = t & 63;
  double lnv0 = 0.0, lnp = 0.0, lref = 0.0;
  if constexpr (SCALED) {
    lnv0 = a.lnv[a.s0];
    if constexpr (CONNECT) {
      int pi = 0;
#pragma unroll
      for (int k = 0; k < kB3Ranks; k++)
        if (lane == k) pi = a.pidx[k];
      if (lane < nr) lnp = a.lnpair[pi];
      lref = *a.ref;
    }
  }

  // issue the loads of rank s0 (and its tangent), g (and v) on the tile widened
  // by nr first: the scale factors below are computed while they are in flight,
  // and no LDS store waits on them before that
  constexpr int kIt = (kR + kB3 - 1) / kB3;
  const int wlo = kH - nr, wy = kTY + 2 * nr, wz = kTZ + 2 * nr;
  const int nreg = (kTX + 2 * nr) * wy * wz;
  T vg[kIt], vgg[kIt], vd[kIt], vv[kIt];
  int li[kIt];
#pragma unroll
  for (int j = 0; j < kIt; j++) {
    const int i = t + j * kB3;
    li[j] = -1;
    if (i < nreg) {
      const int Z = i % wz, r = i / wz, Y = r % wy, X = r / wy;
      const int x = extend(x0 + wlo + X - kH, L.nx, L.periodic[0], L.mirror2_1[0], L.mirror2_2[0]);
      const int y = extend(y0 + wlo + Y - kH, L.ny, L.periodic[1], L.mirror2_1[1], L.mirror2_2[1]);
      const int z = extend(z0 + wlo + Z - kH, L.nz, L.periodic[2], L.mirror2_1[2], L.mirror2_2[2]);
      const long long gi = sidx(L, x, y, z);
      li[j] = (wlo + X) * kRYZ + (wlo + Y) * kRZ + (wlo + Z);
      vg[j] = a.in[gi];
      vgg[j] = a.g[gi];
      if constexpr (TANGENT) {
        vd[j] = a.din[gi];
        vv[j] = a.v[gi];
      }
    }
  }

  // the partner ranks are needed only from the first composition on
  T pr[kB3Ranks], dpr[kB3Ranks];
  if constexpr (CONNECT) {
#pragma unroll
    for (int k = 0; k < kB3Ranks; k++) {
      pr[k] = (owner && k < nr) ? a.pair[k][ogi] : T(0);
      if constexpr (TANGENT) dpr[k] = (owner && k < nr) ? a.dpair[k][ogi] : T(0);
    }
  }

  // scale factors of this launch. They are uniform, so every wave evaluates
  // each FP64 log / exp once in one lane and broadcasts it: lanes 0 and 1 the
  // logs of the two divisors, lane k the composition factor of rank k + 1.
  // The FP64 logs and exponentials are evaluated by the first wave alone,
  // which shares the composition factors through LDS (read after the first
  // rank's barrier); the other waves go on to the LDS stores meanwhile.
  T in_scale = T(1), ginv = T(1);
  double ln_base = 0.0, lng = 0.0;
  __shared__ T cfac[kB3Ranks];
  if constexpr (SCALED) {
    float m0, gm;
    block_step_max2(a.maxv, a.s0, a.gmax, 0, m0, gm);
    if (!(m0 > 0.f)) m0 = 1.f;
    if (!(gm > 0.f)) gm = 1.f;
    in_scale = (T)(1.f / m0);
    ginv = (T)(1.f / gm);
    if (t < 64) {
      const double lv = lane == 0 ? log((double)m0) : (lane == 1 ? log((double)gm) : 0.0);
      ln_base = lnv0 + __shfl(lv, 0, 64);
      lng = __shfl(lv, 1, 64);
      if constexpr (CONNECT) {
        const double c = lane < nr ? exp(ln_base + (lane + 1) * lng + lnp - lref) : 1.0;
        if (lane < kB3Ranks) cfac[lane] = (T)c;
      }
    }
  } else if constexpr (CONNECT) {
    if (t < kB3Ranks) cfac[t] = T(1);
  }

#pragma unroll
  for (int j = 0; j < kIt; j++) {
    if (li[j] >= 0) {
      G0[li[j]] = vg[j] * in_scale;
      gl[li[j]] = vgg[j];
      if constexpr (TANGENT) {
        D0[li[j]] = vd[j] * in_scale;
        vl[li[j]] = vv[j];
      }
    }
  }
  __syncthreads();

  T* src = G0;
  T* dst = G1;
  T* dsrc = D0;
  T* ddst = D1;
  T acc = T(0), mloc = T(0);
  const T sixth = T(1.0 / 6.0);
  for (int k = 1; k <= nr; k++) {
    // rank k on the tile widened by e = nr - k: threads take (y, z) columns,
    // several threads per column split it along x. Consecutive threads cover
    // whole LDS rows of kRZ (threads on the ring outside the widened tile
    // idle), so a wave reads consecutive LDS words, free of bank conflicts.
    const int e = nr - k;
    const int lo = kH - e;
    const int bxn = kTX + 2 * e, byn = kTY + 2 * e, bzn = kTZ + 2 * e;
    if constexpr (sizeof(T) == 4) {
      // FP32: a work item is four consecutive z sites of one (x, y) row, read
      // and written as float4 (rows start on 16-byte boundaries: kRZ and kRYZ
      // are multiples of 4). Items cover the widened z range rounded out to
      // whole quads; the extra sites at the row ends read in-buffer
      // neighbours, and no later rank reads them. Same additions in the same
      // order as one site at a time, so the same bits.
      static_assert(kRZ % 4 == 0 && kRYZ % 4 == 0, "float4 rows");
      const int c0 = lo >> 2, nch = ((lo + bzn + 3) >> 2) - c0;
      const int items = bxn * byn * nch;
      for (int it = t; it < items; it += kB3) {
        const int ch = c0 + it % nch, r = it / nch;
        const int idx = (lo + r / byn) * kRYZ + (lo + r % byn) * kRZ + ch * 4;
        const float4 c = *(const float4*)&src[idx];
        const float4 xm = *(const float4*)&src[idx - kRYZ], xp = *(const float4*)&src[idx + kRYZ];
        const float4 ym = *(const float4*)&src[idx - kRZ], yp = *(const float4*)&src[idx + kRZ];
        const float zm = src[idx - 1], zp = src[idx + 4];
        const float4 gq = *(const float4*)&gl[idx];
        const float4 sv = make_float4((xm.x + xp.x + ym.x + yp.x + zm + c.y) * sixth,
                                      (xm.y + xp.y + ym.y + yp.y + c.x + c.z) * sixth,
                                      (xm.z + xp.z + ym.z + yp.z + c.y + c.w) * sixth,
                                      (xm.w + xp.w + ym.w + yp.w + c.z + zp) * sixth);
        const float4 gv = make_float4(gq.x * ginv, gq.y * ginv, gq.z * ginv, gq.w * ginv);
        *(float4*)&dst[idx] = make_float4(gv.x * sv.x, gv.y * sv.y, gv.z * sv.z, gv.w * sv.w);
        if constexpr (TANGENT) {
          const float4 dc = *(const float4*)&dsrc[idx];
          const float4 dxm = *(const float4*)&dsrc[idx - kRYZ], dxp = *(const float4*)&dsrc[idx + kRYZ];
          const float4 dym = *(const float4*)&dsrc[idx - kRZ], dyp = *(const float4*)&dsrc[idx + kRZ];
          const float dzm = dsrc[idx - 1], dzp = dsrc[idx + 4];
          const float4 vq = *(const float4*)&vl[idx];
          const float4 ds = make_float4((dxm.x + dxp.x + dym.x + dyp.x + dzm + dc.y) * sixth,
                                        (dxm.y + dxp.y + dym.y + dyp.y + dc.x + dc.z) * sixth,
                                        (dxm.z + dxp.z + dym.z + dyp.z + dc.y + dc.w) * sixth,
                                        (dxm.w + dxp.w + dym.w + dyp.w + dc.z + dzp) * sixth);
          *(float4*)&ddst[idx] = make_float4(gv.x * (ds.x - vq.x * sv.x), gv.y * (ds.y - vq.y * sv.y),
                                             gv.z * (ds.z - vq.z * sv.z), gv.w * (ds.w - vq.w * sv.w));
        }
      }
    } else {
      const int ncol = byn * kRZ;
      const int per = kB3 / ncol > 0 ? kB3 / ncol : 1;
      const int c = t % ncol, xs = t / ncol, RZ = c % kRZ;
      if (t < per * ncol && RZ >= lo && RZ < lo + bzn) {
        const int base = (lo + c / kRZ) * kRZ + RZ;
        for (int X = lo + xs; X < lo + bxn; X += per) {
          const int idx = X * kRYZ + base;
          const T s = (src[idx - kRYZ] + src[idx + kRYZ] + src[idx - kRZ] + src[idx + kRZ] + src[idx - 1] +
                       src[idx + 1]) * sixth;
          const T gv = gl[idx] * ginv;
          dst[idx] = gv * s;
          if constexpr (TANGENT) {
            const T ds = (dsrc[idx - kRYZ] + dsrc[idx + kRYZ] + dsrc[idx - kRZ] + dsrc[idx + kRZ] + dsrc[idx - 1] +
                          dsrc[idx + 1]) * sixth;
            ddst[idx] = gv * (ds - vl[idx] * s);
          }
        }
      }
    }
    __syncthreads();
    // the owned tile site of rank k: store, compose, track the last rank's max
    if (owner) {
      const T val = dst[oidx];
      T dval = T(0);
      if constexpr (TANGENT) dval = ddst[oidx];
      T* outk = a.out[k - 1];
      if (outk) outk[ogi] = val;
      if constexpr (TANGENT) {
        T* doutk = a.dout[k - 1];
        if (doutk) doutk[ogi] = dval;
      }
      if constexpr (CONNECT) {
        if constexpr (TANGENT)
          acc += (dval * pr[k - 1] + val * dpr[k - 1]) * cfac[k - 1];
        else
          acc += val * pr[k - 1] * cfac[k - 1];
      }
      if (k == nr) mloc = val;
    }
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
    if (owner) a.acc[ogi] += acc;
  }
  if constexpr (SCALED) {
    const T m = block_max_T<T>(mloc, red);
    if (t == 0) {
      put_step_max(a.maxv, a.s0 + nr, (float)m);
      if (bl == 0)
        for (int k = 1; k <= nr; k++) a.lnv[a.s0 + k] = ln_base + k * lng;
    }
  }
}

template <typename T, bool TANGENT>
static size_t b3_lds() {
  return (TANGENT ? 6 : 3) * (size_t)B3Region<T, TANGENT>::r * sizeof(T);
}

template <typename T, bool TG, bool SC, bool CN>
static void b3_launch(const DevLattice& L, const B3Launch<T>& a, unsigned nb, hipStream_t st, hipEvent_t ev0,
                      hipEvent_t ev1) {
  static bool attr = false;  // dynamic LDS above the default limit needs the attribute once
  const size_t lds = b3_lds<T, TG>();
  if (!attr) {
    SFX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_block3<T, TG, SC, CN>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  // with events: recorded by the dispatch itself at kernel start and end
  hipExtLaunchKernelGGL(k_block3<T, TG, SC, CN>, dim3(nb * (unsigned)a.nchains), dim3(kB3), (uint32_t)lds, st, ev0,
                        ev1, 0u, L, a, (int)nb);
}

template <typename T>
void launch_block3(const DevLattice& L, const B3Args<T>& a, bool tangent, bool scaled, bool connect,
                   hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  B3Launch<T> la{};
  la.a[0] = a;
  la.nchains = 1;
  launch_block3<T>(L, la, tangent, scaled, connect, st, ev0, ev1);
}

template <typename T>
void launch_block3(const DevLattice& L, const B3Launch<T>& a, bool tangent, bool scaled, bool connect,
                   hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  SFX_REQUIRE(a.nchains == 1 || a.nchains == 2, SFX_E_INVALID, "a blocked launch advances one or two chains");
  for (int c = 0; c < a.nchains; c++)
    SFX_REQUIRE(a.a[c].nr >= 1 && a.a[c].nr <= b3_max_ranks<T>(tangent), SFX_E_INVALID,
                "blocked launch of 1..4 ranks (1..2 for tangents and FP64)");
  SFX_REQUIRE(!(tangent && sizeof(T) == 8), SFX_E_UNSUPPORTED, "FP64 tangent sweeps run rank by rank");
  const size_t lds_base = b3_lds<T, false>(), lds_tan = b3_lds<T, true>();
  SFX_REQUIRE((lds_base <= 160 * 1024 && (sizeof(T) == 8 || lds_tan <= 160 * 1024)), SFX_E_UNSUPPORTED,
              "blocked tile exceeds LDS");
  const unsigned nb = (unsigned)(((L.nx + kTX - 1) / kTX) * ((L.ny + kTY - 1) / kTY) * ((L.nz + kTZ - 1) / kTZ));
#define B3(TG, SC, CN) b3_launch<T, TG, SC, CN>(L, a, nb, st, ev0, ev1)
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

// max of each state's FP32 SWF over the interior: 64 workgroups per state,
// one atomic each on the state's word (zeroed first)
__global__ void k_gmax_zero(unsigned* gmax, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gmax[i] = 0u;
}
__global__ void __launch_bounds__(256) k_gmax(DevLattice L, const float* __restrict__ swf32, unsigned* __restrict__ gmax) {
  const float* g = swf32 + (long long)blockIdx.y * L.M;
  float m = 0.f;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < L.V; q += (long long)gridDim.x * blockDim.x) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    m = fmaxf(m, g[sidx(L, x, y, z)]);
  }
  __shared__ float red[4];
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    put_step_max(gmax + (long long)blockIdx.y * kStepWords, 0, m);  // slot blockIdx.x mod kMaxSlots
  }
}

void launch_gmax(const DevLattice& L, const float* swf32, int ns, unsigned* gmax, hipStream_t st, bool zeroed) {
  const int n = ns * kStepWords;
  if (!zeroed) k_gmax_zero<<<grid1(n, 256), 256, 0, st>>>(gmax, n);
  const unsigned nb = (unsigned)std::min<long long>(256, std::max<long long>(1, L.V / 1024));
  k_gmax<<<dim3(nb, ns), 256, 0, st>>>(L, swf32, gmax);
  SFX_HIP(hipGetLastError());
}

template void launch_block3<float>(const DevLattice&, const B3Args<float>&, bool, bool, bool, hipStream_t,
                                   hipEvent_t, hipEvent_t);
template void launch_block3<double>(const DevLattice&, const B3Args<double>&, bool, bool, bool, hipStream_t,
                                    hipEvent_t, hipEvent_t);
template void launch_block3<float>(const DevLattice&, const B3Launch<float>&, bool, bool, bool, hipStream_t,
                                   hipEvent_t, hipEvent_t);
template void launch_block3<double>(const DevLattice&, const B3Launch<double>&, bool, bool, bool, hipStream_t,
                                    hipEvent_t, hipEvent_t);

}  // namespace sfx
