// sfx_device.h -- device-side lattice helpers shared by the kernel files.
#pragma once

#include "sfx_internal.h"

namespace sfx {

__device__ __forceinline__ long long sidx(const DevLattice& L, int x, int y, int z) {
  return ((long long)x * L.MY + y) * L.MZ + z;
}

// interior ordinal q in [0, V) -> coordinates (last used gradient fastest);
// 32-bit arithmetic (V < 2^31 is checked when the lattice is created)
__device__ __forceinline__ void interior_coords(const DevLattice& L, long long q64, int& x, int& y, int& z) {
  const unsigned q = (unsigned)q64;
  if (L.grad == 1) {
    x = (int)q + 1;
    y = 0;
    z = 0;
  } else if (L.grad == 2) {
    unsigned r = q / (unsigned)L.ny;
    y = (int)(q - r * (unsigned)L.ny) + 1;
    x = (int)r + 1;
    z = 0;
  } else {
    unsigned r = q / (unsigned)L.nz;
    z = (int)(q - r * (unsigned)L.nz) + 1;
    unsigned r2 = r / (unsigned)L.ny;
    y = (int)(r - r2 * (unsigned)L.ny) + 1;
    x = (int)r2 + 1;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X. This
// maps the dispatch index to a tile index so that each XCD walks one
// contiguous range of the lattice (its own L2 then holds the planes the next
// rank reads); a bijection on [0, nb).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned per = nb >> 3;
  if (b >= (per << 3)) return b;
  return (b & 7u) * per + (b >> 3);
}

// thread ordinal over the interior of the current launch (XCD-aware)
__device__ __forceinline__ long long site_ordinal() {
  return (long long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ void site_coords(const DevLattice& L, long long i, int& x, int& y, int& z) {
  z = (int)(i % L.MZ);
  long long r = i / L.MZ;
  y = (int)(r % L.MY);
  x = (int)(r / L.MY);
}

__device__ __forceinline__ bool is_interior(const DevLattice& L, int x, int y, int z) {
  if (x < 1 || x > L.nx) return false;
  if (L.grad >= 2 && (y < 1 || y > L.ny)) return false;
  if (L.grad >= 3 && (z < 1 || z > L.nz)) return false;
  return true;
}

// neighbour coordinate of an interior coordinate c (1..n) shifted by d in {-1,0,1}
__device__ __forceinline__ int nb(int c, int d, int n, int b1, int b2) {
  int v = c + d;
  return v == 0 ? b1 : (v == n + 1 ? b2 : v);
}

// Propagator stencil sum without the segment weight: sum_k lambda_k G(r+e_k).
// Surface neighbours contribute zero. 1-D: Lat1DFlat1stO.cpp:44-46 and
// Lat1DCyl1stO.cpp:22-23; 2-D: Lat2DFlat1stO.cpp:44, Lat2DCyl1stO.cpp:27-30;
// 3-D: Lat3D1stO.cpp:151-157, Lat3DFCC1stO.cpp:77-89, Lat3DHEX1stO.cpp:77-89.
template <typename T, int KIND>
__device__ __forceinline__ T stencil(const DevLattice& L, const T* __restrict__ G, int x, int y, int z) {
  if constexpr (KIND == ST_1D) {
    T b = G[x];
    T a = (x == 1) ? (L.surf1[0] ? T(0) : G[L.b1[0]]) : G[x - 1];
    T c = (x == L.nx) ? (L.surf2[0] ? T(0) : G[L.b2[0]]) : G[x + 1];
    if (L.curved) return b + T(L.lm[x]) * (a - b) + T(L.lp[x]) * (c - b);
    return b + T(L.l1) * (a - b - b + c);
  } else if constexpr (KIND == ST_2D) {
    const long long jx = L.jx;
    long long r = (long long)x * jx + y;
    T b = G[r];
    T a = (y == 1) ? (L.surf1[1] ? T(0) : G[(long long)x * jx + L.b1[1]]) : G[r - 1];
    T c = (y == L.ny) ? (L.surf2[1] ? T(0) : G[(long long)x * jx + L.b2[1]]) : G[r + 1];
    T xm = (x == 1) ? (L.surf1[0] ? T(0) : G[(long long)L.b1[0] * jx + y]) : G[r - jx];
    T xp = (x == L.nx) ? (L.surf2[0] ? T(0) : G[(long long)L.b2[0] * jx + y]) : G[r + jx];
    if (L.curved) return b + T(L.lm[x]) * (xm - b) + T(L.lp[x]) * (xp - b) + T(L.l1) * (a + c - b - b);
    return T(L.l0_2d) * b + T(L.l1y) * (a + c) + T(L.l1x) * (xm + xp);
  } else {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    auto at = [&](int dx, int dy, int dz) -> T {
      int X = nb(x, dx, nx, L.b1[0], L.b2[0]);
      int Y = nb(y, dy, ny, L.b1[1], L.b2[1]);
      int Z = nb(z, dz, nz, L.b1[2], L.b2[2]);
      return G[sidx(L, X, Y, Z)];
    };
    if constexpr (KIND == ST_3D_SC) {
      T s = at(-1, 0, 0) + at(1, 0, 0) + at(0, -1, 0) + at(0, 1, 0) + at(0, 0, -1) + at(0, 0, 1);
      return s * T(1.0 / 6.0);
    } else if constexpr (KIND == ST_3D_FCC) {
      T s = at(-1, -1, 0) + at(1, 1, 0) + at(-1, 1, 0) + at(1, -1, 0) + at(0, -1, -1) + at(0, 1, 1) +
            at(0, -1, 1) + at(0, 1, -1) + at(-1, 0, -1) + at(1, 0, 1) + at(-1, 0, 1) + at(1, 0, -1);
      return s * T(1.0 / 12.0);
    } else {
      T s = at(-1, 0, 0) + at(1, 0, 0) + at(-1, 1, 0) + at(0, 1, 0) + at(0, -1, 0) + at(1, -1, 0) +
            at(0, 1, -1) + at(0, 0, -1) + at(1, 0, -1) + at(-1, 0, 1) + at(0, 0, 1) + at(0, -1, 1);
      return s * T(1.0 / 12.0);
    }
  }
}

// Weight of an interior site in ComputeLnGN (Lat1DFlat1stO.cpp:301-316,
// Lat1DCyl1stO.cpp:258-270, Lat2DFlat1stO.cpp:293-330, Lat2DCyl1stO.cpp:388-438,
// Lat3D1stO.cpp:345-370 via SubtractBoundaries).
__device__ __forceinline__ double lngn_weight(const DevLattice& L, int x, int y, int z) {
  double w = 1.0;
  if (L.grad == 1) {
    if (L.curved) {
      w = L.Lw[x];
      if (x == 1 && L.mirror2_1[0]) w *= 0.5;
      if (x == L.nx && L.mirror2_2[0]) w *= 0.5;
    } else {
      if (x == 1) {
        w = L.ladj;
        if (L.mirror2_1[0]) w *= 0.5;
      } else if (x == L.nx) {
        w = L.mirror2_2[0] ? 0.5 : 2 - L.ladj;
      }
    }
    return w;
  }
  int c[3] = {x, y, z};
  int n[3] = {L.nx, L.ny, L.nz};
  for (int k = 0; k < L.grad; k++) {
    if (c[k] == 1 && L.mirror2_1[k]) w *= 0.5;
    if (c[k] == n[k] && L.mirror2_2[k]) w *= 0.5;
  }
  if (L.curved) w *= L.Lw[x];
  return w;
}

// Weight of a site in the free-energy sum: SubtractBoundaries followed by
// MultiplyWithLatticeSites (Lat1DFlat.cpp:265-269, 409-415; Lat1DCylinder.cpp:126-131;
// Lat2DFlat.cpp SubtractBoundaries; Lat3D.cpp:1591-1640).
__device__ __forceinline__ double fe_weight(const DevLattice& L, int x, int y, int z) {
  int c[3] = {x, y, z};
  int n[3] = {L.nx, L.ny, L.nz};
  double w = 1.0;
  for (int k = 0; k < L.grad; k++) {
    if (c[k] == 0) {
      if (!L.surf1[k] || L.grad == 3) return 0.0;
    } else if (c[k] == n[k] + 1) {
      if (!L.surf2[k] || L.grad == 3) return 0.0;
    } else {
      if (c[k] == 1 && L.mirror2_1[k]) w *= 0.5;
      if (c[k] == n[k] && L.mirror2_2[k]) w *= 0.5;
    }
  }
  if (L.grad == 1) {
    if (L.curved)
      w *= L.Lw[x];
    else if (x == 1)
      w *= L.ladj;
    else if (x == L.nx)
      w *= 2 - L.ladj;
  } else if (L.grad == 2 && L.curved) {
    w *= L.Lw[x];
  }
  return w;
}

// contact averages <phi_t> (SideFraction / Sides) on every site
__device__ inline double side_at(const DevLattice& L, const double* __restrict__ P, int x, int y, int z) {
  // remap halo coordinates that are not surfaces
  if (x == 0 && !L.surf1[0]) x = L.b1[0];
  else if (x == L.nx + 1 && !L.surf2[0]) x = L.b2[0];
  if (L.grad >= 2) {
    if (y == 0 && !L.surf1[1]) y = L.b1[1];
    else if (y == L.ny + 1 && !L.surf2[1]) y = L.b2[1];
  }
  if (L.grad >= 3) {
    if (z == 0) z = L.b1[2];
    else if (z == L.nz + 1) z = L.b2[2];
  }
  auto at = [&](int X, int Y, int Z) -> double {
    X = remap(X, L.nx, L.surf1[0] ? 0 : L.b1[0], L.surf2[0] ? L.nx + 1 : L.b2[0]);
    if (L.grad >= 2) Y = remap(Y, L.ny, L.surf1[1] ? 0 : L.b1[1], L.surf2[1] ? L.ny + 1 : L.b2[1]);
    if (L.grad >= 3) Z = remap(Z, L.nz, L.b1[2], L.b2[2]);
    return P[sidx(L, X, Y, Z)];
  };
  if (L.grad == 1) {
    // Lat1DFlat.cpp:510-524, Lat1DCylinder.cpp:160-171
    const long long M = L.MX;
    if (x == 0) {
      double lp = L.curved ? L.lpc[0] : L.l1chi;
      return P[0] + lp * (-P[0] - P[0] + P[1]);
    }
    if (x == L.nx + 1) {
      double lm = L.curved ? L.lmc[M - 1] : L.l1chi;
      return P[M - 1] + lm * (-P[M - 1] - P[M - 1] + P[M - 2]);
    }
    double b = P[x], a = at(x - 1, 0, 0), c = at(x + 1, 0, 0);
    if (L.curved) return b + L.lpc[x] * (c - b) + L.lmc[x] * (a - b);
    return b + L.l1chi * (a - b - b + c);
  }
  if (L.grad == 2) {
    // Lat2DFlat.cpp:996-1053, Lat2DCylinder.cpp:205-320 (standard lattice, l2 = 0)
    double b = P[sidx(L, x, y, 0)];
    double l1 = L.l1;
    auto lam0 = [&](int X) { return 1 - 2 * l1 - L.lp[X] - L.lm[X]; };
    if (x == 0 || x == L.nx + 1) {
      double a = at(x, y - 1, 0), c = at(x, y + 1, 0);
      double in = x == 0 ? P[sidx(L, 1, y, 0)] : P[sidx(L, L.nx, y, 0)];
      if (L.curved) return lam0(x) * b + l1 * (a + c) + (x == 0 ? L.lp[0] : L.lm[x]) * in;
      return L.l0_2d * b + L.l1y * (a + c) + L.l1x * in;
    }
    if (y == 0 || y == L.ny + 1) {
      double xm = at(x - 1, y, 0), xp = at(x + 1, y, 0);
      double in = y == 0 ? P[sidx(L, x, 1, 0)] : P[sidx(L, x, L.ny, 0)];
      if (L.curved) return lam0(x) * b + L.lp[x] * xp + L.lm[x] * xm + l1 * in;
      return L.l0_2d * b + L.l1x * (xm + xp) + L.l1y * in;
    }
    double xm = at(x - 1, y, 0), xp = at(x + 1, y, 0), a = at(x, y - 1, 0), c = at(x, y + 1, 0);
    if (L.curved) return b + L.lp[x] * (xp - b) + L.lm[x] * (xm - b) + l1 * (a + c - b - b);
    return L.l0_2d * b + L.l1x * (xm + xp) + L.l1y * (a + c);
  }
  // 3-D, Lat3D.cpp:1729-1784
  if (L.kind == ST_3D_HEX) {
    const double lam = 1.0 / 12.0;
    return at(x + 1, y, z) * lam + at(x - 1, y, z) * lam + at(x + 1, y - 1, z) * lam + at(x, y - 1, z) * lam +
           at(x, y + 1, z) * lam + at(x - 1, y + 1, z) * lam + at(x, y - 1, z + 1) * lam +
           at(x, y, z + 1) * lam + at(x - 1, y, z + 1) * lam + at(x + 1, y, z - 1) * lam +
           at(x, y, z - 1) * lam + at(x, y + 1, z - 1) * lam;
  }
  if (L.kind == ST_3D_FCC) {
    const double l = 1.0 / 27.0;
    double s = 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) s += l * at(x + dx, y + dy, z + dz);
    return s;
  }
  const double l1 = 1.0 / 6.0;
  return l1 * at(x + 1, y, z) + l1 * at(x - 1, y, z) + l1 * at(x, y + 1, z) + l1 * at(x, y - 1, z) +
         l1 * at(x, y, z + 1) + l1 * at(x, y, z - 1);
}

}  // namespace sfx
