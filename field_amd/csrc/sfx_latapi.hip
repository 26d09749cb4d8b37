// sfx_latapi.hip -- per-rank lattice primitives of the C ABI, the device-pointer
// counterparts of Cuda_tools.h:3-26 and of the Lat1stO virtual interface
// (Lat1stO.h:19-50). One call = one kernel on the caller's stream.
#include <cmath>
#include <mutex>

#include "sfx_device.h"
#include "sfx_kernels.h"

namespace sfx {

// Scratch of sfx_lattice_lngn_noalloc: the per-workgroup partial sums of
// launch_wsum (one per 256 interior sites) and, behind them, the total. One
// copy per GPU, held until the process ends (512 KiB). Every call on a GPU
// shares it, so calls that overlap on one GPU race; each call overwrites
// every word it reads, so nothing needs clearing between calls.
constexpr int kLngnMaxPartials = 65536;  // lattices of up to 65536 * 256 = 16,777,216 interior sites
static __device__ __align__(256) double scratch[kLngnMaxPartials + 1];

// the device address of `scratch` on each GPU (the host-side symbol is only a
// stand-in), looked up once per device with that device current
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static double* scratch_on_current_device() {
  int dev = 0;
  SFX_HIP(hipGetDevice(&dev));
  SFX_REQUIRE(dev >= 0 && dev < kMaxDevices, SFX_E_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    SFX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SFX_REQUIRE(ptr != nullptr, SFX_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = (double*)ptr;
  }
  return g_scratch_ptr[dev];
}

// total of the partial sums in index order (the order of k_sum_partials),
// written to tot; with lngn != null also ln(total) there
__global__ void k_lngn_finish(const double* __restrict__ partial, int n, double* __restrict__ tot,
                              double* __restrict__ lngn) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < n; i++) s += partial[i];
    *tot = s;
    if (lngn) *lngn = log(s);
  }
}

// SetBoundaries (Lat1DFlat.cpp:396-399, Lat2DFlat.cpp SetBoundaries,
// Lat3D.cpp:1510-1562): halo sites copy their remapped source, surface halos
// keep their value.
__global__ void k_set_boundaries(DevLattice L, double* __restrict__ P) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  if (is_interior(L, x, y, z)) return;
  int sx = remap(x, L.nx, L.b1[0], L.b2[0]);
  int sy = L.grad >= 2 ? remap(y, L.ny, L.b1[1], L.b2[1]) : y;
  int sz = L.grad >= 3 ? remap(z, L.nz, L.b1[2], L.b2[2]) : z;
  long long s = sidx(L, sx, sy, sz);
  if (s != i) P[i] = P[s];
}

__global__ void k_connect_all(long long M, const double* __restrict__ A, const double* __restrict__ B,
                              double* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) out[i] += A[i] * B[i];
}

__global__ void k_div_all(long long M, double* __restrict__ in, const double* __restrict__ G) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  double g = G[i];
  in[i] = g > 0 ? in[i] / g : 0.0;  // tools.cpp:86-90
}

__global__ void k_sides_one(DevLattice L, const double* __restrict__ P, double* __restrict__ out) {
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
  out[i] = (inside || surf) ? side_at(L, P, x, y, z) : 0.0;
}

// ---- force ensemble (Lat1DFlat1stO.cpp:52-68, Lat3D1stO.cpp:94-165,
// Lat3DFCC1stO.cpp:102-145, Lat3DHEX1stO.cpp:102-145) ----------------------
// A force f along the last gradient weights a neighbour one layer below by
// e^f and one layer above by e^-f. 1-D: G = g [b + f1 (a - b) + f_1 (c - b)],
// f1 = l1 e^f / n, f_1 = l1 e^-f / n, n = l1 (e^f + e^-f - 2) + 1. 3-D: the
// stencil sum with those weights times fnorm = 1 / (4 + e^f + e^-f) (simple
// cubic) or 1 / (4 + 4 e^f + 4 e^-f) (FCC and HEX, as the reference has it
// for both).
__global__ void k_prop_force(DevLattice L, const double* __restrict__ Gin, const double* __restrict__ g,
                             double* __restrict__ Gout, double ef, double emf, double w0, double wlo, double whi) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r = sidx(L, x, y, z);
  if (L.grad == 1) {
    const double b = Gin[x];
    const double a = (x == 1) ? (L.surf1[0] ? 0.0 : Gin[L.b1[0]]) : Gin[x - 1];
    const double c = (x == L.nx) ? (L.surf2[0] ? 0.0 : Gin[L.b2[0]]) : Gin[x + 1];
    Gout[r] = g[r] * (b + wlo * (a - b) + whi * (c - b));
    return;
  }
  auto at = [&](int dx, int dy, int dz) -> double {
    return Gin[sidx(L, nb(x, dx, L.nx, L.b1[0], L.b2[0]), nb(y, dy, L.ny, L.b1[1], L.b2[1]),
                    nb(z, dz, L.nz, L.b1[2], L.b2[2]))];
  };
  double s;
  if (L.kind == ST_3D_SC) {
    s = at(-1, 0, 0) + at(1, 0, 0) + at(0, -1, 0) + at(0, 1, 0) + ef * at(0, 0, -1) + emf * at(0, 0, 1);
  } else if (L.kind == ST_3D_FCC) {
    s = at(-1, -1, 0) + at(1, 1, 0) + at(-1, 1, 0) + at(1, -1, 0) + ef * at(0, -1, -1) + emf * at(0, 1, 1) +
        emf * at(0, -1, 1) + ef * at(0, 1, -1) + ef * at(-1, 0, -1) + emf * at(1, 0, 1) + emf * at(-1, 0, 1) +
        ef * at(1, 0, -1);
  } else {  // HEX
    s = at(-1, 0, 0) + at(1, 0, 0) + at(-1, 1, 0) + at(0, 1, 0) + at(0, -1, 0) + at(1, -1, 0) +
        ef * at(0, 1, -1) + ef * at(0, 0, -1) + ef * at(1, 0, -1) + emf * at(-1, 0, 1) + emf * at(0, 0, 1) +
        emf * at(0, -1, 1);
  }
  Gout[r] = g[r] * s * w0;
}

// ---- second order, 3-D (Lat3D2ndO.cpp), six direction components -------
// Component k of a site is the weight of a bond into it along e_k,
// e = {+x, +y, +z, -z, -y, -x}; 5 - k is the reversal of k. Transfer
// T_kk = S, T_k,5-k = 0, otherwise (1 - S) / 4 (Lat3D2ndO.cpp:70).
__device__ __forceinline__ int axis_of(int k) { return (k == 0 || k == 5) ? 0 : (k == 1 || k == 4) ? 1 : 2; }

// component k of a six-component array at (x, y, z), which may lie in a halo:
// the halo reads its source site, and a mirror side (not a periodic one)
// reverses the components of its axis (SetBx1(P, Q) etc., Lat3D.cpp:931-972)
__device__ __forceinline__ double comp6(const DevLattice& L, const double* __restrict__ P, int k, int x, int y,
                                        int z) {
  int c[3] = {x, y, z};
  const int n[3] = {L.nx, L.ny, L.nz};
  for (int a = 0; a < 3; a++) {
    if (c[a] == 0) {
      if (!L.periodic[a] && axis_of(k) == a) k = 5 - k;
      c[a] = L.b1[a];
    } else if (c[a] == n[a] + 1) {
      if (!L.periodic[a] && axis_of(k) == a) k = 5 - k;
      c[a] = L.b2[a];
    }
  }
  return P[(long long)k * L.M + sidx(L, c[0], c[1], c[2])];
}

__device__ __forceinline__ double t2(int k, int j, double S, double Lw) {
  return k == j ? S : (j == 5 - k ? 0.0 : Lw);
}

// PropagateF (Lat3D2ndO.cpp:64-152): out^k(r) = g(r) sum_j T_kj in^j(r - e_k);
// PropagateB (Lat3D2ndO.cpp:597-678): out^k(r) = g(r) sum_j T_jk in^j(r + e_j)
__global__ void k_prop2_rank(DevLattice L, const double* __restrict__ Gin, const double* __restrict__ g,
                             double* __restrict__ Gout, double S, int backward) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.V) return;
  int x, y, z;
  interior_coords(L, q, x, y, z);
  const long long r = sidx(L, x, y, z);
  const int dx[6] = {1, 0, 0, 0, 0, -1}, dy[6] = {0, 1, 0, 0, -1, 0}, dz[6] = {0, 0, 1, -1, 0, 0};
  const double Lw = (1.0 - S) / 4.0, gv = g[r];
  for (int k = 0; k < 6; k++) {
    double s = 0;
    for (int j = 0; j < 6; j++) {
      if (!backward)
        s += comp6(L, Gin, j, x - dx[k], y - dy[k], z - dz[k]) * (gv * t2(k, j, S, Lw));
      else
        s += comp6(L, Gin, j, x + dx[j], y + dy[j], z + dz[j]) * (gv * t2(j, k, S, Lw));
    }
    Gout[(long long)k * L.M + r] = s;
  }
}

// halo sites of all six components from their sources, with mirror reversal
__global__ void k_setb6(DevLattice L, double* __restrict__ P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.M) return;
  int x, y, z;
  site_coords(L, i, x, y, z);
  if (is_interior(L, x, y, z)) return;
  double v[6];
  for (int k = 0; k < 6; k++) v[k] = comp6(L, P, k, x, y, z);
  for (int k = 0; k < 6; k++) P[(long long)k * L.M + i] = v[k];
}

// ConnectG with PHI (Lat3D2ndO.cpp:895-955): out += sum_k A^k B^k on every
// site; OUT (3 M, may be null) += A^k B^k per axis of k
__global__ void k_connect2(long long M, const double* __restrict__ A, const double* __restrict__ B,
                           double* __restrict__ out, double* __restrict__ OUT) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  double s = out[i];
  double ax[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 6; k++) {
    const double p = A[(long long)k * M + i] * B[(long long)k * M + i];
    s += p;
    ax[axis_of(k)] += p;
  }
  out[i] = s;
  if (OUT)
    for (int a = 0; a < 3; a++) OUT[(long long)a * M + i] += ax[a];
}

// in /= G where G > 0, else 0, on in and on the three axis parts of IN
// (CorrectDoubleCountG(in, IN, G), Lat3D2ndO.cpp:988-994; tools.cpp:86-90)
__global__ void k_div_axes(long long M, double* __restrict__ in, double* __restrict__ IN, const double* __restrict__ G) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double gv = G[i];
  in[i] = gv > 0 ? in[i] / gv : 0.0;
  if (IN)
    for (int a = 0; a < 3; a++) {
      double& v = IN[(long long)a * M + i];
      v = gv > 0 ? v / gv : 0.0;
    }
}

// sum of w(r) sum_k G^k(r) over the interior, per block (six components of a
// 3-D second-order array; ComputeLnGN, Lat3D2ndO.cpp:996-1025)
__global__ void k_wsum6(DevLattice L, const double* __restrict__ G, double* __restrict__ partial) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (q < L.V) {
    int x, y, z;
    interior_coords(L, q, x, y, z);
    const long long r = sidx(L, x, y, z);
    double s = 0;
    for (int k = 0; k < 6; k++) s += G[(long long)k * L.M + r];
    v = s * lngn_weight(L, x, y, z);
  }
  __shared__ double red[256 / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// phi *= C on every site (and the three axis parts of PHI); with partial
// != null, C / (total of partial) instead (the restricted normalisation)
__global__ void k_norm_phi(long long M, double* __restrict__ phi, double* __restrict__ PHI, double C,
                           const double* __restrict__ total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double c = total ? C / *total : C;
  phi[i] *= c;
  if (PHI)
    for (int a = 0; a < 3; a++) PHI[(long long)a * M + i] *= c;
}

static bool order2_3d(const DevLattice& L) { return L.order == 2 && L.grad == 3; }

}  // namespace sfx

using namespace sfx;



static hipStream_t S(void* st) { return (hipStream_t)st; }

// a device buffer that lives for one call
struct DevBufD {
  double* p = nullptr;
  explicit DevBufD(size_t n) {
    if (hipMalloc(&p, (n ? n : 1) * sizeof(double)) != hipSuccess) throw Error{SFX_E_OOM, "hipMalloc failed"};
  }
  ~DevBufD() {
    if (p) (void)hipFree(p);
  }
};


extern "C" {

int sfx_lattice_propagate(const sfx_lattice* l, const double* Gin, const double* G, double* Gout, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && Gin && G && Gout, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(l->l.d.grad >= 1 && l->l.desc.order != 2, SFX_E_UNSUPPORTED, "first-order lattices only");
    RankScale none{nullptr, nullptr, nullptr, nullptr, 0, 0};
    launch_prop<double, double>(l->l.d, Gin, G, Gout, nullptr, nullptr, none, S(stream));
    k_set_boundaries<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d, Gout);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_connect(const sfx_lattice* l, const double* A, const double* B, double* out, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && A && B && out, SFX_E_INVALID, "NULL argument");
    k_connect_all<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d.M, A, B, out);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_correct_double_count(const sfx_lattice* l, double* in, const double* G, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && in && G, SFX_E_INVALID, "NULL argument");
    k_div_all<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d.M, in, G);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_lngn(const sfx_lattice* l, const double* G, double* lngn, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && G && lngn, SFX_E_INVALID, "NULL argument");
    const DevLattice& L = l->l.d;
    if (L.order == 2) {  // ln(sum_k sum_interior G^k / 6) of a six-component array (Lat3D2ndO.cpp:996-1025)
      SFX_REQUIRE(order2_3d(L), SFX_E_UNSUPPORTED,
                  "second-order per-rank primitives take the 3-D six-component layout only");
      const unsigned nb = grid1(L.V, 256);
      DevBufD partial(nb), out(1);
      k_wsum6<<<nb, 256, 0, S(stream)>>>(L, G, partial.p);
      launch_sum_partials(partial.p, (int)nb, out.p, S(stream));
      double s = 0;
      SFX_HIP(hipMemcpyAsync(&s, out.p, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
      SFX_HIP(hipStreamSynchronize(S(stream)));
      *lngn = std::log(s / 6.0);
      return;
    }
    double* partial = nullptr;
    double* out = nullptr;
    unsigned nb = grid1(L.V, 256);
    SFX_HIP(hipMalloc(&partial, nb * sizeof(double)));
    SFX_HIP(hipMalloc(&out, sizeof(double)));
    launch_wsum<double>(L, G, partial, S(stream));
    launch_sum_partials(partial, (int)nb, out, S(stream));
    double s = 0;
    SFX_HIP(hipMemcpyAsync(&s, out, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
    SFX_HIP(hipStreamSynchronize(S(stream)));
    (void)hipFree(partial);
    (void)hipFree(out);
    *lngn = std::log(s);
  });
}

int sfx_lattice_lngn_noalloc(const sfx_lattice* l, const double* G, double* lngn, int where, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && G && lngn, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(where == SFX_HOST || where == SFX_DEVICE, SFX_E_INVALID, "where must be SFX_HOST or SFX_DEVICE");
    const DevLattice& L = l->l.d;
    const unsigned nb = grid1(L.V, 256);
    SFX_REQUIRE(nb <= (unsigned)kLngnMaxPartials, SFX_E_INVALID,
                "lattice too large for the allocation-free ln G_N (at most 16,777,216 interior sites)");
    double* ptr = scratch_on_current_device();
    double* total = ptr + kLngnMaxPartials;
    launch_wsum<double>(L, G, ptr, S(stream));
    k_lngn_finish<<<1, 64, 0, S(stream)>>>(ptr, (int)nb, total, where == SFX_DEVICE ? lngn : nullptr);
    SFX_HIP(hipGetLastError());
    if (where == SFX_HOST) {
      double s = 0;
      SFX_HIP(hipMemcpyAsync(&s, total, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
      SFX_HIP(hipStreamSynchronize(S(stream)));
      *lngn = std::log(s);
    }
  });
}

int sfx_lattice_propagate_force(const sfx_lattice* l, const double* Gin, const double* G, double* Gout, double f,
                                void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && Gin && G && Gout, SFX_E_INVALID, "NULL argument");
    const DevLattice& L = l->l.d;
    SFX_REQUIRE(L.order == 1, SFX_E_UNSUPPORTED, "first-order lattices only");
    SFX_REQUIRE(std::isfinite(f), SFX_E_INVALID, "the force must be finite");
    const double ef = std::exp(f), emf = std::exp(-f);
    double w0 = 1.0, wlo = 0.0, whi = 0.0;
    if (L.grad == 1) {
      SFX_REQUIRE(!L.curved, SFX_E_UNSUPPORTED,
                  "the force ensemble is built for the flat 1-D lattice (Lat1DFlat1stO.cpp:52-68)");
      const double n = L.l1 * (emf + ef - 2.0) + 1.0;
      wlo = L.l1 * ef / n;
      whi = L.l1 * emf / n;
    } else {
      SFX_REQUIRE(L.grad == 3, SFX_E_UNSUPPORTED,
                  "the reference has no force ensemble on 2-D lattices (Lat2DFlat1stO.cpp:55-57)");
      w0 = L.kind == ST_3D_SC ? 1.0 / (4.0 + ef + emf) : 1.0 / (4.0 + 4.0 * ef + 4.0 * emf);
    }
    k_prop_force<<<grid1(L.V, 256), 256, 0, S(stream)>>>(L, Gin, G, Gout, ef, emf, w0, wlo, whi);
    k_set_boundaries<<<grid1(L.M, 256), 256, 0, S(stream)>>>(L, Gout);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_propagate2(const sfx_lattice* l, const double* Gin, const double* G, double* Gout, double S_,
                           int backward, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && Gin && G && Gout, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(Gin != Gout, SFX_E_INVALID, "Gin and Gout must not overlap");
    const DevLattice& L = l->l.d;
    SFX_REQUIRE(order2_3d(L), SFX_E_UNSUPPORTED,
                "second-order per-rank primitives take the 3-D six-component layout only");
    SFX_REQUIRE(std::isfinite(S_) && S_ >= 0.0 && S_ <= 1.0, SFX_E_INVALID, "S must be in [0, 1]");
    k_prop2_rank<<<grid1(L.V, 256), 256, 0, S(stream)>>>(L, Gin, G, Gout, S_, backward ? 1 : 0);
    k_setb6<<<grid1(L.M, 256), 256, 0, S(stream)>>>(L, Gout);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_connect2(const sfx_lattice* l, const double* A, const double* B, double* out, double* OUT,
                         void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && A && B && out, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(order2_3d(l->l.d), SFX_E_UNSUPPORTED,
                "second-order per-rank primitives take the 3-D six-component layout only");
    k_connect2<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d.M, A, B, out, OUT);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_correct_double_count2(const sfx_lattice* l, double* in, double* IN, const double* G, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && in && G, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(order2_3d(l->l.d), SFX_E_UNSUPPORTED,
                "second-order per-rank primitives take the 3-D six-component layout only");
    k_div_axes<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d.M, in, IN, G);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_norm_phi_free(const sfx_lattice* l, double* phi, double* PHI, double C, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && phi, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(!PHI || order2_3d(l->l.d), SFX_E_INVALID, "PHI is for 3-D second-order lattices only");
    k_norm_phi<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d.M, phi, PHI, C, nullptr);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_norm_phi_restr(sfx_lattice* l, double* phi, double* PHI, const double* Gi, double C, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && phi && Gi, SFX_E_INVALID, "NULL argument");
    const DevLattice& L = l->l.d;
    SFX_REQUIRE(!PHI || order2_3d(L), SFX_E_INVALID, "PHI is for 3-D second-order lattices only");
    SFX_REQUIRE(L.order == 1 || order2_3d(L), SFX_E_UNSUPPORTED,
                "second-order per-rank primitives take the 3-D six-component layout only");
    const unsigned nb = grid1(L.V, 256);
    if (l->nwork < (long long)nb + 1) {
      if (l->d_work) SFX_HIP(hipFree(l->d_work));
      l->d_work = nullptr;
      SFX_HIP(hipMalloc(&l->d_work, ((size_t)nb + 1) * sizeof(double)));
      l->nwork = (long long)nb + 1;
    }
    // first order: C / exp(ComputeLnGN(Gi)) = C / sum w Gi (Lat1DFlat1stO.cpp:324-331);
    // second order: C / (6 exp(ComputeLnGN(Gi))) = C / sum_k sum w Gi^k (Lat3D2ndO.cpp:1036-1047)
    if (L.order == 2)
      k_wsum6<<<nb, 256, 0, S(stream)>>>(L, Gi, l->d_work);
    else
      launch_wsum<double>(L, Gi, l->d_work, S(stream));
    launch_sum_partials(l->d_work, (int)nb, l->d_work + nb, S(stream));
    k_norm_phi<<<grid1(L.M, 256), 256, 0, S(stream)>>>(L.M, phi, PHI, C, l->d_work + nb);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_set_boundaries(const sfx_lattice* l, double* P, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && P, SFX_E_INVALID, "NULL argument");
    k_set_boundaries<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d, P);
    SFX_HIP(hipGetLastError());
  });
}

int sfx_lattice_sides(const sfx_lattice* l, const double* P, double* out, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(l && P && out, SFX_E_INVALID, "NULL argument");
    k_sides_one<<<grid1(l->l.d.M, 256), 256, 0, S(stream)>>>(l->l.d, P, out);
    SFX_HIP(hipGetLastError());
  });
}

}  // extern "C"
