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

}  // namespace sfx

using namespace sfx;



static hipStream_t S(void* st) { return (hipStream_t)st; }


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
