// sfx_internal.h -- internal types of the gfx950 SCF library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sfx.h"

namespace sfx {

// ---------------------------------------------------------------------------
// errors: a thread-local message plus an int status, no exceptions across the ABI
void set_error(int code, const char* fmt, ...);
int last_code();

struct Error {
  int code;
  std::string msg;
};

#define SFX_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      throw ::sfx::Error{SFX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    }                                                                                   \
  } while (0)

// runs f, turning exceptions into a status code and the thread's error message
template <typename F>
int guarded(F&& f) {
  (void)hipGetLastError();  // an error an earlier call already reported must not resurface here
  try {
    f();
    return SFX_OK;
  } catch (const Error& e) {
    set_error(e.code, "%s", e.msg.c_str());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(SFX_E_OOM, "host allocation failed");
    return SFX_E_OOM;
  } catch (const std::exception& e) {
    set_error(SFX_E_INVALID, "%s", e.what());
    return SFX_E_INVALID;
  }
}

#define SFX_REQUIRE(cond, code, text)                   \
  do {                                                  \
    if (!(cond)) throw ::sfx::Error{(code), (text)};    \
  } while (0)

// ---------------------------------------------------------------------------
// Lattice description handed to kernels by value.
enum StencilKind : int {
  ST_1D = 0,     // 3-point, per-layer lambda tables (flat: constant)
  ST_2D = 1,     // 5-point, flat or cylindrical (per-x radial tables)
  ST_3D_SC = 2,  // 6 neighbours x 1/6
  ST_3D_FCC = 3, // 12 neighbours x 1/12, 27-point contact average
  ST_3D_HEX = 4  // 12 neighbours x 1/12
};

struct DevLattice {
  int grad, kind, curved;
  int nx, ny, nz;     // interior layers (1 for unused gradients)
  int MX, MY, MZ;     // padded extents (1 for unused gradients)
  long long M, jx, jy;
  int b1[3], b2[3];   // halo source coordinate per gradient
  int surf1[3], surf2[3];
  int mirror2_1[3], mirror2_2[3];
  int periodic[3];
  int flip1[3], flip2[3];  // 2nd order: a halo read reverses the axis' components
  int order;
  int transfer_mfma;  // 2nd order, FP32: the direction transfer on the matrix cores (0 no, 1 yes, 2 tangents)
  double l1, l1chi, ladj;
  double l0_2d, l1x, l1y;
  // per-x-layer tables (device), curved lattices only
  const double* lm;
  const double* lp;
  const double* lmc;
  const double* lpc;
  const double* Lw;
  long long V;        // interior sites
};

// Host-side lattice: descriptor + owned device tables + host copies of the tables.
struct Lattice {
  sfx_lattice_desc desc;
  DevLattice d;
  std::vector<double> hL, hlm, hlp, hlmc, hlpc;
  double* dtab = nullptr;  // one allocation for the five tables
  ~Lattice();
};

void lattice_init(Lattice& L, const sfx_lattice_desc& d);

// ---------------------------------------------------------------------------
// small device helpers shared by the kernel files

__host__ __device__ inline int remap(int c, int n, int b1, int b2) {
  return c == 0 ? b1 : (c == n + 1 ? b2 : c);
}

inline unsigned grid1(long long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace sfx

// opaque handle of the per-rank lattice primitives (include/sfx.h)
struct sfx_lattice {
  sfx::Lattice l;
  // reduction partials and one result word of the normalisations, allocated
  // on first use and kept (sfx_lattice_norm_phi_restr)
  double* d_work = nullptr;
  long long nwork = 0;
  ~sfx_lattice() {
    if (d_work) (void)hipFree(d_work);
  }
};
