// sfx_lattice.hip -- lattice descriptors and per-rank lattice primitives on gfx950.
//
// Boundary encoding follows Lat1DFlat.cpp:26-88 (1-D), Lat2DFlat.cpp:35-146 (2-D)
// and Lat3D.cpp:69-208 (3-D): a halo layer takes the values of interior layer
// 1 (mirror1), 2 (mirror2) or n (periodic) at the lower side, n, n-1 or 1 at
// the upper side; a surface halo keeps its own value. Kernels never read a
// halo: a neighbour coordinate that falls in a halo is remapped to its source
// (independently per gradient, which is what SetBoundaries does for faces,
// edges and corners, Lat3D.cpp:1510-1562), and a surface neighbour of a
// propagator contributes zero (its statistical weight is zero).
#include <cmath>

#include "sfx_internal.h"

namespace sfx {

static int lower_src(int code, int n) {
  switch (code) {
    case SFX_BOUND_SURFACE: return 0;
    case SFX_BOUND_MIRROR1: return 1;
    case SFX_BOUND_MIRROR2: return 2;
    case SFX_BOUND_PERIODIC: return n;
  }
  return -1;
}
static int upper_src(int code, int n) {
  switch (code) {
    case SFX_BOUND_SURFACE: return n + 1;
    case SFX_BOUND_MIRROR1: return n;
    case SFX_BOUND_MIRROR2: return n - 1;
    case SFX_BOUND_PERIODIC: return 1;
  }
  return -1;
}

Lattice::~Lattice() {
  if (dtab) (void)hipFree(dtab);
}

// Curved 1-D tables: Lat1DCylinder.cpp:370-399 (D = dimensions-1 = 2) and
// Lat1DSphere.cpp:240-270 (D = 3). 2-D cylinder radial tables:
// Lat2DCylinder.cpp:35-55 (lambda11[MX] = 0).
static void curved_tables(Lattice& L, int sphere, int two_d) {
  const int M = L.d.MX;
  const double off = L.desc.offset_first_layer, adj = L.d.ladj;
  const double l1 = L.d.l1, l1c = L.d.l1chi;
  std::vector<double> V(M), S(M);
  L.hL.assign(M, 0);
  L.hlm.assign(M, 0);
  L.hlp.assign(M, 0);
  L.hlmc.assign(M, 0);
  L.hlpc.assign(M, 0);
  const double dims = 3.0;
  const double D = sphere ? dims : dims - 1;
  if (sphere) {
    S[0] = off * off * M_PI * 4;
    V[0] = (4.0 / 3.0) * std::pow(off, 3) * M_PI;
  } else {
    S[0] = off * M_PI * 2;
    V[0] = off * off * M_PI;
  }
  L.hL[0] = V[0];
  const bool adjust_outer = !two_d && L.desc.layer_adjustment > 0 && L.desc.layer_adjustment != 1.0;
  for (int z = 2; z <= M; z++) {
    int i = z - 1;
    double x = z + off - 2 + adj;
    if (z == M - 1 && adjust_outer) x += -adj + 1;
    V[i] = (M_PI * std::pow(2, (D - 1)) / D) * std::pow(x, D);
    L.hL[i] = V[i] - V[i - 1];
    S[i] = std::pow(2, (D - 1)) * M_PI * std::pow(x, (D - 1));
    L.hlm[i] = l1 * S[i - 1] / L.hL[i];
    L.hlp[i] = l1 * S[i] / L.hL[i];
    L.hlmc[i] = l1c * S[i - 1] / L.hL[i];
    L.hlpc[i] = l1c * S[i] / L.hL[i];
  }
  if (two_d) {
    L.hlp[M - 1] = 0;
    L.hlpc[M - 1] = 0;
  }
  if (sphere) {
    if (L.hL[0] > 4.0 * M_PI / 3.0) L.hL[0] = V[0] - std::pow((off - 1), 3) * M_PI * 4.0 / 3.0;
  } else {
    if (L.hL[0] > M_PI) L.hL[0] = V[0] - (off - 1) * (off - 1) * M_PI;
  }
  if (L.hL[0] > 0) {
    L.hlp[0] = L.hL[1] * L.hlm[1] / L.hL[0];
    L.hlpc[0] = L.hL[1] * L.hlmc[1] / L.hL[0];
  }
}

void lattice_init(Lattice& L, const sfx_lattice_desc& d) {
  SFX_REQUIRE(d.gradients >= 1 && d.gradients <= 3, SFX_E_INVALID, "gradients must be 1, 2 or 3");
  SFX_REQUIRE(d.order == 0 || d.order == 1 || d.order == 2, SFX_E_INVALID, "order must be 1 or 2");
  L.desc = d;
  DevLattice& D = L.d;
  D = DevLattice{};
  D.grad = d.gradients;
  int n[3] = {1, 1, 1};
  for (int k = 0; k < D.grad; k++) {
    SFX_REQUIRE(d.n_layers[k] >= 2, SFX_E_INVALID, "n_layers must be >= 2 in every gradient");
    n[k] = d.n_layers[k];
  }
  D.nx = n[0];
  D.ny = n[1];
  D.nz = n[2];
  D.MX = D.nx + 2;
  D.MY = D.grad >= 2 ? D.ny + 2 : 1;
  D.MZ = D.grad >= 3 ? D.nz + 2 : 1;
  D.M = (long long)D.MX * D.MY * D.MZ;
  D.jx = (long long)D.MY * D.MZ;
  D.jy = D.MZ;
  D.V = (long long)D.nx * D.ny * D.nz;
  SFX_REQUIRE(D.M < (1LL << 31), SFX_E_UNSUPPORTED, "lattices are limited to 2^31 sites");
  for (int k = 0; k < 3; k++) {
    D.b1[k] = D.b2[k] = 0;
    if (k >= D.grad) continue;
    int lo = d.lowerbound[k], up = d.upperbound[k];
    SFX_REQUIRE(lo >= 1 && lo <= 4 && up >= 1 && up <= 4, SFX_E_INVALID, "bad boundary code");
    if (lo == SFX_BOUND_PERIODIC || up == SFX_BOUND_PERIODIC) lo = up = SFX_BOUND_PERIODIC;
    if (D.grad == 3)
      SFX_REQUIRE(lo != SFX_BOUND_SURFACE && up != SFX_BOUND_SURFACE, SFX_E_INVALID,
                  "3-D lattices take mirror1, mirror2 or periodic bounds (Lat3D.cpp:69-73)");
    D.b1[k] = lower_src(lo, n[k]);
    D.b2[k] = upper_src(up, n[k]);
    D.surf1[k] = lo == SFX_BOUND_SURFACE;
    D.surf2[k] = up == SFX_BOUND_SURFACE;
    D.mirror2_1[k] = lo == SFX_BOUND_MIRROR2;
    D.mirror2_2[k] = up == SFX_BOUND_MIRROR2;
    D.periodic[k] = lo == SFX_BOUND_PERIODIC;
    // second-order halos (SetBx1(P,Q) / SetBxm(P,Q) and their y, z forms):
    // 3-D swaps the +/- components at every non-periodic bound
    // (Lat3D.cpp:1276-1290, 1428-1440); 1-D and 2-D swap at the lower bound
    // only for surface and mirror1 (`bound1 > 2` copies plainly,
    // Lat1DFlat.cpp:658-666, Lat2DFlat.cpp:1343-1393) and at the upper bound
    // for every non-periodic choice (Lat1DFlat.cpp:673-682, Lat2DFlat.cpp:1363-1410)
    D.flip1[k] = D.grad == 3 ? lo != SFX_BOUND_PERIODIC : (lo == SFX_BOUND_SURFACE || lo == SFX_BOUND_MIRROR1);
    D.flip2[k] = up != SFX_BOUND_PERIODIC;
  }
  D.order = d.order == 2 ? 2 : 1;
  if (D.order == 2)
    SFX_REQUIRE(d.lattice_type == SFX_LAT_STANDARD, SFX_E_UNSUPPORTED,
                "second-order chains run on standard lattices (Lat1D2ndO, Lat1DCyl2ndO, Lat1DSphere2ndO, "
                "Lat2D2ndO, Lat2DCyl2ndO, Lat3D2ndO)");
  D.ladj = d.layer_adjustment > 0 ? d.layer_adjustment : 1.0;
  D.l1 = d.lambda;
  D.l1chi = d.lambda_chi > 0 ? d.lambda_chi : d.lambda;
  if (D.grad == 3) {
    D.kind = d.lattice_type == SFX_LAT_FCC ? ST_3D_FCC : d.lattice_type == SFX_LAT_HEX ? ST_3D_HEX : ST_3D_SC;
    D.l1 = 1.0 / 6.0;  // Lat3D.cpp:231-236
    D.l1chi = D.l1;
    SFX_REQUIRE(d.geometry == SFX_FLAT, SFX_E_UNSUPPORTED, "3-D lattices are flat");
  } else if (D.grad == 2) {
    D.kind = ST_2D;
    SFX_REQUIRE(d.lambda > 0 && d.lambda <= 0.25, SFX_E_INVALID, "2-D lambda must be in (0, 1/4]");
    D.l1x = D.l1y = D.l1;
    D.l0_2d = 1 - 4 * D.l1;
    D.l1chi = D.l1;
    SFX_REQUIRE(d.geometry == SFX_FLAT || d.geometry == SFX_CYLINDER, SFX_E_UNSUPPORTED,
                "2-D lattices are flat or cylindrical");
  } else {
    D.kind = ST_1D;
    SFX_REQUIRE(d.lambda > 0 && d.lambda <= 0.5, SFX_E_INVALID,
                "1-D lambda must be set in (0, 1/2] (Lat1DFlat.cpp:109)");
  }
  D.curved = (D.grad == 1 && d.geometry != SFX_FLAT) || (D.grad == 2 && d.geometry == SFX_CYLINDER);
  if (D.curved) {
    SFX_REQUIRE(!(D.surf1[0] && d.offset_first_layer <= 0), SFX_E_INVALID,
                "a surface at the lowerbound of a curved lattice needs offset_first_layer > 0");
    curved_tables(L, d.geometry == SFX_SPHERE, D.grad == 2);
    const size_t MX = D.MX;
    SFX_HIP(hipMalloc(&L.dtab, 5 * MX * sizeof(double)));
    const std::vector<double>* tabs[5] = {&L.hlm, &L.hlp, &L.hlmc, &L.hlpc, &L.hL};
    for (int t = 0; t < 5; t++)
      SFX_HIP(hipMemcpy(L.dtab + t * MX, tabs[t]->data(), MX * sizeof(double), hipMemcpyHostToDevice));
    D.lm = L.dtab;
    D.lp = L.dtab + MX;
    D.lmc = L.dtab + 2 * MX;
    D.lpc = L.dtab + 3 * MX;
    D.Lw = L.dtab + 4 * MX;
  }
}

}  // namespace sfx
