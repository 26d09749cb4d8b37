// sfx_precond.hip -- Fourier-space preconditioner for Newton-Krylov on
// periodic simple-cubic melts.
//
// At the homogeneous melt (u = 0, every SWF 1, phi_t = phi0_t) the residual's
// Jacobian commutes with lattice translations, so in wave-vector space it is
// one n_itvar x n_itvar matrix J0(q) per q (the lattice random-phase
// approximation). With a(q) = (cos qx + cos qy + cos qz)/3 the eigenvalue of
// the six-neighbour average (Lat3D1stO.cpp:53-260) and of the contact average
// (Lat3D.cpp:1729-1784):
//   dphi_t(q)  = -sum_t' S_tt'(q) du_t'(q),  q != 0
//   S_tt'(q)   = sum_mol theta/(V N) sum_{s in t, s' in t'} a(q)^|s-s'|
//   du'_i      = du_i - sum_j chi_ij (a dphi_j - phi0_j dphi_tot)
//   df_i       = -dphi_tot + du'_i - mean_k du'_k          (SF_Solve.cpp:554-590)
// and at q = 0 the restricted normalisation keeps dphi = 0, so J0(0) is the
// projection I - 11^T/n. The preconditioner applies J0(q)^-1 (the
// pseudo-inverse at q = 0): pack the residual's interior into complex
// fields, DFT along z, y, x, multiply, inverse DFT, unpack. The DFTs are
// direct sums over a line with a twiddle table (any lattice size); one
// thread per output element, consecutive threads along the innermost
// contiguous axis so every load is coalesced.
#include <cmath>
#include <complex>
#include <vector>

#include "sfx_device.h"
#include "sfx_solve.h"

namespace sfx {

constexpr int kPB = 256;

struct cf {
  float x, y;
};

// out[o, k, i] = sum_j in[o, j, i] * tw[(j k) mod n]   (tw = exp(-+2 pi i m / n))
__global__ void __launch_bounds__(kPB) k_dft(const cf* __restrict__ in, cf* __restrict__ out, long long outer,
                                             int n, long long inner, const cf* __restrict__ tw, int fields,
                                             long long field_stride) {
  const long long per = outer * n * inner;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * fields) return;
  const int f = (int)(t / per);
  t -= (long long)f * per;
  const long long i = t % inner;
  const long long r = t / inner;
  const int k = (int)(r % n);
  const long long o = r / n;
  const cf* src = in + f * field_stride + o * n * inner + i;
  float re = 0.f, im = 0.f;
  int m = 0;  // (j k) mod n, advanced incrementally
  for (int j = 0; j < n; j++) {
    const cf v = src[(long long)j * inner];
    const cf w = tw[m];
    re += v.x * w.x - v.y * w.y;
    im += v.x * w.y + v.y * w.x;
    m += k;
    if (m >= n) m -= n;
  }
  out[f * field_stride + o * n * inner + (long long)k * inner + i] = cf{re, im};
}

// interior of the [itvar][M] residual -> complex compact fields [itvar][V]
__global__ void k_pack(DevLattice L, const double* __restrict__ r, int nv, cf* __restrict__ out) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.V * nv) return;
  const int f = (int)(q / L.V);
  const long long c = q - (long long)f * L.V;
  int x, y, z;
  interior_coords(L, c, x, y, z);
  out[q] = cf{(float)r[(long long)f * L.M + sidx(L, x, y, z)], 0.f};
}

// y(q) = Jinv(q) r(q), Jinv stored [i][j][V]
__global__ void k_apply_block(const cf* __restrict__ in, cf* __restrict__ out, const float* __restrict__ jinv,
                              int nv, long long V) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= V) return;
  for (int i = 0; i < nv; i++) {
    float re = 0.f, im = 0.f;
    for (int j = 0; j < nv; j++) {
      const float a = jinv[((long long)i * nv + j) * V + q];
      const cf v = in[(long long)j * V + q];
      re += a * v.x;
      im += a * v.y;
    }
    out[(long long)i * V + q] = cf{re, im};
  }
}

// real part / V -> [itvar][M] interior (halos and off-site entries 0)
__global__ void k_unpack(DevLattice L, const cf* __restrict__ in, int nv, const unsigned char* __restrict__ itmask,
                         double* __restrict__ out) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.M * nv) return;
  const int f = (int)(q / L.M);
  const long long s = q - (long long)f * L.M;
  int x, y, z;
  site_coords(L, s, x, y, z);
  double v = 0.0;
  if (itmask[s]) {
    const long long c = ((long long)(x - 1) * L.ny + (y - 1)) * L.nz + (z - 1);
    v = (double)in[(long long)f * L.V + c].x / (double)L.V;
  }
  out[q] = v;
}

bool precond_supported(const System& S) {
  const DevLattice& L = S.lat.d;
  if (L.grad != 3 || L.kind != ST_3D_SC || L.order != 1 || S.form != SFX_FORM_MELT) return false;
  for (int a = 0; a < 3; a++)
    if (!L.periodic[a]) return false;
  for (int t = 0; t < S.ns; t++)
    if (S.frozen[t]) return false;
  for (const Molecule& mo : S.mols)
    if (mo.freedom != SFX_RESTRICTED) return false;
  return true;
}

struct Precond {
  int nv = 0;
  long long V = 0;
  int n[3] = {0, 0, 0};
  DevBuf<cf> b0, b1, tw;      // tw: forward then inverse table per axis
  DevBuf<float> jinv;
  long long twoff[3][2] = {};

  void init(const System& S) {
    const DevLattice& L = S.lat.d;
    nv = S.n_itvar;
    V = L.V;
    n[0] = L.nx;
    n[1] = L.ny;
    n[2] = L.nz;
    // twiddle tables
    std::vector<cf> h;
    for (int a = 0; a < 3; a++)
      for (int dir = 0; dir < 2; dir++) {
        twoff[a][dir] = (long long)h.size();
        for (int m = 0; m < n[a]; m++) {
          const double ang = (dir == 0 ? -2.0 : 2.0) * M_PI * m / n[a];
          h.push_back(cf{(float)std::cos(ang), (float)std::sin(ang)});
        }
      }
    tw.alloc(h.size());
    SFX_HIP(hipMemcpy(tw.p, h.data(), h.size() * sizeof(cf), hipMemcpyHostToDevice));
    b0.alloc((size_t)nv * V);
    b1.alloc((size_t)nv * V);
    // homogeneous composition and Debye coefficients per molecule and state pair
    const int ns = S.ns;
    std::vector<double> phi0(ns, 0.0);
    int maxN = 1;
    for (const Molecule& mo : S.mols) maxN = std::max(maxN, mo.N);
    std::vector<double> c((size_t)ns * ns * maxN, 0.0);  // sum over molecules of theta/(V N) * pair counts at distance d
    for (const Molecule& mo : S.mols) {
      const double w = mo.theta / ((double)V * mo.N);
      for (int s = 0; s < mo.N; s++) {
        phi0[mo.seq[s]] += w;
        for (int s2 = 0; s2 < mo.N; s2++)
          c[((size_t)mo.seq[s] * ns + mo.seq[s2]) * maxN + std::abs(s - s2)] += w;
      }
    }
    std::vector<std::vector<double>> chi(ns, std::vector<double>(ns));
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) chi[i][j] = S.chi[(size_t)i * ns + j];
    std::vector<float> hj((size_t)nv * nv * V);
    std::vector<double> Sq((size_t)ns * ns), J((size_t)nv * nv), inv((size_t)nv * nv), dphi(ns), dup(nv);
    for (int qx = 0; qx < n[0]; qx++)
      for (int qy = 0; qy < n[1]; qy++)
        for (int qz = 0; qz < n[2]; qz++) {
          const long long q = ((long long)qx * n[1] + qy) * n[2] + qz;
          if (q == 0) {
            for (int i = 0; i < nv; i++)
              for (int j = 0; j < nv; j++) inv[(size_t)i * nv + j] = (i == j ? 1.0 : 0.0) - 1.0 / nv;
          } else {
            const double a = (std::cos(2 * M_PI * qx / n[0]) + std::cos(2 * M_PI * qy / n[1]) +
                              std::cos(2 * M_PI * qz / n[2])) / 3.0;
            for (int t = 0; t < ns; t++)
              for (int t2 = 0; t2 < ns; t2++) {
                double acc = 0;  // Horner in a
                for (int d = maxN - 1; d >= 0; d--) acc = acc * a + c[((size_t)t * ns + t2) * maxN + d];
                Sq[(size_t)t * ns + t2] = acc;
              }
            for (int col = 0; col < nv; col++) {
              double tot = 0;
              for (int t = 0; t < ns; t++) {
                double v = 0;
                for (int t2 = 0; t2 < ns; t2++)
                  if (S.st_itvar[t2] == col) v -= Sq[(size_t)t * ns + t2];
                dphi[t] = v;
                tot += v;
              }
              double mean = 0;
              for (int i = 0; i < nv; i++) {
                const int t = S.itvar_rep[i];
                double v = i == col ? 1.0 : 0.0;
                for (int j = 0; j < ns; j++) v -= chi[t][j] * (a * dphi[j] - phi0[j] * tot);
                dup[i] = v;
                mean += v / nv;
              }
              for (int i = 0; i < nv; i++) J[(size_t)i * nv + col] = -tot + dup[i] - mean;
            }
            invert(J, inv, nv);
          }
          for (int i = 0; i < nv; i++)
            for (int j = 0; j < nv; j++) hj[((size_t)i * nv + j) * V + q] = (float)inv[(size_t)i * nv + j];
        }
    jinv.alloc(hj.size());
    SFX_HIP(hipMemcpy(jinv.p, hj.data(), hj.size() * sizeof(float), hipMemcpyHostToDevice));
  }

  // Gauss-Jordan with partial pivoting; a singular block gets its singular
  // directions left alone (zero rows of the pivot stay zero)
  static void invert(const std::vector<double>& Ain, std::vector<double>& X, int k) {
    std::vector<double> A(Ain);
    X.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++) X[(size_t)i * k + i] = 1.0;
    for (int c = 0; c < k; c++) {
      int piv = c;
      for (int r = c + 1; r < k; r++)
        if (std::fabs(A[(size_t)r * k + c]) > std::fabs(A[(size_t)piv * k + c])) piv = r;
      for (int j = 0; j < k; j++) {
        std::swap(A[(size_t)c * k + j], A[(size_t)piv * k + j]);
        std::swap(X[(size_t)c * k + j], X[(size_t)piv * k + j]);
      }
      double d = A[(size_t)c * k + c];
      if (std::fabs(d) < 1e-12) d = d < 0 ? -1e-12 : 1e-12;
      for (int j = 0; j < k; j++) {
        A[(size_t)c * k + j] /= d;
        X[(size_t)c * k + j] /= d;
      }
      for (int r = 0; r < k; r++) {
        if (r == c) continue;
        const double f = A[(size_t)r * k + c];
        if (f == 0) continue;
        for (int j = 0; j < k; j++) {
          A[(size_t)r * k + j] -= f * A[(size_t)c * k + j];
          X[(size_t)r * k + j] -= f * X[(size_t)c * k + j];
        }
      }
    }
  }

  void dft3(const System& S, cf* a, cf* b, int dir, cf*& result) {
    // z (inner 1), y (inner nz), x (inner ny nz); ping-pong a -> b -> a -> b
    const long long nx = n[0], ny = n[1], nz = n[2];
    const unsigned g = grid1(V * nv, kPB);
    k_dft<<<g, kPB, 0, S.stream>>>(a, b, nx * ny, (int)nz, 1, tw.p + twoff[2][dir], nv, V);
    k_dft<<<g, kPB, 0, S.stream>>>(b, a, nx, (int)ny, nz, tw.p + twoff[1][dir], nv, V);
    k_dft<<<g, kPB, 0, S.stream>>>(a, b, 1, (int)nx, ny * nz, tw.p + twoff[0][dir], nv, V);
    result = b;
  }

  void apply(const System& S, const double* r, double* out) {
    const DevLattice& L = S.lat.d;
    k_pack<<<grid1(V * nv, kPB), kPB, 0, S.stream>>>(L, r, nv, b0.p);
    cf* res = nullptr;
    dft3(S, b0.p, b1.p, 0, res);  // result in b1
    k_apply_block<<<grid1(V, kPB), kPB, 0, S.stream>>>(b1.p, b0.p, jinv.p, nv, V);
    dft3(S, b0.p, b1.p, 1, res);
    k_unpack<<<grid1(L.M * nv, kPB), kPB, 0, S.stream>>>(L, b1.p, nv, S.d_itmask, out);
    SFX_HIP(hipGetLastError());
  }
};

Precond* precond_create(const System& S) {
  auto* p = new Precond();
  try {
    p->init(S);
  } catch (...) {
    delete p;
    throw;
  }
  return p;
}
void precond_apply(Precond* p, const System& S, const double* r, double* out) { p->apply(S, r, out); }
void precond_free(Precond* p) { delete p; }

}  // namespace sfx
