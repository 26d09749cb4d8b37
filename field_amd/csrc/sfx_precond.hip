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
// projection I - 11^T/n. The preconditioner applies (sigma I + J0(q))^+
// (sigma = 1/dt of pseudo-transient continuation, 0 for plain Newton; a
// regularised pseudo-inverse per block): pack the residual's interior into
// complex fields, DFT along z, y, x, solve the small block per q, inverse
// DFT, unpack. A power-of-two side is transformed by an FFT in LDS (k_fft);
// any other side by a direct sum over the line with a twiddle table (k_dft,
// one thread per output element, consecutive threads along the innermost
// contiguous axis so every load is coalesced).
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

// The same transform for a power-of-two line length n (2..1024) as an FFT in
// LDS: radix-2 Stockham auto-sort (Govindaraju et al., SC 2008), log2(n)
// stages ping-ponging two LDS buffers, output in natural order. A workgroup
// takes kFftPts / n consecutive lines (line t = (field, o, i)); loads and
// stores run along i (inner > 1) or along the line (inner == 1) so they
// coalesce. Rows are padded to n + 1 entries against LDS bank conflicts.
constexpr int kFftPts = 1024;
__global__ void __launch_bounds__(kPB) k_fft(const cf* __restrict__ in, cf* __restrict__ out, long long outer,
                                             int n, long long inner, const cf* __restrict__ tw, int fields,
                                             long long field_stride) {
  __shared__ cf buf[2][kFftPts + kFftPts / 2];  // LB rows of n + 1 (LB * (n + 1) <= 1.5 kFftPts for n >= 2)
  const int LB = kFftPts / n, P = n + 1, half = n >> 1;
  const long long nlines = outer * inner, total = nlines * fields;
  const long long t0 = (long long)blockIdx.x * LB;
  for (int e = threadIdx.x; e < LB * n; e += kPB) {
    const int l = inner == 1 ? e / n : e % LB;
    const int j = inner == 1 ? e % n : e / LB;
    const long long T = t0 + l;
    if (T >= total) continue;
    const long long f = T / nlines, t = T - f * nlines, o = t / inner, i = t - o * inner;
    buf[0][l * P + j] = in[f * field_stride + (o * n + j) * inner + i];
  }
  int s = 0;
  for (int Ns = 1; Ns < n; Ns <<= 1) {
    __syncthreads();
    const cf* src = buf[s];
    cf* dst = buf[s ^ 1];
    for (int b = threadIdx.x; b < LB * half; b += kPB) {
      const int l = b / half, jb = b - l * half;
      const cf v0 = src[l * P + jb], v1 = src[l * P + jb + half];
      const int k = jb & (Ns - 1);
      const cf w = tw[k * (n / (2 * Ns))];
      const float tx = v1.x * w.x - v1.y * w.y, ty = v1.x * w.y + v1.y * w.x;
      const int d = ((jb - k) << 1) + k;
      dst[l * P + d] = cf{v0.x + tx, v0.y + ty};
      dst[l * P + d + Ns] = cf{v0.x - tx, v0.y - ty};
    }
    s ^= 1;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < LB * n; e += kPB) {
    const int l = inner == 1 ? e / n : e % LB;
    const int j = inner == 1 ? e % n : e / LB;
    const long long T = t0 + l;
    if (T >= total) continue;
    const long long f = T / nlines, t = T - f * nlines, o = t / inner, i = t - o * inner;
    out[f * field_stride + (o * n + j) * inner + i] = buf[s][l * P + j];
  }
}

static bool fft_len(int n) { return n >= 2 && n <= kFftPts && (n & (n - 1)) == 0; }

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

// y(q) = (sigma I + J0(q))^+ r(q), J0 stored [i][j][V]. The block A is
// solved in the Tikhonov form (A^T A + mu^2 I) y = A^T r with mu = 1e-6 max|A|
// in FP64: an exact inverse for a regular block, and a pseudo-inverse for a
// singular one (the projection J0(0); the checkerboard q = (pi,pi,pi) of an
// even lattice, where an even chain's response cancels; sigma I + J0 where
// J0 has the eigenvalue -sigma), which leaves null directions alone instead of
// amplifying rounding noise along them.
// NV (the number of iteration variables) is a template argument so the small
// matrices live in registers.
template <int NV>
__global__ void __launch_bounds__(kPB) k_apply_block(const cf* __restrict__ in, cf* __restrict__ out,
                                                     const float* __restrict__ j0, long long V, float sigma) {
  constexpr int nv = NV;
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= V) return;
  double A[NV][NV], B[NV][NV], re[NV], im[NV];
  double amax = 0;
  for (int i = 0; i < nv; i++)
    for (int j = 0; j < nv; j++) {
      A[i][j] = (double)j0[((long long)i * nv + j) * V + q] + (i == j ? (double)sigma : 0.0);
      amax = fmax(amax, fabs(A[i][j]));
    }
  const double mu2 = (1e-6 * amax) * (1e-6 * amax) + 1e-300;
  for (int i = 0; i < nv; i++) {
    double sr = 0, si = 0;
    for (int k = 0; k < nv; k++) {
      const cf v = in[(long long)k * V + q];
      sr += A[k][i] * v.x;
      si += A[k][i] * v.y;
    }
    re[i] = sr;
    im[i] = si;
    for (int j = 0; j < nv; j++) {
      double s = 0;
      for (int k = 0; k < nv; k++) s += A[k][i] * A[k][j];
      B[i][j] = s + (i == j ? mu2 : 0.0);
    }
  }
  // B is symmetric positive definite: Gauss-Jordan without pivoting
  for (int c = 0; c < nv; c++) {
    const double inv = 1.0 / B[c][c];
    for (int j = 0; j < nv; j++) B[c][j] *= inv;
    re[c] *= inv;
    im[c] *= inv;
    for (int r = 0; r < nv; r++) {
      if (r == c) continue;
      const double f = B[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < nv; j++) B[r][j] -= f * B[c][j];
      re[r] -= f * re[c];
      im[r] -= f * im[c];
    }
  }
  for (int i = 0; i < nv; i++) out[(long long)i * V + q] = cf{(float)re[i], (float)im[i]};
}

// ---------------------------------------------------------------------------
// Fused path for power-of-two sides with a y-z plane of at most kPlane
// points: three passes instead of nine.
//   k_plane_fwd: pack one (field, x) plane of the residual and FFT it along z,
//                then along y, in LDS;
//   k_xline:     per group of (y, z) lines, all fields: FFT along x, the block
//                solve per wave vector, inverse FFT along x;
//   k_plane_inv: inverse FFT of one plane along y, then z, real part / V into
//                the [itvar][M] layout (halos and off-site entries 0).
// The butterflies are those of k_fft (radix-2 Stockham, natural order).
constexpr int kPlane = 4096;
constexpr int kPlaneThreads = 1024;
constexpr int kXTw = 512;   // longest x side of the fused path (twiddles in LDS)
constexpr int kXPts = 2048;  // points (all fields) per k_xline workgroup

// one radix-2 Stockham stage over `lines` lines of n points: element j of line
// l sits at l * ls + j * js in src and dst
__device__ __forceinline__ void stockham_stage(const cf* src, cf* dst, int lines, int n, int Ns, int ls, int js,
                                               const cf* tw, int nthreads) {
  const int half = n >> 1;
  for (int b = threadIdx.x; b < lines * half; b += nthreads) {
    // lines fastest when they are the contiguous direction (js > 1)
    const int l = js > 1 ? b % lines : b / half;
    const int jb = js > 1 ? b / lines : b % half;
    const cf v0 = src[l * ls + jb * js], v1 = src[l * ls + (jb + half) * js];
    const int k = jb & (Ns - 1);
    const cf w = tw[k * (n / (2 * Ns))];
    const float tx = v1.x * w.x - v1.y * w.y, ty = v1.x * w.y + v1.y * w.x;
    const int d = ((jb - k) << 1) + k;
    dst[l * ls + d * js] = cf{v0.x + tx, v0.y + ty};
    dst[l * ls + (d + Ns) * js] = cf{v0.x - tx, v0.y - ty};
  }
}

// FFT of every line of an LDS array in place (result back in buf[0] or
// buf[1]; returns which)
__device__ __forceinline__ int lds_fft(cf* buf0, cf* buf1, int cur, int lines, int n, int ls, int js,
                                       const cf* tw, int nthreads) {
  cf* b[2] = {buf0, buf1};
  for (int Ns = 1; Ns < n; Ns <<= 1) {
    __syncthreads();
    stockham_stage(b[cur], b[cur ^ 1], lines, n, Ns, ls, js, tw, nthreads);
    cur ^= 1;
  }
  __syncthreads();
  return cur;
}

// the twiddles of the y and z transforms of a plane (ny + nz <= kPlane / 2 + 2
// entries) copied into LDS once per workgroup
__device__ __forceinline__ void load_plane_twiddles(cf* twl, const cf* __restrict__ twy, int ny,
                                                    const cf* __restrict__ twz, int nz) {
  for (int e = threadIdx.x; e < ny + nz; e += blockDim.x) twl[e] = e < ny ? twy[e] : twz[e - ny];
}

__global__ void __launch_bounds__(kPlaneThreads) k_plane_fwd(DevLattice L, const double* __restrict__ r, int nv,
                                                             cf* __restrict__ out, const cf* __restrict__ twy,
                                                             const cf* __restrict__ twz) {
  __shared__ cf buf[2][kPlane + kPlane / 2];  // rows of nz + 1
  __shared__ cf twl[kPlane / 2 + 2];
  const int ny = L.ny, nz = L.nz, P = nz + 1;
  const int f = blockIdx.x / L.nx, x = blockIdx.x % L.nx;
  load_plane_twiddles(twl, twy, ny, twz, nz);
  for (int e = threadIdx.x; e < ny * nz; e += kPlaneThreads) {
    const int y = e / nz, z = e % nz;
    buf[0][y * P + z] = cf{(float)r[(long long)f * L.M + sidx(L, x + 1, y + 1, z + 1)], 0.f};
  }
  int cur = lds_fft(buf[0], buf[1], 0, ny, nz, P, 1, twl + ny, kPlaneThreads);  // rows: along z
  cur = lds_fft(buf[0], buf[1], cur, nz, ny, 1, P, twl, kPlaneThreads);         // columns: along y
  cf* o = out + (long long)f * L.V + (long long)x * ny * nz;
  for (int e = threadIdx.x; e < ny * nz; e += kPlaneThreads) o[e] = buf[cur][(e / nz) * P + e % nz];
}

template <int NV>
__global__ void __launch_bounds__(kPB) k_xline(const cf* __restrict__ in, cf* __restrict__ out,
                                               const float* __restrict__ j0, int nx, long long nyz, int LB,
                                               const cf* __restrict__ twf, const cf* __restrict__ twi, float sigma) {
  __shared__ cf buf[2][kXPts + kXPts / 2];  // NV * LB lines of nx + 1
  __shared__ cf twl[2][kXTw];                // forward and inverse twiddles
  const long long V = nx * nyz;
  const int P = nx + 1, lines = NV * LB;
  const long long c0 = (long long)blockIdx.x * LB;  // first (y, z) line
  for (int e = threadIdx.x; e < nx; e += kPB) {
    twl[0][e] = twf[e];
    twl[1][e] = twi[e];
  }
  // line (f, l) at (f * LB + l) * P; loads coalesce along l (consecutive y-z)
  for (int e = threadIdx.x; e < lines * nx; e += kPB) {
    const int l = e % LB, j = (e / LB) % nx, f = e / (LB * nx);
    const long long c = c0 + l;
    buf[0][(f * LB + l) * P + j] = c < nyz ? in[(long long)f * V + j * nyz + c] : cf{0.f, 0.f};
  }
  int cur = lds_fft(buf[0], buf[1], 0, lines, nx, P, 1, twl[0], kPB);
  // (sigma I + J0(q))^+ per wave vector, as k_apply_block
  for (int e = threadIdx.x; e < LB * nx; e += kPB) {
    const int l = e % LB, j = e / LB;
    const long long c = c0 + l;
    if (c >= nyz) continue;
    const long long q = j * nyz + c;
    double A[NV][NV], B[NV][NV], re[NV], im[NV];
    double amax = 0;
    for (int i = 0; i < NV; i++)
      for (int k = 0; k < NV; k++) {
        A[i][k] = (double)j0[((long long)i * NV + k) * V + q] + (i == k ? (double)sigma : 0.0);
        amax = fmax(amax, fabs(A[i][k]));
      }
    const double mu2 = (1e-6 * amax) * (1e-6 * amax) + 1e-300;
    for (int i = 0; i < NV; i++) {
      double sr = 0, si = 0;
      for (int k = 0; k < NV; k++) {
        const cf v = buf[cur][(k * LB + l) * P + j];
        sr += A[k][i] * v.x;
        si += A[k][i] * v.y;
      }
      re[i] = sr;
      im[i] = si;
      for (int m = 0; m < NV; m++) {
        double sacc = 0;
        for (int k = 0; k < NV; k++) sacc += A[k][i] * A[k][m];
        B[i][m] = sacc + (i == m ? mu2 : 0.0);
      }
    }
    for (int cc = 0; cc < NV; cc++) {
      const double inv = 1.0 / B[cc][cc];
      for (int m = 0; m < NV; m++) B[cc][m] *= inv;
      re[cc] *= inv;
      im[cc] *= inv;
      for (int rr = 0; rr < NV; rr++) {
        if (rr == cc) continue;
        const double fct = B[rr][cc];
        if (fct == 0.0) continue;
        for (int m = 0; m < NV; m++) B[rr][m] -= fct * B[cc][m];
        re[rr] -= fct * re[cc];
        im[rr] -= fct * im[cc];
      }
    }
    // each thread reads and writes only its own (l, j) across the fields
    for (int i = 0; i < NV; i++) buf[cur][(i * LB + l) * P + j] = cf{(float)re[i], (float)im[i]};
  }
  cur = lds_fft(buf[0], buf[1], cur, lines, nx, P, 1, twl[1], kPB);
  for (int e = threadIdx.x; e < lines * nx; e += kPB) {
    const int l = e % LB, j = (e / LB) % nx, f = e / (LB * nx);
    const long long c = c0 + l;
    if (c < nyz) out[(long long)f * V + j * nyz + c] = buf[cur][(f * LB + l) * P + j];
  }
}

__global__ void __launch_bounds__(kPlaneThreads) k_plane_inv(DevLattice L, const cf* __restrict__ in, int nv,
                                                             const unsigned char* __restrict__ itmask,
                                                             double* __restrict__ out, const cf* __restrict__ twy,
                                                             const cf* __restrict__ twz) {
  __shared__ cf buf[2][kPlane + kPlane / 2];
  __shared__ cf twl[kPlane / 2 + 2];
  const int ny = L.ny, nz = L.nz, P = nz + 1;
  const int f = blockIdx.x / L.nx, x = blockIdx.x % L.nx;
  load_plane_twiddles(twl, twy, ny, twz, nz);
  const cf* src = in + (long long)f * L.V + (long long)x * ny * nz;
  for (int e = threadIdx.x; e < ny * nz; e += kPlaneThreads) buf[0][(e / nz) * P + e % nz] = src[e];
  int cur = lds_fft(buf[0], buf[1], 0, nz, ny, 1, P, twl, kPlaneThreads);  // columns: along y
  cur = lds_fft(buf[0], buf[1], cur, ny, nz, P, 1, twl + ny, kPlaneThreads);  // rows: along z
  // the whole M-layout plane x + 1 (halo rows and columns 0); the first x
  // also clears the two halo planes
  double* o = out + (long long)f * L.M;
  const long long plane = (long long)L.MY * L.MZ;
  for (long long e = threadIdx.x; e < plane; e += kPlaneThreads) {
    const int y = (int)(e / L.MZ), z = (int)(e % L.MZ);
    const long long s = (long long)(x + 1) * plane + e;
    double v = 0.0;
    if (y >= 1 && y <= ny && z >= 1 && z <= nz && itmask[s])
      v = (double)buf[cur][(y - 1) * P + (z - 1)].x / (double)L.V;
    o[s] = v;
    if (x == 0) {
      o[e] = 0.0;
      o[(long long)(L.nx + 1) * plane + e] = 0.0;
    }
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

// J0(q) of the homogeneous melt, one thread per wave vector q (layout
// [i][j][V]); coef[(t*ns + t2)*maxN + d] holds sum_mol theta/(V N) times the
// number of (t, t2) segment pairs at chain distance d, so S_tt'(q) is a
// polynomial in a(q) evaluated by Horner's rule. chi, the iteration-variable
// map and its representatives come from sp, so a chi change only reruns this.
struct J0Shape {
  int n[3];
  int nv, maxN;
  double phi0[SFX_MAXS];
};
__global__ void __launch_bounds__(kPB) k_build_j0(SysParams sp, J0Shape sh, const double* __restrict__ coef,
                                                  float* __restrict__ j0, long long V) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= V) return;
  const int ns = sp.ns, nv = sh.nv;
  double J[SFX_MAXS][SFX_MAXS];
  if (q == 0) {
    for (int i = 0; i < nv; i++)
      for (int j = 0; j < nv; j++) J[i][j] = (i == j ? 1.0 : 0.0) - 1.0 / nv;
  } else {
    const int qz = (int)(q % sh.n[2]);
    const int qy = (int)((q / sh.n[2]) % sh.n[1]);
    const int qx = (int)(q / ((long long)sh.n[1] * sh.n[2]));
    const double a = (cos(2.0 * M_PI * qx / sh.n[0]) + cos(2.0 * M_PI * qy / sh.n[1]) +
                      cos(2.0 * M_PI * qz / sh.n[2])) / 3.0;
    double Sq[SFX_MAXS][SFX_MAXS];
    for (int t = 0; t < ns; t++)
      for (int t2 = 0; t2 < ns; t2++) {
        const double* c = coef + ((long long)t * ns + t2) * sh.maxN;
        double acc = 0;
        for (int d = sh.maxN - 1; d >= 0; d--) acc = acc * a + c[d];
        Sq[t][t2] = acc;
      }
    for (int col = 0; col < nv; col++) {
      double dphi[SFX_MAXS], dup[SFX_MAXS], tot = 0, mean = 0;
      for (int t = 0; t < ns; t++) {
        double v = 0;
        for (int t2 = 0; t2 < ns; t2++)
          if (sp.st_itvar[t2] == col) v -= Sq[t][t2];
        dphi[t] = v;
        tot += v;
      }
      for (int i = 0; i < nv; i++) {
        const int t = sp.itvar_rep[i];
        double v = i == col ? 1.0 : 0.0;
        for (int j = 0; j < ns; j++) v -= sp.chi[t * SFX_MAXS + j] * (a * dphi[j] - sh.phi0[j] * tot);
        dup[i] = v;
        mean += v / nv;
      }
      for (int i = 0; i < nv; i++) J[i][col] = -tot + dup[i] - mean;
    }
  }
  for (int i = 0; i < nv; i++)
    for (int j = 0; j < nv; j++) j0[((long long)i * nv + j) * V + q] = (float)J[i][j];
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
  DevBuf<float> j0;
  DevBuf<double> coef;        // Debye polynomial coefficients (k_build_j0)
  J0Shape sh{};
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
    sh = J0Shape{};
    sh.n[0] = n[0];
    sh.n[1] = n[1];
    sh.n[2] = n[2];
    sh.nv = nv;
    sh.maxN = 1;
    for (const Molecule& mo : S.mols) sh.maxN = std::max(sh.maxN, mo.N);
    std::vector<double> c((size_t)ns * ns * sh.maxN, 0.0);
    for (const Molecule& mo : S.mols) {
      const double w = mo.theta / ((double)V * mo.N);
      for (int s = 0; s < mo.N; s++) {
        sh.phi0[mo.seq[s]] += w;
        for (int s2 = 0; s2 < mo.N; s2++)
          c[((size_t)mo.seq[s] * ns + mo.seq[s2]) * sh.maxN + std::abs(s - s2)] += w;
      }
    }
    coef.alloc(c.size());
    SFX_HIP(hipMemcpy(coef.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
    j0.alloc((size_t)nv * nv * V);
    build_j0(S);
  }

  // J0 at the system's current chi
  void build_j0(const System& S) {
    k_build_j0<<<grid1(V, kPB), kPB, 0, S.stream>>>(S.sp, sh, coef.p, j0.p, V);
    SFX_HIP(hipGetLastError());
  }

  // one axis: FFT for power-of-two lengths, the direct sum otherwise
  void axis(const System& S, const cf* a, cf* b, long long outer, int len, long long inner, const cf* t) {
    if (fft_len(len)) {
      const long long lines = outer * inner * nv, lb = kFftPts / len;
      k_fft<<<(unsigned)((lines + lb - 1) / lb), kPB, 0, S.stream>>>(a, b, outer, len, inner, t, nv, V);
    } else {
      k_dft<<<grid1(V * nv, kPB), kPB, 0, S.stream>>>(a, b, outer, len, inner, t, nv, V);
    }
  }

  void dft3(const System& S, cf* a, cf* b, int dir, cf*& result) {
    // z (inner 1), y (inner nz), x (inner ny nz); ping-pong a -> b -> a -> b
    const long long nx = n[0], ny = n[1], nz = n[2];
    axis(S, a, b, nx * ny, (int)nz, 1, tw.p + twoff[2][dir]);
    axis(S, b, a, nx, (int)ny, nz, tw.p + twoff[1][dir]);
    axis(S, a, b, 1, (int)nx, ny * nz, tw.p + twoff[0][dir]);
    result = b;
  }

  // the three-pass path: power-of-two sides, a y-z plane within kPlane
  // points and every field of a x-line group within kXPts points
  bool fused() const {
    for (int a = 0; a < 3; a++)
      if (!fft_len(n[a])) return false;
    return (long long)n[1] * n[2] <= kPlane && nv * n[0] <= kXPts && n[0] <= kXTw;
  }

  template <int NV>
  void xline(const System& S, double sigma) {
    const int LB = kXPts / (NV * n[0]);
    const long long nyz = (long long)n[1] * n[2];
    const unsigned g = (unsigned)((nyz + LB - 1) / LB);
    k_xline<NV><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, n[0], nyz, LB, tw.p + twoff[0][0], tw.p + twoff[0][1],
                                         (float)sigma);
  }

  void apply(const System& S, const double* r, double* out, double sigma) {
    const DevLattice& L = S.lat.d;
    if (fused()) {
      const unsigned planes = (unsigned)(n[0] * nv);
      k_plane_fwd<<<planes, kPlaneThreads, 0, S.stream>>>(L, r, nv, b1.p, tw.p + twoff[1][0], tw.p + twoff[2][0]);
      switch (nv) {
        case 1: xline<1>(S, sigma); break;
        case 2: xline<2>(S, sigma); break;
        case 3: xline<3>(S, sigma); break;
        case 4: xline<4>(S, sigma); break;
        case 5: xline<5>(S, sigma); break;
        case 6: xline<6>(S, sigma); break;
        case 7: xline<7>(S, sigma); break;
        default: xline<8>(S, sigma); break;
      }
      k_plane_inv<<<planes, kPlaneThreads, 0, S.stream>>>(L, b0.p, nv, S.d_itmask, out, tw.p + twoff[1][1],
                                                          tw.p + twoff[2][1]);
      SFX_HIP(hipGetLastError());
      return;
    }
    k_pack<<<grid1(V * nv, kPB), kPB, 0, S.stream>>>(L, r, nv, b0.p);
    cf* res = nullptr;
    dft3(S, b0.p, b1.p, 0, res);  // result in b1
    const unsigned g = grid1(V, kPB);
    switch (nv) {
      case 1: k_apply_block<1><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      case 2: k_apply_block<2><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      case 3: k_apply_block<3><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      case 4: k_apply_block<4><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      case 5: k_apply_block<5><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      case 6: k_apply_block<6><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      case 7: k_apply_block<7><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
      default: k_apply_block<8><<<g, kPB, 0, S.stream>>>(b1.p, b0.p, j0.p, V, (float)sigma); break;
    }
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
void precond_apply(Precond* p, const System& S, const double* r, double* out, double sigma) {
  p->apply(S, r, out, sigma);
}
void precond_update(Precond* p, const System& S) {
  if (p) p->build_j0(S);
}
void precond_free(Precond* p) { delete p; }

}  // namespace sfx
