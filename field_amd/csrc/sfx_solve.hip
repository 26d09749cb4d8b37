// sfx_solve.hip -- device-resident SCF solvers.
//
// DIIS follows MPS.cpp:228-284 (memory m, step eta, sliding window, the
// coefficients solve A c = 1 then normalise) with every vector on the device:
// per iteration one residual evaluation, one multi-dot kernel, one host solve
// of an m x m system, one combination kernel. Vectors use the internal
// [itvar][M] layout with zeros off the iteration sites, so their dot products
// equal those of the reference's iteration-variable vectors.
#include <chrono>
#include <cmath>
#include <cstring>

#include "sfx_solve.h"

namespace sfx {

constexpr int kDotBlock = 256;
constexpr int kMaxDots = 16;
constexpr int kMaxVecs = 32;  // VecList capacity
constexpr int kGsWords = 2 * kMaxVecs + 1;  // two Gram-Schmidt passes and a norm

// partial[d][block] = sum_i a_d[i] * b_d[i]
static __global__ void __launch_bounds__(kDotBlock) k_multidot(DotList dl, long long n, double* partial) {
  __shared__ double sm[kMaxDots][kDotBlock / 64];
  double acc[kMaxDots];
  for (int d = 0; d < dl.count; d++) acc[d] = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    for (int d = 0; d < dl.count; d++) acc[d] += dl.a[d][i] * dl.b[d][i];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int d = 0; d < dl.count; d++) {
    double v = acc[d];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sm[d][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < dl.count) {
    double s = 0;
    for (int k = 0; k < kDotBlock / 64; k++) s += sm[threadIdx.x][k];
    partial[(long long)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// out[r] = sum_c partial[r][c]: one workgroup per row, coalesced strided
// loads, then a fixed-order wave and workgroup reduction; with a mirror
// (coherent pinned host memory) the sum is stored there too, so the host
// reads it after a stream synchronisation without a copy
static __global__ void __launch_bounds__(kDotBlock) k_reduce_rows(const double* partial, int ncols, double* out,
                                                                  double* mirror) {
  __shared__ double sm[kDotBlock / 64];
  const int r = blockIdx.x;
  double s = 0;
  for (int c = threadIdx.x; c < ncols; c += kDotBlock) s += partial[(long long)r * ncols + c];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int k = 0; k < kDotBlock / 64; k++) t += sm[k];
    out[r] = t;
    if (mirror) mirror[r] = t;
  }
}

// y = sum_k c_k v_k  (c passed by value)
static __global__ void k_combine(VecList vl, long long n, double* y) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int k = 0; k < vl.count; k++) s += vl.c[k] * vl.v[k][i];
  y[i] = s;
}

// y = v_last - sum_{k < count-1} dc[k] v_k  (dc in device memory): one
// Gram-Schmidt pass with the projections still on the device
static __global__ void k_subtract_projections(VecList vl, const double* __restrict__ dc, long long n, double* y) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  const int last = vl.count - 1;
  for (int k = 0; k < last; k++) s += (-dc[k]) * vl.v[k][i];
  s += 1.0 * vl.v[last][i];
  y[i] = s;
}

// y = a*x + b*y
static __global__ void k_axpby(double a, const double* x, double b, double* y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}
// y = a*x (the old y is not read, so a NaN left in it does not survive)
static __global__ void k_scale(double a, const double* x, double* y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}
static __global__ void k_zero_vec(double* y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = 0.0;
}
// z = x - y
static __global__ void k_sub(const double* x, const double* y, double* z, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] - y[i];
}
static __global__ void k_copy(const double* x, double* y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

void VecOps::init(long long n_, hipStream_t st) {
  stream = st;
  if (partial && n == n_) return;
  if (partial) (void)hipFree(partial);
  if (dres) (void)hipFree(dres);
  partial = dres = nullptr;
  n = n_;
  nblk = (int)std::min<long long>(grid1(n, kDotBlock), 1024);
  SFX_HIP(hipMalloc(&partial, sizeof(double) * kMaxDots * nblk));
  SFX_HIP(hipMalloc(&dres, sizeof(double) * kMaxDots));
  if (!hres) SFX_HIP(hipHostMalloc(&hres, sizeof(double) * kHostWords, hipHostMallocCoherent));
}
VecOps::~VecOps() {
  if (partial) (void)hipFree(partial);
  if (dres) (void)hipFree(dres);
  if (hres) (void)hipHostFree(hres);
}
void VecOps::dots_dev(const DotList& dl, double* dout, int mirror_at) {
  SFX_REQUIRE(dl.count <= kMaxDots, SFX_E_INVALID, "too many dot products");
  SFX_REQUIRE(mirror_at < 0 || mirror_at + dl.count <= kHostWords, SFX_E_INVALID, "host mirror out of range");
  if (dl.count <= 0) return;
  k_multidot<<<nblk, kDotBlock, 0, stream>>>(dl, n, partial);
  k_reduce_rows<<<dl.count, kDotBlock, 0, stream>>>(partial, nblk, dout, mirror_at < 0 ? nullptr : hres + mirror_at);
}
void VecOps::wait_host(int count, double* host_out) {
  SFX_REQUIRE(count >= 0 && count <= kHostWords, SFX_E_INVALID, "too many words to read back");
  SFX_HIP(hipStreamSynchronize(stream));
  std::memcpy(host_out, hres, sizeof(double) * count);
}
void VecOps::fetch(const double* dsrc, int count, double* host_out) {
  SFX_REQUIRE(count >= 0 && count <= kHostWords, SFX_E_INVALID, "too many words to read back");
  if (count == 0) return;
  SFX_HIP(hipMemcpyAsync(hres, dsrc, sizeof(double) * count, hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipStreamSynchronize(stream));
  std::memcpy(host_out, hres, sizeof(double) * count);
}
void VecOps::dots(const DotList& dl, double* host_out) {
  if (dl.count <= 0) return;
  dots_dev(dl, dres, 0);
  wait_host(dl.count, host_out);
}
double VecOps::dot(const double* a, const double* b) {
  DotList dl{};
  dl.a[0] = a;
  dl.b[0] = b;
  dl.count = 1;
  double r;
  dots(dl, &r);
  return r;
}
void VecOps::combine(const VecList& vl, double* y) {
  SFX_REQUIRE(vl.count >= 1 && vl.count <= kMaxVecs, SFX_E_INVALID, "combination of 1..32 vectors");
  k_combine<<<grid1(n, 256), 256, 0, stream>>>(vl, n, y);
}
void VecOps::subtract_projections(const VecList& vl, const double* dc, double* y) {
  SFX_REQUIRE(vl.count >= 1 && vl.count <= kMaxVecs, SFX_E_INVALID, "combination of 1..32 vectors");
  k_subtract_projections<<<grid1(n, 256), 256, 0, stream>>>(vl, dc, n, y);
}
SolveWork& solve_work(System& S, long long n, size_t nvec) {
  if (!S.work) S.work = new SolveWork();
  SolveWork& w = *S.work;
  if (w.n != n || w.nvec < nvec) {  // grows only: a hipFree waits for the device
    w.vecs.reset((size_t)n * nvec);
    w.n = n;
    w.nvec = nvec;
  }
  if (!w.gs.p) w.gs.alloc(kGsWords);
  w.V.init(n, S.stream);
  return w;
}
void solve_work_free(SolveWork* w) { delete w; }

void VecOps::axpby(double a, const double* x, double b, double* y) {
  k_axpby<<<grid1(n, 256), 256, 0, stream>>>(a, x, b, y, n);
}
void VecOps::scale(double a, const double* x, double* y) {
  k_scale<<<grid1(n, 256), 256, 0, stream>>>(a, x, y, n);
}
void VecOps::zero(double* y) { k_zero_vec<<<grid1(n, 256), 256, 0, stream>>>(y, n); }
void VecOps::sub(const double* x, const double* y, double* z) { k_sub<<<grid1(n, 256), 256, 0, stream>>>(x, y, z, n); }
void VecOps::copy(const double* x, double* y) { k_copy<<<grid1(n, 256), 256, 0, stream>>>(x, y, n); }

// X = A^{-1} 1 for the symmetric DIIS matrix: Jacobi eigen-decomposition,
// X = Q diag(1/w) Q^T 1 (the pseudo-inverse of MPS.cpp:200-226 for a
// non-singular A).
void diis_coeffs(const std::vector<double>& Ain, int n, std::vector<double>& X) {
  std::vector<double> A(Ain), Q(n * n, 0.0), w(n);
  for (int i = 0; i < n; i++) Q[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p * n + q];
        if (apq == 0) continue;
        double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        double t = (th >= 0 ? 1 : -1) / (std::fabs(th) + std::sqrt(th * th + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double qkp = Q[k * n + p], qkq = Q[k * n + q];
          Q[k * n + p] = c * qkp - s * qkq;
          Q[k * n + q] = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = A[i * n + i];
  X.assign(n, 0.0);
  for (int k = 0; k < n; k++) {
    double qt1 = 0;
    for (int i = 0; i < n; i++) qt1 += Q[i * n + k];
    double c = qt1 / w[k];
    for (int i = 0; i < n; i++) X[i] += Q[i * n + k] * c;
  }
}

void solve_diis(System& S, const sfx_solver_opts& o, sfx_solve_stats& st) {
  const long long n = (long long)S.n_itvar * S.lat.d.M;
  const int m = o.diis_m;
  const double eta = o.diis_eta;
  SFX_REQUIRE(m >= 1 && m <= 15, SFX_E_INVALID, "diis_m must be in 1..15");
  SolveWork& W = solve_work(S, n, 2 * (size_t)m + 1);
  VecOps& V = W.V;
  double* x0 = W.vec(0);
  auto xR = [&](int k) { return W.vec(1 + (size_t)k); };
  auto x_x0 = [&](int k) { return W.vec(1 + (size_t)m + k); };
  double* x = S.d_u;
  double* g = S.d_f;
  hipEvent_t e0, e1;
  SFX_HIP(hipEventCreate(&e0));
  SFX_HIP(hipEventCreate(&e1));
  SFX_HIP(hipEventRecord(e0, S.stream));
  std::vector<double> Aij(m * m, 0.0), Apij, Ci;
  int it = 0, k_diis = 1, k = 0;
  V.copy(x, x0);  // remember the starting point for x_x0 = x - x0 (x0 = 0 in MPS.cpp:264)
  S.residual_replay(SFX_FP64);
  st.residual_evals++;
  V.axpby(-eta, g, 1.0, x);
  V.sub(x, x0, x_x0(0));
  V.copy(x, xR(0));
  double err = std::sqrt(V.dot(g, g));
  if (err <= o.tolerance) V.copy(x0, x);  // the start already meets the tolerance: return it, unmoved
  while (err > o.tolerance && it < o.max_iter) {
    it++;
    V.copy(x, x0);
    S.residual_replay(SFX_FP64);
    st.residual_evals++;
    k = it % m;
    k_diis++;
    V.axpby(-eta, g, 1.0, x);
    V.copy(x, xR(k));
    V.sub(x, x0, x_x0(k));
    if (k_diis > m) {
      k_diis = m;
      for (int i = 1; i < m; i++)
        for (int j = 1; j < m; j++) Aij[m * (i - 1) + j - 1] = Aij[m * i + j];
    }
    DotList dl{};
    dl.count = k_diis + 1;
    for (int i = 0; i < k_diis; i++) {
      int posi = k - k_diis + 1 + i;
      if (posi < 0) posi += m;
      dl.a[i] = x_x0(posi);
      dl.b[i] = x_x0(k);
    }
    dl.a[k_diis] = g;
    dl.b[k_diis] = g;
    double d[kMaxDots];
    V.dots(dl, d);
    for (int i = 0; i < k_diis; i++) Aij[i + m * (k_diis - 1)] = Aij[k_diis - 1 + m * i] = d[i];
    err = std::sqrt(d[k_diis]);
    if (err <= o.tolerance) {
      V.copy(x0, x);  // return the point whose residual was measured
      break;
    }
    Apij.assign(k_diis * k_diis, 0.0);
    for (int i = 0; i < k_diis; i++)
      for (int j = 0; j < k_diis; j++) Apij[j + k_diis * i] = Aij[j + m * i];
    diis_coeffs(Apij, k_diis, Ci);
    double normC = 0;
    for (int i = 0; i < k_diis; i++) normC += Ci[i];
    for (int i = 0; i < k_diis; i++) Ci[i] /= normC;
    VecList vl{};
    vl.count = k_diis;
    for (int i = 0; i < k_diis; i++) {
      int posi = k - k_diis + 1 + i;
      if (posi < 0) posi += m;
      vl.v[i] = xR(posi);
      vl.c[i] = Ci[i];
    }
    V.combine(vl, x);
    if (!std::isfinite(err)) throw Error{SFX_E_NONFINITE, "non-finite residual in DIIS"};
    if (o.verbose) std::fprintf(stderr, "diis it=%d |f|=%.6e\n", it, err);
  }
  SFX_HIP(hipEventRecord(e1, S.stream));
  SFX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SFX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the residual of the returned x (leaving phi / free energy consistent
  // with it): at the iteration limit x is the last, untested DIIS
  // combination, so the norm and the convergence flag come from this evaluation
  S.residual_replay(SFX_FP64);
  const double fin = std::sqrt(V.dot(g, g));
  st.iterations = it;
  st.residual_norm = fin;
  st.seconds = ms * 1e-3;
  st.converged = fin <= o.tolerance;
}

}  // namespace sfx
