// sfx_batch.hip -- many independent 1-D adsorption SCF problems at once
// (the parameter sweeps of Vary.cpp:3-80 / SF_Box.cpp:145-176, solved
// concurrently instead of one calculation after the other).
//
// Each problem: Lat1DFlat1stO with n layers, lowerbound surface holding a
// frozen wall W in layer 0, upperbound mirror1, step weight lambda; a free
// homopolymer A_N (phi_b) and a monomeric solvent S; chi_AS, chi_AW = chi_s,
// chi_SW = 0. One workgroup owns one problem. Inside it run the half-stored
// homopolymer sweep (SF_Homopol1stO.cpp:552-622) with the running rank in LDS
// and the stored ranks in HBM, the solvent closure (SF_MolList1stO.cpp:280-292),
// the contact averages (Lat1DFlat.cpp:510-524) and the residual
// (SF_Solve.cpp:554-590), then one DIIS step (MPS.cpp:228-284); a solve is
// one launch in which every workgroup loops over its own problem's iterations
// until that problem is done.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sfx_batch.h"

namespace sfx {

constexpr int kMaxLayers = 1024;
constexpr int kMaxDiis = 15;

template <int BT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0;
#pragma unroll
  for (int i = 0; i < BT / 64; i++) r += red[i];
  return r;
}

// The residual of problem b at x, by the workgroup that owns it; every
// thread of the workgroup calls it. LDS holds two ping-pong rank buffers of
// ldsw doubles; index 0 of both stays 0 (the frozen surface layer), so the
// lower neighbour needs no branch; the mirror1 neighbour of layer n is n.
template <int BT>
__device__ void batch_residual(const BatchProb& p, int b, const double* __restrict__ x, double* __restrict__ f,
                               double* __restrict__ store, double* __restrict__ phi_out,
                               double* __restrict__ lngn, double* lds, int ldsw) {
  constexpr int K = kMaxLayers / BT;  // layers per thread
  __shared__ double red[BT / 64];
  const int n = p.n, N = p.N;
  const long long M = n + 2;
  const double l = p.lambda;
  const double* xb = x + p.xoff;
  double* st = store + p.soff;
  double* cur = lds;
  double* nxt = lds + ldsw;
  const int t = threadIdx.x;

  double gA[K], gS[K], acc[K], pr[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int z = 1 + t + k * BT;
    gA[k] = gS[k] = acc[k] = pr[k] = 0.0;
    if (z <= n) {
      gA[k] = exp(-xb[2 * (z - 1)]);      // SWF of A (SF_Solve.cpp:794-852)
      gS[k] = exp(-xb[2 * (z - 1) + 1]);  // SWF of S
      cur[z] = gA[k];                     // rank 1
      st[z] = gA[k];
    }
  }
  if (t == 0) {
    cur[0] = 0.0;
    nxt[0] = 0.0;
  }
  __syncthreads();

  // G_s(z) = g(z) [G(z) + l (G(z-1) - 2 G(z) + G(z+1))] on the previous rank
#define SFX_STEP(k, z) (gA[k] * (cur[z] + l * (cur[(z) - 1] - 2.0 * cur[z] + ((z) == n ? cur[z] : cur[(z) + 1]))))

  const int h = N / 2;
  for (int s = 2; s <= h; s++) {  // stored ranks 2..N/2
    double* row = st + (long long)(s - 1) * M;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int z = 1 + t + k * BT;
      if (z <= n) {
        const double v = SFX_STEP(k, z);
        nxt[z] = v;
        row[z] = v;
      }
    }
    __syncthreads();
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (N & 1) {  // middle rank of an odd chain meets itself
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int z = 1 + t + k * BT;
      if (z <= n) {
        const double v = SFX_STEP(k, z);
        nxt[z] = v;
        acc[k] += 0.5 * v * v;
      }
    }
    __syncthreads();
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  // ranks N/2+1..N pair with the stored rank N-s+1; each thread reads back
  // only the layers it wrote itself, one rank ahead
  const int s0 = (N + 3) / 2;
  {
    const double* row = st + (long long)(N - s0) * M;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int z = 1 + t + k * BT;
      if (z <= n) pr[k] = row[z];
    }
  }
  for (int s = s0; s <= N; s++) {
    const double* next_row = st + (long long)(N - s - 1) * M;  // read only while s < N
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int z = 1 + t + k * BT;
      if (z <= n) {
        const double v = SFX_STEP(k, z);
        nxt[z] = v;
        acc[k] += v * pr[k];
        if (s < N) pr[k] = next_row[z];
      }
    }
    __syncthreads();
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
#undef SFX_STEP

  // ln G_N over the interior (Lat1DFlat1stO.cpp:301-316, layer weight 1)
  double w = 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int z = 1 + t + k * BT;
    if (z <= n) w += cur[z];
  }
  const double GN = block_sum<BT>(w, red);
  if (t == 0 && lngn) lngn[b] = log(GN);

  // phi_A = 2 (phi_b / N) acc / g (free homopolymer, half store), phi_S = (1 - phi_b) g_S
  const double CA = 2.0 * p.phib / N, phibS = 1.0 - p.phib;
  double* phA = cur;  // block_sum synchronised: every thread is done with the rank buffers
  double* phS = nxt;
  double vA[K], vS[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int z = 1 + t + k * BT;
    vA[k] = vS[k] = 0.0;
    if (z <= n) {
      vA[k] = gA[k] > 0 ? CA * acc[k] / gA[k] : 0.0;
      vS[k] = phibS * gS[k];
      phA[z] = vA[k];
      phS[z] = vS[k];
    }
  }
  __syncthreads();
  if (phi_out) {
    double* po = phi_out + p.poff;
    for (int z = t; z < M; z += BT) {
      const bool in = z >= 1 && z <= n;
      po[z] = in ? phA[z] : 0.0;
      po[M + z] = in ? phS[z] : 0.0;
    }
  }
  // contact averages and the residual (two iteration variables, A and S)
  const double cbA = p.chi_as * phibS, cbS = p.chi_as * p.phib;
  double* fb = f + p.xoff;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int z = 1 + t + k * BT;
    if (z <= n) {
      const double upA = z == n ? phA[n] : phA[z + 1];
      const double upS = z == n ? phS[n] : phS[z + 1];
      const double sA = vA[k] + l * (phA[z - 1] - 2.0 * vA[k] + upA);
      const double sS = vS[k] + l * (phS[z - 1] - 2.0 * vS[k] + upS);
      const double sW = z == 1 ? l : 0.0;  // the wall (phi 1 in layer 0) seen from layer 1
      const double pt = vA[k] + vS[k];
      const double uA = -log(gA[k]) - (p.chi_as * sS + p.chi_s * sW) / pt + cbA;
      const double uS = -log(gS[k]) - (p.chi_as * sA) / pt + cbS;
      const double av = 0.5 * (uA + uS);
      fb[2 * (z - 1)] = -1.0 + 1.0 / pt + uA - av;
      fb[2 * (z - 1) + 1] = -1.0 + 1.0 / pt + uS - av;
    }
  }
}

template <int BT>
__global__ void __launch_bounds__(BT) k_batch_residual(const BatchProb* __restrict__ P, const double* __restrict__ x,
                                                       double* __restrict__ f, double* __restrict__ store,
                                                       double* __restrict__ phi_out, double* __restrict__ lngn,
                                                       const int* __restrict__ active, int ldsw) {
  extern __shared__ double lds[];
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  const BatchProb p = P[b];
  batch_residual<BT>(p, b, x, f, store, phi_out, lngn, lds, ldsw);
}

// Sums four values over the workgroup in a fixed order; every thread gets
// the totals. red holds 4 * (BT / 64) doubles.
template <int BT>
__device__ __forceinline__ void block_reduce4(double (&v)[4], double* red) {
#pragma unroll
  for (int j = 0; j < 4; j++)
    for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_xor(v[j], o, 64);
  __syncthreads();  // red may still be read from the previous call
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int j = 0; j < 4; j++) red[j * (BT / 64) + (threadIdx.x >> 6)] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double r = 0;
    for (int w = 0; w < BT / 64; w++) r += red[j * (BT / 64) + w];
    v[j] = r;
  }
}

// X = A^{-1} 1 by Gaussian elimination with partial pivoting; A (k x k, row
// major), rhs and X in LDS, one thread.
__device__ void diis_solve(double* A, double* rhs, int k, double* X) {
  for (int i = 0; i < k; i++) rhs[i] = 1.0;
  for (int c = 0; c < k; c++) {
    int piv = c;
    for (int r = c + 1; r < k; r++)
      if (fabs(A[r * k + c]) > fabs(A[piv * k + c])) piv = r;
    if (piv != c) {
      for (int j = 0; j < k; j++) {
        const double tmp = A[c * k + j];
        A[c * k + j] = A[piv * k + j];
        A[piv * k + j] = tmp;
      }
      const double tmp = rhs[c];
      rhs[c] = rhs[piv];
      rhs[piv] = tmp;
    }
    const double d = A[c * k + c];
    for (int r = c + 1; r < k; r++) {
      const double fct = A[r * k + c] / d;
      for (int j = c; j < k; j++) A[r * k + j] -= fct * A[c * k + j];
      rhs[r] -= fct * rhs[c];
    }
  }
  for (int i = k - 1; i >= 0; i--) {
    double s = rhs[i];
    for (int j = i + 1; j < k; j++) s -= A[i * k + j] * X[j];
    X[i] = s / A[i * k + i];
  }
}

// One DIIS step per active problem given g = f(x) at its current x
// (MPS.cpp:263-282; the start plays x0, so the first step is x_x0[0]).
// History per problem: x0 | xR[m] | x_x0[m] | A[m*m] | best x. With
// backtracking (diis_backtrack) a non-finite residual, or one kGrow times the
// best so far, sends the problem back to its best point with its history
// cleared and its step halved; below kEtaFloor times the initial step it
// stops there.
constexpr double kGrow = 100.0;
constexpr double kEtaFloor = 1e-4;
// One DIIS step of problem b by the workgroup that owns it; returns true
// (in every thread) when the problem is done. Every thread calls it.
template <int BT>
__device__ bool batch_diis(const BatchProb& p, int b, double* __restrict__ x, const double* __restrict__ g,
                           double* __restrict__ hist, BatchState* __restrict__ S, int m, double eta, double tol,
                           int max_iter, int backtrack) {
  __shared__ double red[4 * (BT / 64)];
  __shared__ double dts[kMaxDiis + 1];
  __shared__ double Ap[kMaxDiis * kMaxDiis], rhs[kMaxDiis], coef[kMaxDiis];
  __shared__ int decision, save_best, finished;
  const long long nv = 2LL * p.n;
  double* xb = x + p.xoff;
  const double* gb = g + p.xoff;
  double* x0 = hist + p.hoff;
  double* xR = x0 + nv;
  double* xx0 = xR + (long long)m * nv;
  double* Aij = xx0 + (long long)m * nv;
  double* xbest = Aij + (long long)m * m;
  const BatchState st = S[b];
  const bool first = st.k_diis == 0;
  const int it = st.it + 1;  // evaluations so far (the reset state holds -1)
  const int k = it % m;
  const int kd = first ? 1 : std::min(st.k_diis + 1, m);
  const int t = threadIdx.x;
  const double eta_p = backtrack ? st.eta : eta;

  // new point x - eta g and its step; |g|^2 and the step's own product
  double gg = 0.0, dd = 0.0;
  for (long long q = t; q < nv; q += BT) {
    const double xi = xb[q], gi = gb[q];
    const double xn = xi - eta_p * gi;
    const double d = xn - xi;
    x0[q] = xi;
    xR[(long long)k * nv + q] = xn;
    xx0[(long long)k * nv + q] = d;
    gg += gi * gi;
    dd += d * d;
  }
  {
    double two[4] = {gg, dd, 0.0, 0.0};
    block_reduce4<BT>(two, red);
    if (t == 0) {
      dts[kMaxDiis] = two[0];
      dts[kd - 1] = two[1];
    }
  }
  // the products of the step with the kept steps, four per pass (a thread
  // reads back only the step entries it wrote itself)
  for (int i0 = 0; i0 < kd - 1; i0 += 4) {
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
    for (long long q = t; q < nv; q += BT) {
      const double d = xx0[(long long)k * nv + q];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (i0 + j < kd - 1) {
          int posi = k - kd + 1 + i0 + j;
          if (posi < 0) posi += m;
          a4[j] += xx0[(long long)posi * nv + q] * d;
        }
      }
    }
    block_reduce4<BT>(a4, red);
    if (t == 0)
      for (int j = 0; j < 4 && i0 + j < kd - 1; j++) dts[i0 + j] = a4[j];
  }
  __syncthreads();
  if (t == 0) {
    const double* dots = dts;
    const double err = sqrt(dots[kMaxDiis]);
    BatchState ns = st;
    ns.err = err;
    // 0: combine, 1: stop at the evaluated point, 2: take the plain step,
    // 3: return to the best point (history cleared)
    int v = 0, save = 0;
    if (backtrack && isfinite(st.best) && (!isfinite(err) || err > kGrow * st.best)) {
      v = 3;
      ns.err = st.best;
      ns.k_diis = 0;
      ns.eta = 0.5 * st.eta;
      if (ns.eta < kEtaFloor * eta) ns.done = 3;
    } else if (!isfinite(err)) {
      ns.done = 3;
      v = 1;
    } else if (err <= tol) {
      ns.done = 1;
      v = 1;
    } else if (first) {
      v = 2;
    }
    if (backtrack && v != 3 && err < st.best) {
      ns.best = err;
      save = 1;
    }
    ns.it = it;
    if (v == 3) {
      if (it >= max_iter) ns.done = 2;
    } else if (first) {
      ns.k = 0;
      ns.k_diis = 1;
      // A[0][0] is never filled in: the first step's self product is not
      // computed in MPS.cpp:263-270 either
      for (int i = 0; i < m * m; i++) Aij[i] = 0.0;
      if (v == 2 && it >= max_iter) ns.done = 2;
    } else if (v == 0) {
      if (st.k_diis + 1 > m)  // drop the oldest row and column
        for (int i = 1; i < m; i++)
          for (int j = 1; j < m; j++) Aij[m * (i - 1) + j - 1] = Aij[m * i + j];
      for (int i = 0; i < kd; i++) Aij[i + m * (kd - 1)] = Aij[kd - 1 + m * i] = dots[i];
      for (int i = 0; i < kd; i++)
        for (int j = 0; j < kd; j++) Ap[j + kd * i] = Aij[j + m * i];
      diis_solve(Ap, rhs, kd, coef);
      double nc = 0;
      for (int i = 0; i < kd; i++) nc += coef[i];
      for (int i = 0; i < kd; i++) coef[i] /= nc;
      ns.k = k;
      ns.k_diis = kd;
      if (it >= max_iter) ns.done = 2;
    }
    S[b] = ns;
    decision = v;
    save_best = save;
    finished = ns.done != 0;
  }
  __syncthreads();
  const int vd = decision;
  const bool done = finished != 0;
  if (save_best)
    for (long long q = t; q < nv; q += BT) xbest[q] = x0[q];  // x0 holds the evaluated point
  if (vd == 3) {
    for (long long q = t; q < nv; q += BT) xb[q] = xbest[q];
  } else if (vd == 2) {
    for (long long q = t; q < nv; q += BT) xb[q] = xR[(long long)k * nv + q];
  } else if (vd == 0) {
    for (long long q = t; q < nv; q += BT) {
      double s = 0;
      for (int i = 0; i < kd; i++) {
        int posi = k - kd + 1 + i;
        if (posi < 0) posi += m;
        s += coef[i] * xR[(long long)posi * nv + q];
      }
      xb[q] = s;
    }
  }  // vd == 1: x stays the point whose residual was measured
  // the shared decision words are rewritten by the next step's thread 0 only
  // after every thread has read them here
  __syncthreads();
  return done;
}

// The whole solve in one launch: each workgroup owns one problem and loops
// residual -> DIIS step until that problem is done (converged, at the
// iteration limit, non-finite or given up). Problems are independent, so no
// workgroup ever waits on another and the launch ends when the slowest
// problem does; the host neither synchronises nor relaunches per iteration.
template <int BT>
__global__ void __launch_bounds__(BT) k_batch_solve(const BatchProb* __restrict__ P, double* __restrict__ x,
                                                    double* __restrict__ f, double* __restrict__ store,
                                                    double* __restrict__ phi_out, double* __restrict__ lngn,
                                                    double* __restrict__ hist, BatchState* __restrict__ S, int m,
                                                    double eta, double tol, int max_iter, int backtrack, int ldsw) {
  extern __shared__ double lds[];
  const int b = blockIdx.x;
  const BatchProb p = P[b];
  while (true) {
    batch_residual<BT>(p, b, x, f, store, phi_out, lngn, lds, ldsw);
    __syncthreads();  // f and the rank buffers are read by other threads below
    if (batch_diis<BT>(p, b, x, f, hist, S, m, eta, tol, max_iter, backtrack)) break;
  }
}

__global__ void k_batch_reset(BatchState* __restrict__ S, int n, double eta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) S[i] = BatchState{-1, 0, 0, 0, 0.0, eta, INFINITY};
}

// ---------------------------------------------------------------------------
Batch::~Batch() {
  void* ptrs[] = {d_prob, d_x, d_f, d_store, d_hist, d_phi, d_lngn, d_state};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (stream) (void)hipStreamDestroy(stream);
}

void Batch::create(const sfx_adsorption_desc* d, int count, const sfx_options* opt) {
  SFX_REQUIRE(d && count >= 1, SFX_E_INVALID, "empty batch");
  SFX_REQUIRE(!opt || opt->precision == SFX_FP64, SFX_E_UNSUPPORTED,
              "batched adsorption problems are computed in FP64 (rank scaling cannot span their range in FP32)");
  B = count;
  probs.resize(B);
  long long xo = 0, so = 0, po = 0;
  n_max = 0;
  for (int i = 0; i < B; i++) {
    const sfx_adsorption_desc& q = d[i];
    SFX_REQUIRE(q.n_layers >= 2 && q.n_layers <= kMaxLayers, SFX_E_UNSUPPORTED, "n_layers must be in 2..1024");
    SFX_REQUIRE(q.N >= 2, SFX_E_INVALID, "chain length must be >= 2 (the solvent is the monomer)");
    SFX_REQUIRE(q.lambda > 0 && q.lambda <= 0.5, SFX_E_INVALID, "lambda must be in (0, 1/2]");
    SFX_REQUIRE(q.phibulk > 0 && q.phibulk < 1, SFX_E_INVALID, "phibulk must be in (0, 1)");
    SFX_REQUIRE(q.chi_as != 0 || q.chi_s != 0, SFX_E_UNSUPPORTED,
                "with chi_AS = chi_s = 0 polymer and solvent share one iteration variable");
    BatchProb& p = probs[i];
    p.n = q.n_layers;
    p.N = q.N;
    p.lambda = q.lambda;
    p.chi_as = q.chi_as;
    p.chi_s = q.chi_s;
    p.phib = q.phibulk;
    p.xoff = xo;
    p.soff = so;
    p.poff = po;
    p.hoff = 0;
    xo += 2LL * p.n;
    so += (long long)(p.N / 2) * (p.n + 2);
    po += 2LL * (p.n + 2);
    n_max = std::max(n_max, p.n);
  }
  nvar = xo;
  store_words = so;
  phi_words = po;
  SFX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  SFX_HIP(hipMalloc(&d_prob, B * sizeof(BatchProb)));
  SFX_HIP(hipMalloc(&d_x, nvar * sizeof(double)));
  SFX_HIP(hipMalloc(&d_f, nvar * sizeof(double)));
  SFX_HIP(hipMalloc(&d_store, store_words * sizeof(double)));
  SFX_HIP(hipMalloc(&d_phi, phi_words * sizeof(double)));
  SFX_HIP(hipMalloc(&d_lngn, B * sizeof(double)));
  SFX_HIP(hipMalloc(&d_state, B * sizeof(BatchState)));
  SFX_HIP(hipMemcpy(d_prob, probs.data(), B * sizeof(BatchProb), hipMemcpyHostToDevice));
  SFX_HIP(hipMemset(d_x, 0, nvar * sizeof(double)));
  tune();
}

void Batch::launch_residual(int bt, const int* active) {
  const int ldsw = n_max + 2;
  const size_t shm = 2 * (size_t)ldsw * sizeof(double);
  switch (bt) {
    case 64:
      k_batch_residual<64><<<B, 64, shm, stream>>>(d_prob, d_x, d_f, d_store, d_phi, d_lngn, active, ldsw);
      break;
    case 128:
      k_batch_residual<128><<<B, 128, shm, stream>>>(d_prob, d_x, d_f, d_store, d_phi, d_lngn, active, ldsw);
      break;
    case 256:
      k_batch_residual<256><<<B, 256, shm, stream>>>(d_prob, d_x, d_f, d_store, d_phi, d_lngn, active, ldsw);
      break;
    default:
      k_batch_residual<512><<<B, 512, shm, stream>>>(d_prob, d_x, d_f, d_store, d_phi, d_lngn, active, ldsw);
      break;
  }
  SFX_HIP(hipGetLastError());
}

void Batch::launch_solve(int bt, int m, double eta, double tol, int max_iter, int backtrack) {
  const int ldsw = n_max + 2;
  const size_t shm = 2 * (size_t)ldsw * sizeof(double);
#define SOLVE(BT)                                                                                              \
  k_batch_solve<BT><<<B, BT, shm, stream>>>(d_prob, d_x, d_f, d_store, d_phi, d_lngn, d_hist, d_state, m, eta, \
                                            tol, max_iter, backtrack, ldsw)
  switch (bt) {
    case 64: SOLVE(64); break;
    case 128: SOLVE(128); break;
    case 256: SOLVE(256); break;
    default: SOLVE(512); break;
  }
#undef SOLVE
  SFX_HIP(hipGetLastError());
}

// Threads per problem: time one residual of the whole batch at 64, 128, 256
// and 512 threads (at the current x) and keep the fastest. A solve iteration
// is one residual plus a DIIS step of a few vector passes, so the residual
// decides.
void Batch::tune() {
  hipEvent_t e0, e1;
  SFX_HIP(hipEventCreate(&e0));
  SFX_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  block = 256;
  for (int bt : {64, 128, 256, 512}) {
    launch_residual(bt, nullptr);  // warm-up
    SFX_HIP(hipEventRecord(e0, stream));
    for (int r = 0; r < 3; r++) launch_residual(bt, nullptr);
    SFX_HIP(hipEventRecord(e1, stream));
    SFX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SFX_HIP(hipEventElapsedTime(&ms, e0, e1));
    tuned_ms[bt == 64 ? 0 : bt == 128 ? 1 : bt == 256 ? 2 : 3] = ms / 3;
    if (ms < best) {
      best = ms;
      block = bt;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void Batch::ensure_history(int m) {
  if (hist_m == m) return;
  if (d_hist) SFX_HIP(hipFree(d_hist));
  d_hist = nullptr;
  long long ho = 0;
  for (int i = 0; i < B; i++) {
    probs[i].hoff = ho;
    ho += 2LL * probs[i].n * (2 + 2LL * m) + (long long)m * m;  // x0, xR[m], x_x0[m], A, best x
  }
  SFX_HIP(hipMalloc(&d_hist, ho * sizeof(double)));
  SFX_HIP(hipMemcpy(d_prob, probs.data(), B * sizeof(BatchProb), hipMemcpyHostToDevice));
  hist_m = m;
}

void Batch::residual(const double* x, double* f, int where) {
  const bool host = where == SFX_HOST;
  SFX_HIP(hipMemcpyAsync(d_x, x, nvar * sizeof(double), host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                         stream));
  launch_residual(block, nullptr);
  SFX_HIP(hipMemcpyAsync(f, d_f, nvar * sizeof(double), host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                         stream));
  SFX_HIP(hipStreamSynchronize(stream));
}

void Batch::solve(const sfx_solver_opts& o, double* x, int where, int* iters, double* norms, sfx_solve_stats* out) {
  SFX_REQUIRE(o.method == SFX_SOLVER_DIIS, SFX_E_UNSUPPORTED, "batched problems are solved with DIIS");
  SFX_REQUIRE(o.diis_m >= 1 && o.diis_m <= kMaxDiis, SFX_E_INVALID, "diis_m must be in 1..15");
  SFX_REQUIRE(o.max_iter >= 0, SFX_E_INVALID, "max_iter must be >= 0");
  ensure_history(o.diis_m);
  const bool host = where == SFX_HOST;
  SFX_HIP(hipMemcpyAsync(d_x, x, nvar * sizeof(double), host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                         stream));
  k_batch_reset<<<grid1(B, 256), 256, 0, stream>>>(d_state, B, o.diis_eta);
  SFX_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  SFX_HIP(hipEventCreate(&e0));
  SFX_HIP(hipEventCreate(&e1));
  SFX_HIP(hipEventRecord(e0, stream));
  // one launch: every problem stops after at most max_iter + 1 evaluations
  launch_solve(block, o.diis_m, o.diis_eta, o.tolerance, o.max_iter, o.diis_backtrack);
  SFX_HIP(hipEventRecord(e1, stream));
  SFX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SFX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  last.resize(B);
  SFX_HIP(hipMemcpyAsync(last.data(), d_state, B * sizeof(BatchState), hipMemcpyDeviceToHost, stream));
  SFX_HIP(hipMemcpyAsync(x, d_x, nvar * sizeof(double), host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                         stream));
  SFX_HIP(hipStreamSynchronize(stream));
  sfx_solve_stats st{};
  int conv = 0;
  double worst = 0;
  for (int i = 0; i < B; i++) {
    if (iters) iters[i] = last[i].it;
    if (norms) norms[i] = last[i].err;
    st.iterations = std::max(st.iterations, last[i].it);
    st.residual_evals += last[i].it + 1;
    conv += last[i].done == 1;
    worst = std::max(worst, std::isfinite(last[i].err) ? last[i].err : (double)INFINITY);
  }
  st.residual_norm = worst;
  st.seconds = ms * 1e-3;
  st.converged = conv == B;
  if (out) *out = st;
}

}  // namespace sfx

// ---- extern "C" boundary of the batch (include/sfx.h) -----------------------
using sfx::Error;
using sfx::guarded;

extern "C" {

int sfx_batch_create(const sfx_adsorption_desc* problems, int count, const sfx_options* opt, sfx_batch** out) {
  return guarded([&] {
    SFX_REQUIRE(problems && out, SFX_E_INVALID, "NULL argument");
    auto* h = new sfx_batch();
    try {
      h->b.create(problems, count, opt);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int sfx_batch_destroy(sfx_batch* b) {
  delete b;
  return SFX_OK;
}

int sfx_batch_sizes(const sfx_batch* b, long long* nvar, long long* offsets) {
  return guarded([&] {
    SFX_REQUIRE(b && nvar, SFX_E_INVALID, "NULL argument");
    *nvar = b->b.nvar;
    if (offsets) {
      for (int i = 0; i < b->b.B; i++) offsets[i] = b->b.probs[i].xoff;
      offsets[b->b.B] = b->b.nvar;
    }
  });
}

int sfx_batch_tuning(const sfx_batch* b, int* threads, float* ms) {
  return guarded([&] {
    SFX_REQUIRE(b && threads, SFX_E_INVALID, "NULL argument");
    *threads = b->b.block;
    if (ms)
      for (int i = 0; i < 4; i++) ms[i] = b->b.tuned_ms[i];
  });
}

int sfx_batch_residual(sfx_batch* b, const double* x, double* f, int where) {
  return guarded([&] {
    SFX_REQUIRE(b && x && f, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(where == SFX_HOST || where == SFX_DEVICE, SFX_E_INVALID, "where must be SFX_HOST or SFX_DEVICE");
    b->b.residual(x, f, where);
  });
}

int sfx_batch_solve(sfx_batch* b, const sfx_solver_opts* o, double* x, int where, int* iterations, double* norms,
                    sfx_solve_stats* stats) {
  return guarded([&] {
    SFX_REQUIRE(b && o && x, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(where == SFX_HOST || where == SFX_DEVICE, SFX_E_INVALID, "where must be SFX_HOST or SFX_DEVICE");
    sfx_solve_stats st{};
    b->b.solve(*o, x, where, iterations, norms, &st);
    if (stats) *stats = st;
    if (!st.converged) throw Error{SFX_E_NOTCONVERGED, "some problems did not converge"};
  });
}

int sfx_batch_get_lngn(const sfx_batch* b, double* out) {
  return guarded([&] {
    SFX_REQUIRE(b && out, SFX_E_INVALID, "NULL argument");
    SFX_HIP(hipMemcpy(out, b->b.d_lngn, b->b.B * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int sfx_batch_get_phi(const sfx_batch* b, int problem, int state, double* out) {
  return guarded([&] {
    SFX_REQUIRE(b && out, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(problem >= 0 && problem < b->b.B, SFX_E_INVALID, "problem index out of range");
    SFX_REQUIRE(state == 0 || state == 1, SFX_E_INVALID, "state must be 0 (A) or 1 (S)");
    const sfx::BatchProb& p = b->b.probs[problem];
    const long long M = p.n + 2;
    SFX_HIP(hipMemcpy(out, b->b.d_phi + p.poff + state * M, M * sizeof(double), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
