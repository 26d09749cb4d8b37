// sfx_chain1d.hip -- a whole stretch of a chain's ranks on a 1-D lattice in
// one workgroup launch.
//
// A 1-D lattice has at most a few thousand sites, so one rank is a fraction
// of one workgroup's work and a launch per rank (SF_Copol1stO.cpp:1204-1232,
// one PropagateG per rank) is all launch overhead. Here one workgroup keeps
// the running rank in LDS (two ping-pong buffers indexed by layer, halos
// never stored: the stencil reads the remapped interior layer or 0 across a
// surface) and advances it through every step of the list with one barrier
// per rank; a step stores its rank to HBM when the backward pass needs it
// (`out`) and composes acc += G * pair (ConnectG) when it has a partner. The
// per-site arithmetic is that of k_prop, so the results are the same bits.
#include "sfx_chain1d.h"
#include "sfx_device.h"

namespace sfx {

constexpr int kC1Threads = 512;

__global__ void __launch_bounds__(kC1Threads) k_chain1d(DevLattice L, const double* __restrict__ in0,
                                                       const Chain1Step* __restrict__ steps, int n,
                                                       double* __restrict__ last) {
  extern __shared__ double lds[];
  const int nx = L.nx;
  double* cur = lds;
  double* nxt = lds + (nx + 2);
  for (int x = 1 + (int)threadIdx.x; x <= nx; x += kC1Threads) cur[x] = in0[sidx(L, x, 0, 0)];
  __syncthreads();
  for (int k = 0; k < n; k++) {
    const Chain1Step s = steps[k];
    for (int x = 1 + (int)threadIdx.x; x <= nx; x += kC1Threads) {
      const long long r = sidx(L, x, 0, 0);
      const double v = s.g[r] * stencil<double, ST_1D>(L, cur, x, 0, 0);
      nxt[x] = v;
      if (s.out) s.out[r] = v;
      if (s.acc) {
        const double cfac = 1.0;
        s.acc[r] = s.acc[r] + (double)(v * s.pair[r] * cfac);
      }
    }
    __syncthreads();
    double* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (last)
    for (int x = 1 + (int)threadIdx.x; x <= nx; x += kC1Threads) last[sidx(L, x, 0, 0)] = cur[x];
}

bool chain1d_ok(const DevLattice& L) { return L.grad == 1 && L.order == 1 && L.nx >= 1 && L.nx <= kChain1MaxSites; }

void launch_chain1d(const DevLattice& L, const double* in0, const Chain1Step* steps, int n, double* last,
                    hipStream_t st) {
  const size_t lds = 2 * (size_t)(L.nx + 2) * sizeof(double);
  k_chain1d<<<1, kC1Threads, lds, st>>>(L, in0, steps, n, last);
}

}  // namespace sfx
