// sfx_order2.h -- launchers of the second-order (semiflexible) 3-D kernels.
#pragma once

#include "sfx_kernels.h"

namespace sfx {

// Forward ranks are stored as W = T G (the transfer matrix applied at each
// site): the next forward rank then reads one component per neighbour, and a
// rank's G is recovered from the previous rank's W where it is needed.
// one forward rank W_{s-1} -> W_s
template <typename T>
void launch_prop2f(const DevLattice& L, bool tangent, const T* Win, const T* dWin, const T* g, const T* v, T* Wout,
                   T* dWout, double S, const RankScale& sc, hipStream_t st);
// one backward rank G_{s+1} -> G_s, composed with the forward G_f(s) recovered
// from Wp = W_f(s-1) (null for s = 1, where G_f = g); sc.pmaxv: the forward maxima
template <typename T>
void launch_prop2b(const DevLattice& L, bool tangent, const T* Gin, const T* dGin, const T* g, const T* v, T* Gout,
                   T* dGout, const T* Wp, const T* dWp, T* acc, double S, const RankScale& sc, hipStream_t st);
// forward rank 1 in W form (G = g, tangent -v g)
template <typename T>
void launch_init2f(const DevLattice& L, bool tangent, const T* g, const T* v, T* W, T* dW, double S,
                   const RankScale& sc, hipStream_t st);
// backward rank N (G = g, tangent -v g) composed with G_f(N) recovered from Wp = W_f(N-1)
template <typename T>
void launch_init2b(const DevLattice& L, bool tangent, const T* g, const T* v, T* G, T* dG, const T* Wp,
                   const T* dWp, T* acc, const RankScale& sc, hipStream_t st);
// per-block weighted sums over the six components of G (and dG)
template <typename T>
int launch_wsum6(const DevLattice& L, const T* G, const T* dG, double* p1, double* p2, hipStream_t st);

}  // namespace sfx
