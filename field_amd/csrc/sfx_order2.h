// sfx_order2.h -- launchers of the second-order (semiflexible) 3-D kernels.
#pragma once

#include "sfx_kernels.h"

namespace sfx {

// one rank of the six-component propagator; backward composes acc += G'.Gp
template <typename T>
void launch_prop2(const DevLattice& L, bool backward, bool tangent, const T* Gin, const T* dGin, const T* g,
                  const T* v, T* Gout, T* dGout, const T* Gp, const T* dGp, T* acc, double S, const RankScale& sc,
                  hipStream_t st);
// all components = g (tangent: -v g), optional composition with Gp
template <typename T>
void launch_init2(const DevLattice& L, bool tangent, const T* g, const T* v, T* G, T* dG, const T* Gp,
                  const T* dGp, T* acc, const RankScale& sc, hipStream_t st);
// per-block weighted sums over the six components of G (and dG)
template <typename T>
int launch_wsum6(const DevLattice& L, const T* G, const T* dG, double* p1, double* p2, hipStream_t st);

}  // namespace sfx
