// sfx_chain1d.h -- several ranks of a chain on a 1-D lattice in one launch
// (FP64, unscaled, full store; sfx_chain1d.hip).
#pragma once

#include "sfx_kernels.h"

namespace sfx {

constexpr int kChain1MaxSites = 4000;  // two LDS rank buffers stay below 64 KiB

// one rank step: next = g * stencil(current); stored to out (may be null);
// acc += next * pair when acc is set
struct Chain1Step {
  const double* g;
  double* out;
  const double* pair;
  double* acc;
};

bool chain1d_ok(const DevLattice& L);
// advances the rank in0 through steps[0..n) (a device array); the final rank
// goes to last (may be null)
void launch_chain1d(const DevLattice& L, const double* in0, const Chain1Step* steps, int n, double* last,
                    hipStream_t st);

}  // namespace sfx
