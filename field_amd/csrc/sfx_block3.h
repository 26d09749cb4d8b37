// sfx_block3.h -- several propagator ranks per launch on the 3-D simple-cubic
// lattice (temporal blocking through LDS tiles).
#pragma once

#include <utility>
#include <vector>

#include "sfx_kernels.h"

namespace sfx {

constexpr int kB3Ranks = 4;     // ranks per launch at most
constexpr int kB3TanRanks = 2;  // ranks per tangent or FP64 launch at most (that region fits two workgroups per CU)
// ranks per launch at most of a variant
template <typename T>
constexpr int b3_max_ranks(bool tangent) {
  return (tangent || sizeof(T) == 8) ? kB3TanRanks : kB3Ranks;
}

// Ranks s0+1 .. s0+nr of one chain, all of one segment state.
template <typename T>
struct B3Args {
  const T* in;                 // rank s0 (stored scale)
  const T* din;                // its tangent (tangent sweeps)
  const T* g;                  // SWF of the state
  const T* v;                  // tangent direction of the state
  T* out[kB3Ranks];            // where rank s0+k goes (nullptr: not kept)
  T* dout[kB3Ranks];
  const T* pair[kB3Ranks];     // composition partner of rank s0+k
  const T* dpair[kB3Ranks];
  T* acc;                      // accumulator of the state (phi, or dphi for tangents)
  int nr;
  // FP32 log-scaling: the input is divided by max(rank s0) (maxv[s0]) and
  // every rank by the lattice maximum of g (gmax); lnv[s] is the log of the
  // divisor accumulated up to rank s, lnpair[pidx[k]] that of the partner
  unsigned* maxv;
  double* lnv;
  const double* lnpair;
  const double* ref;
  const unsigned* gmax;        // step-max slots of the state's lattice maximum of g
  int s0;
  int pidx[kB3Ranks];
};

// One launch advancing up to two independent chains (the forward and the
// backward chain of a molecule): workgroups [0, nb) take chain 0, [nb, 2 nb)
// chain 1. Both run the same kernel variant (tangent / scaled / connect).
template <typename T>
struct B3Launch {
  B3Args<T> a[2];
  int nchains;
};

// Splits chain steps a..b (ascending) into launches: a launch ends at a grid
// line ((step - g0 + 1) % grid == 0; grid is b3_max_ranks of the variant), before a step of another segment
// state, or before a step that split() marks. f(first, last) per launch. The
// grid makes the partition of any stretch that starts on a grid line equal
// to that of the whole chain, so a recomputed stretch reproduces it bit for bit.
template <class State, class Split, class F>
void chain_blocks(int a, int b, int g0, State state, Split split, F f, int grid = kB3Ranks) {
  int s = a;
  while (s <= b) {
    int e = s;
    while (e < b && (e + 1 - g0) % grid != 0 && state(e + 1) == state(s) && !split(e + 1)) e++;
    f(s, e);
    s = e + 1;
  }
}

// the launches of chain_blocks(a, b, g0, state, no extra splits) as a list
template <class State>
std::vector<std::pair<int, int>> chain_launches(int a, int b, int g0, State state, int grid = kB3Ranks) {
  std::vector<std::pair<int, int>> out;
  chain_blocks(
      a, b, g0, state, [](int) { return false; }, [&](int s, int e) { out.emplace_back(s, e); }, grid);
  return out;
}

// true when the blocked sweep applies: 3-D simple cubic, first order, every
// side at least kB3Ranks + 1 layers (one reflection reaches the interior)
bool block3_ok(const DevLattice& L);

// ev0/ev1 (optional): timing events the dispatch records at kernel start and end
template <typename T>
void launch_block3(const DevLattice& L, const B3Args<T>& a, bool tangent, bool scaled, bool connect,
                   hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
template <typename T>
void launch_block3(const DevLattice& L, const B3Launch<T>& la, bool tangent, bool scaled, bool connect,
                   hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// max over the interior of each state's FP32 SWF, into step-max slots
// gmax + t * kStepWords (zeroed here; read with get_step_max(.., 0))
// zeroed: gmax[0, ns * kStepWords) is already 0 on st (launch_swf can clear it)
void launch_gmax(const DevLattice& L, const float* swf32, int ns, unsigned* gmax, hipStream_t st,
                 bool zeroed = false);

}  // namespace sfx
