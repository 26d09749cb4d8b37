// sfx_kernels.h -- launch wrappers and kernel parameter blocks.
#pragma once

#include "sfx_internal.h"

#define SFX_MAXS 8  // segment states per system
#define SFX_MAXM 8  // molecules per system

namespace sfx {

// per-molecule device scalars
enum { MS_LNGN = 0, MS_LNC, MS_PHIB, MS_PREF, MS_REF, MS_RANGE_C, MS_COUNT = 8 };

// FP32 per-rank log-scaling context of one sweep. Step k writes max|G| of its
// output into maxv[k] and the cumulative log scale into lnv[k]; its output is
// multiplied by 1/maxv[k-1]. Composition multiplies by
// exp(lnv[k] + lnpair[pair] - *ref). maxv == nullptr disables scaling.
// The maximum of a step is spread over kMaxSlots words (workgroup b writes
// word b mod kMaxSlots), each in a cache line of its own (kSlotStride words
// apart): atomics on one line serialise, and a per-rank launch on a large
// lattice has thousands of workgroups. A reader reduces the words with one
// wave load. A step occupies kStepWords words.
constexpr int kMaxSlots = 64;
constexpr int kSlotStride = 32;  // 128-byte lines
constexpr int kStepWords = kMaxSlots * kSlotStride;

__device__ __forceinline__ void put_step_max(unsigned* maxv, int step, float m) {
  atomicMax(maxv + (long long)step * kStepWords + (blockIdx.x & (kMaxSlots - 1)) * kSlotStride,
            __float_as_uint(m));
}

// the maxima of two steps (a second pointer may be null: its value is 0),
// read once per workgroup by its first wave (one slot line per lane) and shared
// through LDS; every thread of the workgroup must call it
__device__ __forceinline__ void block_step_max2(const unsigned* maxa, int sa, const unsigned* maxb, int sb,
                                                float& ma, float& mb) {
  __shared__ float sh[2];
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    float a = __uint_as_float(maxa[(long long)sa * kStepWords + lane * kSlotStride]);
    float b = maxb ? __uint_as_float(maxb[(long long)sb * kStepWords + lane * kSlotStride]) : 0.f;
    for (int o = 32; o > 0; o >>= 1) {
      a = fmaxf(a, __shfl_xor(a, o, 64));
      b = fmaxf(b, __shfl_xor(b, o, 64));
    }
    if (lane == 0) {
      sh[0] = a;
      sh[1] = b;
    }
  }
  __syncthreads();
  ma = sh[0];
  mb = sh[1];
}

// called by every lane of a wave; all lanes get the maximum
__device__ __forceinline__ float get_step_max(const unsigned* maxv, int step) {
  const int lane = threadIdx.x & 63;
  float v = lane < kMaxSlots ? __uint_as_float(maxv[(long long)step * kStepWords + lane * kSlotStride]) : 0.f;
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

struct RankScale {
  unsigned* maxv;
  double* lnv;
  const double* lnpair;
  const double* ref;
  int step;
  int pair;
  const unsigned* pmaxv;  // step maxima of the partner chain (second order recovers its rank)
};

struct MolParams {
  int index, N, freedom, kind;  // kind: 0 monomer, 1 full sweep, 2 half-stored (palindrome), 3 second order
  double phib, theta;
};

struct SysParams {
  int ns, nm, n_itvar, melt;
  double chi[SFX_MAXS * SFX_MAXS];
  int frozen[SFX_MAXS];
  int st_itvar[SFX_MAXS];
  int itvar_rep[SFX_MAXS];
  double melt_phib[SFX_MAXS];
  double mol_frac[SFX_MAXM * SFX_MAXS];
  double mol_N[SFX_MAXM];
  double mol_ref[SFX_MAXM];
};

// up to eight device arrays of 4-byte words cleared by one launch
struct ZeroList {
  void* p[8];
  long long words[8];
  int n = 0;
};
void launch_zero_list(const ZeroList& z, hipStream_t st);
// adds an array, launching the list first when it is full
void zero_list_add(ZeroList& z, void* p, long long words, hipStream_t st);

template <typename T, typename A>
void launch_prop(const DevLattice& L, const T* Gin, const T* g, T* Gout, const T* Gpair, A* acc,
                 const RankScale& sc, hipStream_t st);
template <typename T>
void launch_init_rank(const DevLattice& L, const T* g, T* out, const RankScale& sc, hipStream_t st);
// rank 1 of two chains (the forward and backward ends of a middle-out sweep) in one launch
template <typename T>
struct InitRank {
  const T* g;
  T* out;
  RankScale sc;
};
template <typename T>
void launch_init_rank2(const DevLattice& L, const InitRank<T>& c0, const InitRank<T>& c1, hipStream_t st);
template <typename T, typename A>
void launch_connect(const DevLattice& L, const T* a, const T* b, A* acc, const RankScale& sc, double half,
                    hipStream_t st);
template <typename T>
void launch_copy_interior(const DevLattice& L, const T* a, T* b, hipStream_t st);
template <typename T>
int launch_wsum(const DevLattice& L, const T* G, double* partial, hipStream_t st);
void launch_mol_finalize(const MolParams& mp, const double* partial, int nparts, double* scal,
                         const double* lnfN, const double* allscal, int nmol, hipStream_t st);
// phi_m = pref (acc + acc2) / g; acc2 (the backward chain's accumulator of a
// middle-out sweep) may be null
template <typename A>
void launch_mol_phi(const DevLattice& L, const A* acc, const double* g, const double* scal, int monomer,
                    double* phi_m, hipStream_t st, const A* acc2 = nullptr);
// phi_m of every state set in `uses` in one launch (state-major arrays from the given bases)
template <typename A>
void launch_mol_phi_states(const DevLattice& L, unsigned uses, int ns, const A* acc, const A* acc2, const double* g,
                           const double* scal, double* phi_m, hipStream_t st);
// scal[MS_REF] = lna[ia] + lnb[ib] (0 when lna is null): the reference log of
// a middle-out sweep's compositions
void launch_set_ref_pair(double* scal, const double* lna, int ia, const double* lnb, int ib, hipStream_t st);
// zero != null: the same launch also clears zero[0..nzero)
void launch_swf(const DevLattice& L, const double* u, const unsigned char* itmask, int nstates,
                const int* st_itvar, double* swf, float* swf32, hipStream_t st, unsigned* zero = nullptr,
                int nzero = 0);
// sp != null: the same launch also writes the bulk fractions and ChemIntBulk
void launch_phi_total(const DevLattice& L, int nstates, int nmol, const double* phi_m, const int* mol_uses,
                      const double* fprof, const int* frozen, double* phi, double* phitot, hipStream_t st,
                      const SysParams* sp = nullptr, const double* molscal = nullptr, double* phib_state = nullptr,
                      double* cbulk = nullptr);
void launch_sides(const DevLattice& L, int nstates, const double* phi, double* rho, hipStream_t st);

void launch_residual(const DevLattice& L, const SysParams& sp, const unsigned char* itmask, const double* swf,
                     const double* rho, const double* phitot, const double* cbulk, double* f, hipStream_t st);
int launch_free_energy(const DevLattice& L, const SysParams& sp, const double* phi, const double* phi_m,
                       const double* rho, const double* swf, const double* u, const double* molscal,
                       double* partial, hipStream_t st);
int launch_grand_potential(const DevLattice& L, const SysParams& sp, const double* phi, const double* phi_m,
                           const double* rho, const double* swf, const double* molscal, const double* phib_state,
                           const double* cbulk, double* partial, hipStream_t st);
void launch_sum_partials(const double* partial, int n, double* out, hipStream_t st);

// range_restricted molecules (SF_Copol1stO.cpp:1246-1280): phi_m of the used
// states summed over the interior sites of the box lo..hi (partial sums per
// block), then C = phi_range / sum into the molecule's scalars (pref, ln C,
// phi_b = N C) and phi_m *= C
struct RangeBox {
  int lo[3], hi[3];
};
int launch_range_sum(const DevLattice& L, RangeBox box, unsigned uses, int ns, const double* phi_m, double* partial,
                     hipStream_t st);
void launch_range_scale(const DevLattice& L, unsigned uses, int ns, const double* partial, int nparts,
                        double phi_range, int N, double* scal, double* phi_m, hipStream_t st);

}  // namespace sfx
