// sfx_system.h -- one SCF system on one GPU (host side).
#pragma once

#include <algorithm>
#include <array>
#include <memory>
#include <utility>
#include <vector>

#include "sfx_chain1d.h"
#include "sfx_kernels.h"

namespace sfx {

// Where the forward ranks 1..n of a chain live while the backward pass needs
// them: every rank (c == 0), or every c-th rank (c ~ sqrt(n)) plus one segment
// of up to c-1 ranks recomputed from its checkpoint when the backward pass
// enters it. sqrt(n) checkpointing replaces Matrix1Long's cube-root scheme
// (SF_Homopol1stO.cpp:49-51, 731-843): 2 sqrt(n) slots for one extra forward
// pass.
struct ChainPlan {
  int n = 0, c = 0, nck = 0, sets = 1;
  // gran: c is a multiple of it, so segments start where launches of several
  // ranks start; such launches may need two consecutive segments at once, so
  // recomputed segments alternate between two sets of slots
  void init(int n_, bool checkpoint, int gran = 1) {
    n = n_;
    c = 0;
    nck = 0;
    sets = gran > 1 ? 2 : 1;
    if (!checkpoint || n < 4) return;
    c = 1;
    while (c * c < n) c++;
    c = (c + gran - 1) / gran * gran;
    if (c >= n) {
      c = 0;
      return;
    }
    nck = (n + c - 1) / c;
  }
  int seglen() const { return std::max(2, c - 1); }
  int slots() const { return c ? nck + sets * seglen() : n; }
  int seg(int i) const { return nck + i; }
  int seg_in(int set, int i) const { return nck + set * seglen() + i; }
};

// Forward pass over ranks 2..n; rank 1 is already in slot 0.
// step(s, src_slot, dst_slot) computes rank s from rank s-1.
template <class Step>
void chain_forward(const ChainPlan& P, Step&& step) {
  if (!P.c) {
    for (int s = 2; s <= P.n; s++) step(s, s - 2, s - 1);
    return;
  }
  int prev = 0;
  for (int s = 2; s <= P.n; s++) {
    const int dst = (s - 1) % P.c == 0 ? (s - 1) / P.c : P.seg(s & 1);
    step(s, prev, dst);
    prev = dst;
  }
}

// Slot of rank s for a backward pass that asks for ranks in non-increasing
// order; recomputes the segment that holds s when it is not resident.
template <class Step>
struct ChainReader {
  const ChainPlan& P;
  Step& step;
  int resident = -1;
  int get(int s) {
    if (!P.c) return s - 1;
    const int j = (s - 1) / P.c, start = 1 + j * P.c;
    if (s == start) return j;
    if (resident != j) {
      const int end = std::min(start + P.c - 1, P.n);
      int prev = j;
      for (int r = start + 1; r <= end; r++) {
        const int dst = P.seg(r - start - 1);
        step(r, prev, dst);
        prev = dst;
      }
      resident = j;
    }
    return P.seg(s - start - 1);
  }
};

struct Molecule {
  int N = 1, freedom = SFX_FREE, kind = 1;  // kind: 0 monomer, 1 full sweep, 2 half-stored, 3 second order
  std::vector<int> seq;
  std::vector<int> uses;  // per state: 1 if the chain contains it
  double phib = 0, theta = 0, chem_int_ref = 0;
  RangeBox range{};       // range_restricted: the box and the amount in it
  double phi_range = 0;
  int nstore = 0;         // forward ranks the backward pass needs
  ChainPlan plan;         // how they are kept
  MolParams params(int index) const {
    return MolParams{index, N, freedom, kind, phib, theta};
  }
  unsigned uses_mask() const {  // one bit per state the chain contains
    unsigned u = 0;
    for (size_t t = 0; t < uses.size(); t++)
      if (uses[t]) u |= 1u << t;
    return u;
  }
};

// HIP-event timing of propagator launches (include/sfx_profile.h)
struct Profiler {
  bool on = false;
  unsigned select = ~0u;  // bit 2*kind + precision: that pair is timed
  bool open = false;      // begin() recorded a start event that end() closes
  bool wants(int kind, int prec) const { return on && (select >> (2 * kind + prec) & 1u); }
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Mark {
    int kind, prec;
    size_t ev;
    int ranks;
  };
  std::vector<Mark> marks;
  void begin(int kind, int prec, hipStream_t st, int ranks = 1);
  void end(hipStream_t st);
  // a start/stop pair for a launch that records them itself (hipExtLaunchKernelGGL)
  std::pair<hipEvent_t, hipEvent_t> events(int kind, int prec, int ranks);
  ~Profiler();
};

struct Tangent;  // JVP / VJP buffers (sfx_adjoint.hip)
void tangent_free(Tangent* t);
struct Precond;
struct SolveWork;  // solver vectors (sfx_solve.hip)
void solve_work_free(SolveWork* w);

struct System {
  Lattice lat;
  int ns = 0, nm = 0, form = SFX_FORM_SFBOX, n_itvar = 0, precision = SFX_FP64;
  size_t budget = 0;
  double stiffness = 0;
  std::vector<double> chi, fprof, melt_phib;
  std::vector<int> frozen, st_itvar, itvar_rep;
  std::vector<Molecule> mols;
  std::vector<long long> sites;  // iteration sites (SF_Solve layout order)
  std::vector<unsigned char> itmask;
  long long nvar = 0;
  int maxN = 1, maxstore = 1;
  int store_slots = 1;     // rank slots of the propagator store (all molecules share it)
  bool checkpointed = false;
  bool use_block3 = false;  // blocked 3-D simple-cubic sweeps (sfx_options.sweep_kernel)
  bool use_chain1 = false;  // 1-D FP64 chains in one launch per stretch (sfx_options.sweep_kernel)
  // per molecule and stretch (0 forward, 1 backward, 2 half-store second half):
  // the step list on the device, built on the first sweep (its pointers never change)
  std::vector<std::array<Chain1Step*, 3>> chain1_dev;
  std::vector<std::array<int, 3>> chain1_n;
  // one launch of the stretch `which` of molecule m (building its list on
  // first use from the host steps)
  void chain1_run(int m, int which, const std::vector<Chain1Step>& host, const double* in0, double* last);
  SysParams sp{};
  hipStream_t stream = nullptr;      // where every call of this system runs
  hipStream_t own_stream = nullptr;  // created with the system, used unless a caller sets one

  // device buffers
  double* d_u = nullptr;      // n_itvar * M
  double* d_f = nullptr;      // n_itvar * M (internal residual layout)
  double* d_swf = nullptr;    // ns * M
  float* d_swf32 = nullptr;   // ns * M
  double* d_phi_m = nullptr;  // nm * ns * M
  void* d_acc = nullptr;      // nm * ns * M accumulators (double or float)
  void* d_acc2 = nullptr;     // the backward chain's, in a middle-out sweep (blocked lattices)
  double* d_phi = nullptr;    // ns * M
  double* d_rho = nullptr;    // ns * M
  double* d_phitot = nullptr; // M
  double* d_fprof = nullptr;  // ns * M
  int* d_frozen = nullptr;
  int* d_st_itvar = nullptr;
  int* d_mol_uses = nullptr;  // nm * ns
  unsigned char* d_itmask = nullptr;
  long long* d_sites = nullptr;
  void* d_store = nullptr;    // maxstore * M propagator words
  void* d_pp = nullptr;       // 2 * M running propagator
  void* d_pp2 = nullptr;      // 2 * M, the backward chain's in a middle-out sweep
  double* d_molscal = nullptr;  // nm * MS_COUNT
  double* d_lnf = nullptr;      // nm * (maxN+2)
  double* d_lnb = nullptr;
  unsigned* d_maxf = nullptr;
  unsigned* d_maxb = nullptr;
  double* d_partial = nullptr;  // reduction partials
  unsigned* d_gmax = nullptr;   // ns x kStepWords: lattice maximum of each FP32 SWF (blocked sweeps)
  double* d_phib_state = nullptr;
  double* d_cbulk = nullptr;
  double* d_scalar = nullptr;   // misc device scalars
  double* d_xabi = nullptr;     // nvar: ABI-layout staging
  double* d_abi_in = nullptr;   // n_itvar * M each: the direction and the product of
  double* d_abi_out = nullptr;  // sfx_jvp / sfx_vjp / sfx_precondition, allocated on first use
  int npart = 0;
  Tangent* tan = nullptr;
  Precond* precond = nullptr;  // Newton-Krylov preconditioner, built on first use
  SolveWork* work = nullptr;   // solver vectors, allocated on first use
  Profiler prof;
  bool phi_valid = false;
  int last_precision = SFX_FP64;

  ~System();
  void create(const sfx_lattice_desc& ld, const sfx_system_desc& sd, const sfx_options* opt);
  long long M() const { return lat.d.M; }
  // replaces chi_ab = chi_ba between solves (a Vary-style parameter step)
  void set_chi(int a, int b, double value);

  // x (ABI layout, host or device) -> d_u
  void load_x(const double* x, int where);
  // d_f -> f (ABI layout, host or device)
  void store_f(const double* src_internal, double* f, int where);
  // the two internal-layout vectors of the product entry points (kept, so a
  // call allocates nothing after the first)
  void abi_vectors(double*& in, double*& out);
  void compute_phi(int precision);  // from d_u
  void residual_internal(int precision);  // d_u -> d_f
  void finish_range(int m);  // range_restricted: the range condition fixes molecule m's constant
  bool has_range() const;
  // residual_internal replayed from a HIP graph captured on the first call
  // (per precision): one submission instead of a launch per rank. Plain
  // launches while kernel timing is on. Kernels take their parameters by
  // value, so set_chi drops the graphs.
  void residual_replay(int precision);
  void drop_graphs();
  hipGraphExec_t res_graph[2] = {nullptr, nullptr};
  bool res_graph_gmax[2] = {false, false};
  bool graph_ok = true;  // false once a capture was refused: plain launches from then on
  long long graph_replays = 0;  // residuals replayed from a graph so far
  double free_energy();
  double grand_potential();
  double chemical_potential(int m);
  void sync();
  // later calls run on st (nullptr: the system's own stream); work already
  // queued on the previous stream is finished first
  void set_stream(hipStream_t st);

  template <typename T, typename A>
  void sweep(int m, bool scaled);
  template <typename T>
  void sweep_blocked(int m, bool scaled);
  // a first-order chain with a full store on a blocked lattice is swept
  // middle-out: its forward and backward chains advance in the same launches.
  // The base and the tangent sweeps must agree on this (their compositions
  // share the reference log).
  bool middle_out(const Molecule& mo) const {
    return use_block3 && schedule == SFX_SCHEDULE_AUTO && mo.kind == 1 && !mo.plan.c && mo.N >= 2;
  }
  int schedule = SFX_SCHEDULE_AUTO;  // sfx_options.chain_schedule
  bool gmax_valid = false;           // d_gmax holds the maxima of the current FP32 SWF
};

// Fourier-space Newton-Krylov preconditioner (sfx_precond.hip)
bool precond_supported(const System& S);
Precond* precond_create(const System& S);
void precond_apply(Precond* p, const System& S, const double* r, double* out, double sigma);
void precond_update(Precond* p, const System& S);  // after a chi change
void precond_free(Precond* p);

}  // namespace sfx
