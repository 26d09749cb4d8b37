// sfx_batch.h -- many independent 1-D adsorption SCF problems solved concurrently.
#pragma once

#include <vector>

#include "sfx_internal.h"

namespace sfx {

// One problem as the kernels see it (copied to the device once).
struct BatchProb {
  int n, N;            // layers, chain length
  double lambda;
  double chi_as, chi_s;
  double phib;         // homopolymer bulk volume fraction
  long long xoff;      // offset of its iteration vector (2n values)
  long long soff;      // offset of its propagator store (N/2 ranks x (n+2))
  long long hoff;      // offset of its DIIS history
  long long poff;      // offset of its phi output (2 x (n+2))
};

// DIIS bookkeeping per problem (MPS.cpp:263-282 variables).
struct BatchState {
  int it, k, k_diis, done;  // done: 0 running, 1 converged, 2 iteration limit, 3 non-finite / given up
  double err;
  double eta, best;         // this problem's step and best residual norm (backtracking)
};

struct Batch {
  int B = 0;
  int n_max = 0;
  int block = 0;          // threads per problem, chosen by tune()
  float tuned_ms[4] = {0, 0, 0, 0};  // residual time at 64, 128, 256, 512 threads
  long long nvar = 0, store_words = 0, phi_words = 0;
  std::vector<BatchProb> probs;
  BatchProb* d_prob = nullptr;
  double* d_x = nullptr;
  double* d_f = nullptr;
  double* d_store = nullptr;
  double* d_hist = nullptr;
  double* d_phi = nullptr;
  double* d_lngn = nullptr;
  BatchState* d_state = nullptr;
  int hist_m = 0;
  hipStream_t stream = nullptr;
  // results of the last solve
  std::vector<BatchState> last;
  int residual_launches = 0;

  ~Batch();
  void create(const sfx_adsorption_desc* d, int count, const sfx_options* opt);
  void tune();
  void residual(const double* x, double* f, int where);
  void solve(const sfx_solver_opts& o, double* x, int where, int* iters, double* norms, sfx_solve_stats* st);

 private:
  void ensure_history(int m);
  void launch_residual(int bt, const int* active);
  void launch_solve(int bt, int m, double eta, double tol, int max_iter, int backtrack);
};

}  // namespace sfx

struct sfx_batch {
  sfx::Batch b;
};
