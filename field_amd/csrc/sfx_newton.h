// sfx_newton.h -- the reference's default SCF iteration: Newton with a dense
// pseudo-Hessian and the Scheutjens line search, for systems small enough to
// hold an n x n float factorisation on the host.
#pragma once

#include <functional>
#include <vector>

namespace sfx {

// L D U factorisation of the pseudo-Hessian, in float as the reference keeps
// it (sfnewton.h:147 `float *h`). L is unit lower, U unit upper. Column i of
// L and row i of U are stored contiguously (lo[i n + j] = L(j, i),
// up[i n + j] = U(i, j) for j > i), the two that the rank-one update walks
// at step i.
struct Ldu {
  int n = 0;
  std::vector<float> lo, up, d;
  void resize(int n_);
  void reset(float diag);                                  // L = U = I, D = diag I
  int sign() const;                                        // sign of det D
  void solve(const double* g, double* p) const;            // p = -(L D U)^-1 g
  void multiply(const double* w, double alpha, double* v) const;  // v = alpha L D U w
  // L D U <- L D U + alpha w v^T; w and v are overwritten
  void update(double* w, double* v, double alpha);
};

struct NewtonParams {
  double tolerance = 1e-7;
  int max_iter = 1000;
  double deltamax = 0.5;              // largest step length (trust region)
  double deltamin = 0.0;
  int linesearch_limit = 20;          // residual evaluations per line search
  double reset_criterion = 1e3;       // reset when accuracy > this x the best so far
  double small_alpha = 1e-5;          // a step this short counts as small
  int max_small_alpha = 50;           // this many small steps in a row reset the Hessian
  double hessian_scaling_accuracy = 0.1;
};

struct NewtonResult {
  int iterations = 0;
  int evaluations = 0;
  int hessian_resets = 0;
  int reversed_directions = 0;
  bool time_dependent = false;  // a re-evaluation at the same point changed the residual
  bool converged = false;
  double accuracy = 0;          // sqrt(|p| |g| / (1 + |x|))
  double residual_norm = 0;     // |g|
};

// Solves g(x) = 0 from x (in/out, n values on the host). resid(x, g) fills g.
NewtonResult pseudo_newton(const std::function<void(const double*, double*)>& resid, double* x, int n,
                           const NewtonParams& prm);

}  // namespace sfx
