// sfx_newton.hip -- pseudo-Hessian Newton with the Scheutjens line search
// (host code: the factorisation is updated one row at a time, a chain of n
// dependent steps that one CPU core runs faster than a GPU launch for the
// sizes this method is for).
//
// Reference: SFNewton::iterate with pseudohessian on and the Scheutjens line
// search (sfnewton.cpp:844-1143): newhessian (:1433-1521), startderivatives
// (:1523-1530), direction (:1916-1940), zero / linesearch / linecriterion /
// stepchange / newtrustregion (:1946-2038), residue (:2040-2042); the
// factorisation routines updatpos / gausa / gausb / multiply /
// signdeterminant (newttool.cpp:12-75,170-243); and the reset rules of
// SF_Solve::inneriteration (SF_Solve.cpp:659-760).
#include "sfx_newton.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "sfx_solve.h"

namespace sfx {

void Ldu::resize(int n_) {
  n = n_;
  lo.assign((size_t)n * n, 0.f);
  up.assign((size_t)n * n, 0.f);
  d.assign(n, 0.f);
}

void Ldu::reset(float diag) {
  std::fill(lo.begin(), lo.end(), 0.f);
  std::fill(up.begin(), up.end(), 0.f);
  std::fill(d.begin(), d.end(), diag);
}

int Ldu::sign() const {
  int s = 1;
  for (int i = 0; i < n; i++)
    if (d[i] < 0) s = -s;
  return s;
}

// forward substitution with L in row order, then D and U from the last row up
// (the summation order of gausa / gausb)
void Ldu::solve(const double* g, double* p) const {
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < i; j++) s += lo[(size_t)j * n + i] * p[j];
    p[i] = -g[i] - s;
  }
  for (int i = n - 1; i >= 0; i--) {
    const float* u = &up[(size_t)i * n];
    double s = 0;
    for (int j = i + 1; j < n; j++) s += u[j] * p[j];
    p[i] = p[i] / d[i] - s;
  }
}

void Ldu::multiply(const double* w, double alpha, double* v) const {
  std::vector<double> t(n);
  for (int i = 0; i < n; i++) {
    const float* u = &up[(size_t)i * n];
    double s = 0;
    for (int j = i + 1; j < n; j++) s += w[j] * u[j];
    t[i] = (s + w[i]) * d[i];
    s = 0;
    for (int j = 0; j < i; j++) s += t[j] * lo[(size_t)j * n + i];
    v[i] = alpha * (s + t[i]);
  }
}

// rank-one update of a factorisation, one pivot at a time: pivot i fixes
// D(i), then corrects column i of L and row i of U and carries the rest of w
// and v to the later pivots
void Ldu::update(double* w, double* v, double alpha) {
  for (int i = 0; i < n; i++) {
    float* l = &lo[(size_t)i * n];
    float* u = &up[(size_t)i * n];
    const double vi = v[i], wi = w[i];
    double dd = d[i];
    const double piv = dd + (alpha * wi) * vi;
    d[i] = (float)piv;
    dd /= piv;
    const double cl = vi * alpha / piv, cu = wi * alpha / piv;
    alpha *= dd;
    for (int j = i + 1; j < n; j++) {
      const double wj = w[j], vj = v[j];
      w[j] -= wi * l[j];
      l[j] *= dd;
      l[j] += cl * wj;
      v[j] -= vi * u[j];
      u[j] *= dd;
      u[j] += cu * vj;
    }
  }
}

namespace {

double norm(const double* a, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * a[i];
  return std::sqrt(s);
}
double dotp(const double* a, const double* b, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

class Iteration {
 public:
  Iteration(const std::function<void(const double*, double*)>& f, double* x, int n, const NewtonParams& prm)
      : f_(f), x_(x), n_(n), P(prm), x0_(n), g_(n), g0_(n), p_(n), p0_(n), y_(n), hp_(n) {
    H.resize(n);
    deltamax_ = P.deltamax;
    deltamin_ = P.deltamin > 0 && P.deltamin <= P.deltamax ? P.deltamin : 0.0;
    trustregion_ = deltamax_;
  }

  NewtonResult run() {
    evaluate();
    minimum_ = sq(norm(g_.data(), n_));
    new_direction();
    double ng = std::sqrt(minimum_);
    while (!done(ng) && iterations_ < P.max_iter) {
      iterations_++;
      lineiterations_ = 0;
      new_trust_region();
      const double bound = trustregion_ / (norm(p_.data(), n_) + 1.0 / std::pow(2.0, 53));
      std::memcpy(x0_.data(), x_, sizeof(double) * n_);
      std::memcpy(g0_.data(), g_.data(), sizeof(double) * n_);
      alpha_ = bound < 1 ? bound : 1;
      trial(alpha_);
      trustfactor_ *= step_change();
      trustfactor_ *= alpha_ / bound;
      new_hessian();
      new_direction();
      ng = std::sqrt(minimum_);
    }
    R.iterations = iterations_;
    R.accuracy = accuracy_;
    R.residual_norm = ng;
    R.converged = !done_pending(ng);
    return R;
  }

 private:
  static double sq(double a) { return a * a; }
  // the reference's loop test: stop when accuracy and 10 tolerance bound
  // accuracy and |g|, or when accuracy is not a number
  bool done_pending(double ng) const {
    return P.tolerance < accuracy_ || P.tolerance * 10 < ng || accuracy_ != std::fabs(accuracy_);
  }
  bool done(double ng) const {
    return !(P.tolerance < accuracy_ || P.tolerance * 10 < ng) || accuracy_ != std::fabs(accuracy_);
  }

  void evaluate() {
    f_(x_, g_.data());
    R.evaluations++;
  }

  void reset_hessian() {
    trouble_ = 0;
    H.reset((float)(1 + norm(g_.data(), n_)));
    reset_iteration_ = iterations_;
    R.hessian_resets++;
  }

  void set_deltamax(double v) {
    if (v > deltamin_) deltamax_ = v;
  }

  void new_hessian() {
    if (alpha_ == 0 || iterations_ == 0) {
      reset_hessian();
      return;
    }
    const double dmin = 1.0 / std::pow(2.0, 53);
    double py = 0;
    for (int i = 0; i < n_; i++) {
      y_[i] = g_[i] - g0_[i];
      py += p_[i] * y_[i];
    }
    if (!newton_dir_)
      H.multiply(p_.data(), 1, hp_.data());
    else
      for (int i = 0; i < n_; i++) hp_[i] = -g0_[i];
    double php = dotp(p_.data(), hp_.data(), n_);
    double theta = py / (10 * dmin + alpha_ * php);
    if (theta > 0 && iterations_ == reset_iteration_ + 1 && accuracy_ > P.hessian_scaling_accuracy) {
      alpha_ *= theta;
      py /= theta;
      php /= theta;
      for (int i = 0; i < n_; i++) {
        p_[i] /= theta;
        H.d[i] *= theta;
      }
    }
    trustfactor_ *= 4 / (sq(theta - 1) + 1) + 0.5;
    double sum = alpha_ * sq(norm(p_.data(), n_));
    theta = std::fabs(py / (alpha_ * php));
    if (theta < .01)
      sum /= 0.8;
    else if (theta > 100)
      sum *= theta / 50;
    for (int i = 0; i < n_; i++) y_[i] -= alpha_ * hp_[i];
    H.update(y_.data(), p_.data(), 1.0 / sum);
    trouble_ -= H.sign();
    if (trouble_ < 0)
      trouble_ = 0;
    else if (trouble_ >= 3)
      reset_hessian();
  }

  // SF_Solve::inneriteration
  void reset_rules() {
    const double acc = accuracy_;
    const bool number = acc == std::fabs(acc);
    if (acc < min_accuracy_ && iterations_ > 0 && number) min_accuracy_ = acc;
    if (acc > min_accuracy_ * P.reset_criterion && number) {
      reset_hessian();
      min_accuracy_ *= 1.5;
      if (deltamax_ > 0.005) set_deltamax(0.9 * deltamax_);
    }
    small_count_ = alpha_ < P.small_alpha ? small_count_ + 1 : 0;
    if (small_count_ == P.max_small_alpha) {
      small_count_ = 0;
      reset_hessian();
      if (deltamax_ > 0.005) set_deltamax(0.9 * deltamax_);
    }
  }

  void new_direction() {
    reset_rules();
    p0_ = p_;
    if (iterations_ == 0) new_hessian();
    H.solve(g_.data(), p_.data());
    newton_dir_ = H.sign() > 0;
    if (!newton_dir_) {
      for (double& v : p_) v = -v;
      R.reversed_directions++;
    }
    accuracy_ = std::sqrt(norm(p_.data(), n_) * norm(g_.data(), n_) / (1 + norm(x_, n_)));
  }

  // x = x0 + a p and its residual; the fifth evaluation of one line search
  // re-evaluates x0 first and flags a residual that changed
  void trial(double a) {
    lineiterations_++;
    if (lineiterations_ == 5) {
      std::memcpy(x_, x0_.data(), sizeof(double) * n_);
      evaluate();
      for (int i = 0; i < n_; i++) {
        if (g_[i] != g0_[i]) {
          R.time_dependent = true;
          break;
        }
        if (!std::isfinite(g_[i])) {
          a = -a;
          break;
        }
      }
    }
    for (int i = 0; i < n_; i++) x_[i] = x0_[i] + a * p_[i];
    evaluate();
    for (int i = 0; i < n_; i++)
      if (!std::isfinite(g_[i])) {
        g_[i] = 1;
        break;
      }
    minimum_ = sq(norm(g_.data(), n_));
  }

  double line_criterion() const {
    double ng = norm(g0_.data(), n_);
    const double gg0 = dotp(g_.data(), g0_.data(), n_) / ng / ng;
    ng = sq(norm(g_.data(), n_) / ng);
    if ((gg0 > 1 || ng > 1) && ng - gg0 * std::fabs(gg0) < 0.2) ng = 1.5 * ng;
    if (gg0 < 0 && ng < 2) return 1;
    if (ng > 10) return .01;
    return 0.4 * (1 + 0.75 * ng) / (ng - gg0 * std::fabs(gg0) + 0.1);
  }

  double step_change() {
    double crit = line_criterion(), change = crit;
    while (crit < 0.35 && lineiterations_ < P.linesearch_limit) {
      alpha_ /= 4;
      trial(alpha_);
      crit = line_criterion();
      change = 1;
    }
    return change;
  }

  void new_trust_region() {
    const double np0 = norm(p0_.data(), n_);
    if (np0 > 0 && trustregion_ > 2 * alpha_ * np0) trustregion_ = 2 * alpha_ * np0;
    trustregion_ *= trustfactor_;
    trustfactor_ = 1.0;
    if (trustregion_ > deltamax_) trustregion_ = deltamax_;
    if (trustregion_ < deltamin_) trustregion_ = deltamin_;
  }

  const std::function<void(const double*, double*)>& f_;
  double* x_;
  int n_;
  NewtonParams P;
  std::vector<double> x0_, g_, g0_, p_, p0_, y_, hp_;
  Ldu H;
  NewtonResult R;
  double alpha_ = 1, trustregion_ = 0, trustfactor_ = 1, deltamax_ = 0, deltamin_ = 0;
  double accuracy_ = 0, minimum_ = 0, min_accuracy_ = 1e30;
  int iterations_ = 0, lineiterations_ = 0, reset_iteration_ = 0, trouble_ = 0, small_count_ = 0;
  bool newton_dir_ = false;
};

}  // namespace

NewtonResult pseudo_newton(const std::function<void(const double*, double*)>& resid, double* x, int n,
                           const NewtonParams& prm) {
  Iteration it(resid, x, n, prm);
  return it.run();
}

namespace {
// two pinned host vectors of n doubles (hipHostMalloc), freed on scope exit
struct PinnedPair {
  double* x = nullptr;
  double* g = nullptr;
  explicit PinnedPair(size_t n) {
    if (hipHostMalloc((void**)&x, n * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&g, n * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      release();
      throw Error{SFX_E_OOM, "hipHostMalloc failed"};
    }
  }
  void release() {
    if (x) (void)hipHostFree(x);
    if (g) (void)hipHostFree(g);
    x = g = nullptr;
  }
  ~PinnedPair() { release(); }
  PinnedPair(const PinnedPair&) = delete;
  PinnedPair& operator=(const PinnedPair&) = delete;
};
}  // namespace

NewtonParams newton_params(const sfx_solver_opts& o) {
  SFX_REQUIRE(o.newton_deltamax > 0 && o.newton_linesearch_limit >= 1 && o.newton_reset_criterion > 0,
              SFX_E_INVALID, "newton_deltamax, newton_linesearch_limit and newton_reset_criterion must be positive");
  SFX_REQUIRE(o.tolerance > 0 && o.max_iter >= 0, SFX_E_INVALID, "tolerance must be > 0 and max_iter >= 0");
  NewtonParams P;
  P.tolerance = o.tolerance;
  P.max_iter = o.max_iter;
  P.deltamax = o.newton_deltamax;
  P.linesearch_limit = o.newton_linesearch_limit;
  P.reset_criterion = o.newton_reset_criterion;
  return P;
}

void fill_stats(const NewtonResult& R, sfx_solve_stats& st) {
  st.iterations = R.iterations;
  st.residual_evals = R.evaluations;
  st.residual_norm = R.residual_norm;
  st.accuracy = R.accuracy;
  st.converged = R.converged;
}

// sfx_solve with SFX_SOLVER_NEWTON: the iteration vector in the ABI layout
// (the reference's iteration variables), every residual in FP64 on the GPU
void solve_newton(System& S, const sfx_solver_opts& o, sfx_solve_stats& st) {
  const long long n = S.nvar;
  SFX_REQUIRE(n >= 2, SFX_E_INVALID, "Newton needs at least two iteration variables");
  SFX_REQUIRE(n <= kNewtonMaxVars, SFX_E_UNSUPPORTED,
              "the dense pseudo-Hessian Newton takes at most 16384 iteration variables; use SFX_SOLVER_NK");
  const NewtonParams P = newton_params(o);
  std::vector<double> x((size_t)n);
  PinnedPair io((size_t)n);
  hipEvent_t e0, e1;
  SFX_HIP(hipEventCreate(&e0));
  SFX_HIP(hipEventCreate(&e1));
  SFX_HIP(hipEventRecord(e0, S.stream));
  S.store_f(S.d_u, x.data(), SFX_HOST);
  // x and g pass through pinned staging, the residual through its graph
  auto resid = [&](const double* xv, double* g) {
    std::memcpy(io.x, xv, sizeof(double) * (size_t)n);
    S.load_x(io.x, SFX_HOST);
    S.residual_replay(SFX_FP64);
    S.store_f(S.d_f, io.g, SFX_HOST);
    std::memcpy(g, io.g, sizeof(double) * (size_t)n);
  };
  NewtonResult R;
  try {
    R = pseudo_newton(resid, x.data(), (int)n, P);
  } catch (const std::bad_alloc&) {
    throw Error{SFX_E_OOM, "no host memory for the pseudo-Hessian"};
  }
  // leave the system at the returned point: its last evaluation may be a
  // re-evaluation of the line search's start
  S.load_x(x.data(), SFX_HOST);
  SFX_HIP(hipEventRecord(e1, S.stream));
  SFX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SFX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  fill_stats(R, st);
  st.seconds = ms * 1e-3;
}

}  // namespace sfx
