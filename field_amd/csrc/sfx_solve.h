// sfx_solve.h -- device vector helpers and solver entry points.
#pragma once

#include "sfx_system.h"

namespace sfx {

struct DotList {
  const double* a[16];
  const double* b[16];
  int count;
};
struct VecList {
  const double* v[32];
  double c[32];  // coefficients of combine(), passed by value to the kernel
  int count;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n_) { alloc(n_); }
  void alloc(size_t n_) {
    n = n_;
    if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) throw Error{SFX_E_OOM, "hipMalloc failed"};
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void reset(size_t n_) {
    if (p) (void)hipFree(p);
    p = nullptr;
    alloc(n_);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// vector operations of length n on one stream; reductions are deterministic
struct VecOps {
  static constexpr int kHostWords = 128;
  long long n = 0;
  int nblk = 0;
  hipStream_t stream = nullptr;
  double* partial = nullptr;
  double* dres = nullptr;
  double* hres = nullptr;  // pinned host words for results read back
  void init(long long n, hipStream_t st);  // allocates only when n changes
  ~VecOps();
  void dots(const DotList& dl, double* host_out);  // waits for the results
  // results stay on the device; with mirror_at >= 0 also in hres[mirror_at..]
  void dots_dev(const DotList& dl, double* dout, int mirror_at = -1);
  void wait_host(int count, double* host_out);  // synchronise, then copy hres[0..count) out
  void fetch(const double* dsrc, int count, double* host_out);  // one copy, then wait
  double dot(const double* a, const double* b);
  void combine(const VecList& vl, double* y);  // y = sum_k vl.c[k] vl.v[k]
  // y = v_last - sum_{k < count-1} dc[k] v_k, dc on the device (Gram-Schmidt)
  void subtract_projections(const VecList& vl, const double* dc, double* y);
  void axpby(double a, const double* x, double b, double* y);
  void scale(double a, const double* x, double* y);  // y = a x (y's old contents never read)
  void zero(double* y);                             // y = 0 (stores zeros, clears NaN/Inf)
  void sub(const double* x, const double* y, double* z);
  void copy(const double* x, double* y);
};

// the solvers' device workspace, kept by the System between solves so that a
// solve allocates and frees nothing once the sizes have been seen
struct SolveWork {
  VecOps V;
  DevBuf<double> vecs;  // nvec vectors of length n
  DevBuf<double> gs;    // device-side Gram-Schmidt coefficients
  long long n = 0;
  size_t nvec = 0;
  double* vec(size_t k) { return vecs.p + k * (size_t)n; }
};
SolveWork& solve_work(System& S, long long n, size_t nvec);

void diis_coeffs(const std::vector<double>& A, int n, std::vector<double>& X);
void solve_diis(System& S, const sfx_solver_opts& o, sfx_solve_stats& st);
void solve_nk(System& S, const sfx_solver_opts& o, sfx_solve_stats& st);
constexpr long long kNewtonMaxVars = 16384;  // a 1 GiB float pseudo-Hessian
void solve_newton(System& S, const sfx_solver_opts& o, sfx_solve_stats& st);
struct NewtonParams;
struct NewtonResult;
NewtonParams newton_params(const sfx_solver_opts& o);
void fill_stats(const NewtonResult& R, sfx_solve_stats& st);

}  // namespace sfx
