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
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// vector operations of length n on one stream; reductions are deterministic
struct VecOps {
  long long n = 0;
  int nblk = 0;
  hipStream_t stream = nullptr;
  double* partial = nullptr;
  double* dres = nullptr;
  double* dcoef = nullptr;
  void init(long long n, hipStream_t st);
  ~VecOps();
  void dots(const DotList& dl, double* host_out);
  double dot(const double* a, const double* b);
  void combine(const VecList& vl, const double* host_c, double* y);
  void axpby(double a, const double* x, double b, double* y);
  void sub(const double* x, const double* y, double* z);
  void copy(const double* x, double* y);
};

void diis_coeffs(const std::vector<double>& A, int n, std::vector<double>& X);
void solve_diis(System& S, const sfx_solver_opts& o, sfx_solve_stats& st);
void solve_nk(System& S, const sfx_solver_opts& o, sfx_solve_stats& st);

}  // namespace sfx
