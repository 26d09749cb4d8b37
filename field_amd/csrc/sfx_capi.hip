// sfx_capi.hip -- the extern "C" boundary (include/sfx.h). No exception
// crosses it: every entry point converts failures into a status code and a
// thread-local message.
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <new>
#include <string>

#include "../../include/sfx_profile.h"
#include "sfx_newton.h"
#include "sfx_solve.h"

namespace sfx {

static thread_local std::string g_msg;
static thread_local int g_code = 0;

void set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_msg = buf;
  g_code = code;
}
int last_code() { return g_code; }


// implemented in sfx_adjoint.hip
void jvp_internal(System& S, const double* v_internal, double* jv_internal, int precision);
void vjp_internal(System& S, const double* w_internal, double* jtw_internal, int precision);
void load_vec(System& S, const double* x, int where, double* dst_internal);

}  // namespace sfx

using namespace sfx;

struct sfx_system {
  System s;
};

extern "C" {

int sfx_abi_version(void) { return SFX_ABI_VERSION; }

int sfx_last_error(char* buf, size_t size) {
  if (!buf || size == 0) return SFX_E_INVALID;
  std::strncpy(buf, g_msg.c_str(), size - 1);
  buf[size - 1] = 0;
  return SFX_OK;
}

int sfx_set_device(int device) {
  return guarded([&] { SFX_HIP(hipSetDevice(device)); });
}

int sfx_device_count(int* count) {
  return guarded([&] {
    SFX_REQUIRE(count, SFX_E_INVALID, "count is NULL");
    SFX_HIP(hipGetDeviceCount(count));
  });
}

int sfx_system_create(const sfx_lattice_desc* lat, const sfx_system_desc* sys, const sfx_options* opt,
                      sfx_system** out) {
  return guarded([&] {
    SFX_REQUIRE(lat && sys && out, SFX_E_INVALID, "NULL descriptor");
    *out = nullptr;
    auto* h = new sfx_system();
    try {
      h->s.create(*lat, *sys, opt);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int sfx_system_destroy(sfx_system* s) {
  return guarded([&] { delete s; });
}

int sfx_system_sizes(const sfx_system* s, long* nvar, long* M) {
  return guarded([&] {
    SFX_REQUIRE(s, SFX_E_INVALID, "NULL system");
    if (nvar) *nvar = (long)s->s.nvar;
    if (M) *M = (long)s->s.lat.d.M;
  });
}

int sfx_compute_phi(sfx_system* s, const double* x, int where, int precision) {
  return guarded([&] {
    SFX_REQUIRE(s && x, SFX_E_INVALID, "NULL argument");
    s->s.load_x(x, where);
    s->s.compute_phi(precision);
    s->s.sync();
  });
}

int sfx_residual(sfx_system* s, const double* x, double* f, int where, int precision) {
  return guarded([&] {
    SFX_REQUIRE(s && x && f, SFX_E_INVALID, "NULL argument");
    s->s.load_x(x, where);
    s->s.residual_internal(precision);
    s->s.store_f(s->s.d_f, f, where);
  });
}

int sfx_jvp(sfx_system* s, const double* x, const double* v, double* jv, int where, int precision) {
  return guarded([&] {
    SFX_REQUIRE(s && x && v && jv, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    S.load_x(x, where);
    double *vin, *out;
    S.abi_vectors(vin, out);
    load_vec(S, v, where, vin);
    jvp_internal(S, vin, out, precision);
    S.store_f(out, jv, where);
  });
}

int sfx_vjp(sfx_system* s, const double* x, const double* w, double* jtw, int where, int precision) {
  return guarded([&] {
    SFX_REQUIRE(s && x && w && jtw, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    S.load_x(x, where);
    double *win, *out;
    S.abi_vectors(win, out);
    load_vec(S, w, where, win);
    vjp_internal(S, win, out, precision);
    S.store_f(out, jtw, where);
  });
}

int sfx_get_phi(sfx_system* s, int state, double* phi_out, int where) {
  return guarded([&] {
    SFX_REQUIRE(s && phi_out, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    SFX_REQUIRE(state >= 0 && state < S.ns, SFX_E_INVALID, "state out of range");
    SFX_REQUIRE(S.phi_valid, SFX_E_INVALID, "no phi computed yet");
    SFX_HIP(hipMemcpyAsync(phi_out, S.d_phi + (size_t)state * S.lat.d.M, S.lat.d.M * sizeof(double),
                           where == SFX_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, S.stream));
    S.sync();
  });
}

static int get_molscal(sfx_system* s, int molecule, int which, double* out) {
  return guarded([&] {
    SFX_REQUIRE(s && out, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    SFX_REQUIRE(molecule >= 0 && molecule < S.nm, SFX_E_INVALID, "molecule out of range");
    SFX_REQUIRE(S.phi_valid, SFX_E_INVALID, "no phi computed yet");
    SFX_HIP(hipMemcpyAsync(out, S.d_molscal + (size_t)molecule * MS_COUNT + which, sizeof(double),
                           hipMemcpyDeviceToHost, S.stream));
    S.sync();
  });
}

int sfx_get_lngn(sfx_system* s, int molecule, double* lngn) { return get_molscal(s, molecule, MS_LNGN, lngn); }
int sfx_get_phibulk(sfx_system* s, int molecule, double* phibulk) {
  return get_molscal(s, molecule, MS_PHIB, phibulk);
}

int sfx_free_energy(sfx_system* s, double* F) {
  return guarded([&] {
    SFX_REQUIRE(s && F, SFX_E_INVALID, "NULL argument");
    *F = s->s.free_energy();
  });
}

int sfx_grand_potential(sfx_system* s, double* omega) {
  return guarded([&] {
    SFX_REQUIRE(s && omega, SFX_E_INVALID, "NULL argument");
    *omega = s->s.grand_potential();
  });
}

int sfx_chemical_potential(sfx_system* s, int molecule, double* mu) {
  return guarded([&] {
    SFX_REQUIRE(s && mu, SFX_E_INVALID, "NULL argument");
    *mu = s->s.chemical_potential(molecule);
  });
}

int sfx_set_chi(sfx_system* s, int a, int b, double value) {
  return guarded([&] {
    SFX_REQUIRE(s, SFX_E_INVALID, "NULL argument");
    s->s.set_chi(a, b, value);
  });
}

int sfx_stream_set(sfx_system* s, void* stream) {
  return guarded([&] {
    SFX_REQUIRE(s, SFX_E_INVALID, "NULL argument");
    s->s.set_stream((hipStream_t)stream);
  });
}

int sfx_solver_defaults(sfx_solver_opts* o) {
  if (!o) return SFX_E_INVALID;
  std::memset(o, 0, sizeof *o);
  o->method = SFX_SOLVER_DIIS;
  o->tolerance = 1e-7;    // SF_Solve.cpp:1836, MPS.cpp:13
  o->max_iter = 1000;     // MPS.cpp:273
  o->diis_m = 10;         // MPS.cpp:306
  o->diis_eta = 0.1;      // MPS.cpp:13
  o->gmres_restart = 30;
  o->gmres_max_iter = 200;
  o->gmres_rtol = 1e-3;
  o->jvp_precision = SFX_FP32;
  o->fp64_switch = 1e-3;
  o->precondition = 1;
  o->ptc_dt0 = 1.0;
  o->verbose = 0;
  o->diis_backtrack = 0;
  o->newton_deltamax = 0.5;          // SF_Solve.cpp:1888 (Scheutjens line search)
  o->newton_linesearch_limit = 20;   // sfnewton.cpp:31
  o->newton_reset_criterion = 1e3;   // SF_Solve.cpp:2176
  return SFX_OK;
}

int sfx_precondition(sfx_system* s, const double* r, double* out, int where) {
  return guarded([&] {
    SFX_REQUIRE(s && r && out, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    SFX_REQUIRE(precond_supported(S), SFX_E_UNSUPPORTED,
                "the Fourier-space preconditioner needs a periodic simple-cubic melt");
    if (!S.precond) S.precond = precond_create(S);
    double *a, *b;
    S.abi_vectors(a, b);
    load_vec(S, r, where, a);
    precond_apply(S.precond, S, a, b, 0.0);
    S.store_f(b, out, where);
  });
}

int sfx_solve(sfx_system* s, const sfx_solver_opts* o, double* x, int where, sfx_solve_stats* stats) {
  return guarded([&] {
    SFX_REQUIRE(s && o && x, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    sfx_solve_stats st{};
    S.load_x(x, where);
    const long long replays0 = S.graph_replays;
    if (o->method == SFX_SOLVER_DIIS)
      solve_diis(S, *o, st);
    else if (o->method == SFX_SOLVER_NK)
      solve_nk(S, *o, st);
    else if (o->method == SFX_SOLVER_NEWTON)
      solve_newton(S, *o, st);
    else
      throw Error{SFX_E_INVALID, "unknown solver method"};
    S.store_f(S.d_u, x, where);
    st.graph_replays = (int)(S.graph_replays - replays0);
    if (stats) *stats = st;
    if (!st.converged) throw Error{SFX_E_NOTCONVERGED, "solver reached its iteration limit"};
  });
}

int sfx_newton_iterate(sfx_residual_fn f, void* ctx, double* x, long n, const sfx_solver_opts* o,
                       sfx_solve_stats* stats) {
  return guarded([&] {
    SFX_REQUIRE(f && x && o, SFX_E_INVALID, "NULL argument");
    SFX_REQUIRE(n >= 2 && n <= kNewtonMaxVars, SFX_E_INVALID, "n must be in 2..16384");
    const NewtonParams P = newton_params(*o);
    const auto t0 = std::chrono::steady_clock::now();
    NewtonResult R;
    try {
      R = pseudo_newton([&](const double* xv, double* g) { f(ctx, xv, g); }, x, (int)n, P);
    } catch (const std::bad_alloc&) {
      throw Error{SFX_E_OOM, "no host memory for the pseudo-Hessian"};
    }
    sfx_solve_stats st{};
    fill_stats(R, st);
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    if (!st.converged) throw Error{SFX_E_NOTCONVERGED, "Newton stopped before the tolerance"};
  });
}

// ---- kernel timing (include/sfx_profile.h) ----------------------------------
int sfx_profile_start(sfx_system* s) { return sfx_profile_start_select(s, ~0u); }

int sfx_profile_start_select(sfx_system* s, unsigned select) {
  return guarded([&] {
    SFX_REQUIRE(s, SFX_E_INVALID, "NULL system");
    Profiler& p = s->s.prof;
    p.marks.clear();
    p.used = 0;
    p.select = select;
    p.open = false;
    p.on = true;
  });
}

int sfx_profile_stop(sfx_system* s, sfx_profile_report* rep) {
  return guarded([&] {
    SFX_REQUIRE(s && rep, SFX_E_INVALID, "NULL argument");
    System& S = s->s;
    Profiler& p = S.prof;
    S.sync();
    p.on = false;
    std::memset(rep, 0, sizeof *rep);
    // algorithmic words per site-rank: first order reads the previous rank and
    // g and writes the new rank (compose adds the partner and the accumulator's
    // read and write; the tangent doubles the propagator words and adds v);
    // second order moves all six direction components of every propagator
    const double words1[SFX_K_COUNT] = {3, 6, 6, 10};
    const double words2[SFX_K_COUNT] = {13, 21, 26, 46};
    const double* words = S.lat.d.order == 2 ? words2 : words1;
    for (int k = 0; k < SFX_K_COUNT; k++)
      for (int pr = 0; pr < 2; pr++) {
        sfx_profile_entry& e = rep->entries[k * 2 + pr];
        e.kind = k;
        e.precision = pr;
      }
    for (const Profiler::Mark& m : p.marks) {
      float ms = 0;
      SFX_HIP(hipEventElapsedTime(&ms, p.pool[m.ev], p.pool[m.ev + 1]));
      sfx_profile_entry& e = rep->entries[m.kind * 2 + m.prec];
      e.launches++;
      e.ranks += m.ranks;
      e.total_ms += ms;
    }
    for (sfx_profile_entry& e : rep->entries)
      if (e.launches)
        e.bytes_per_launch = words[e.kind] * (e.precision == SFX_FP32 ? 4.0 : 8.0) * (double)S.lat.d.V *
                             (double)e.ranks / (double)e.launches;
    rep->count = 2 * SFX_K_COUNT;
    p.marks.clear();
    p.used = 0;
  });
}

// ---- per-rank lattice primitives --------------------------------------------
int sfx_lattice_create(const sfx_lattice_desc* d, sfx_lattice** out) {
  return guarded([&] {
    SFX_REQUIRE(d && out, SFX_E_INVALID, "NULL argument");
    auto* h = new sfx_lattice();
    try {
      lattice_init(h->l, *d);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int sfx_lattice_destroy(sfx_lattice* l) {
  return guarded([&] { delete l; });
}

int sfx_lattice_sites(const sfx_lattice* l, long* M) {
  return guarded([&] {
    SFX_REQUIRE(l && M, SFX_E_INVALID, "NULL argument");
    *M = (long)l->l.d.M;
  });
}

}  // extern "C"
