// sfx_nk.hip -- matrix-free Newton-Krylov on the device.
//
// Newton step J(x) dx = -f(x) solved by restarted GMRES whose products J v are
// the exact tangent-propagator JVPs (sfx_adjoint.hip), run in FP32 with
// per-rank log-scaling. The residual f is evaluated with FP32 sweeps while
// ||f|| (measured, or predicted by the step's linear model) is above
// fp64_switch and with FP64 sweeps below it; convergence is
// declared on the FP64 residual only, so the FP32 Jacobian only steers the
// step (iterative refinement). Backtracking on ||f|| globalises the step.
// Where it applies, GMRES is right-preconditioned with the Fourier-space
// Jacobian of the homogeneous melt (sfx_precond.hip): it solves (J P) z = -f
// and steps dx = P z. With ptc_dt0 > 0 the step is pseudo-transient
// continuation instead: (I/dt + J) dx = -f with the preconditioner shifted
// alike, dt grown by the ratio of successive ||f|| (switched evolution
// relaxation), a step that raises ||f|| fivefold or gives a non-finite
// residual rejected and retried with dt/4. Small dt follows the relaxation
// dx/dt = -f that fixed-point iterations such as DIIS follow; large dt is
// Newton.
#include <cmath>

#include "sfx_solve.h"

namespace sfx {

void jvp_at_base(System& S, const double* v, double* jv, int precision);

void solve_nk(System& S, const sfx_solver_opts& o, sfx_solve_stats& st) {
  SFX_REQUIRE(o.fp64_switch >= 0 && std::isfinite(o.fp64_switch), SFX_E_INVALID,
              "fp64_switch must be finite and >= 0 (0: every residual in FP64)");
  const long long n = (long long)S.n_itvar * S.lat.d.M;
  const int m = std::max(1, std::min(o.gmres_restart, 30));
  Precond* pc = nullptr;
  if (o.precondition && precond_supported(S)) {
    if (!S.precond) S.precond = precond_create(S);
    pc = S.precond;
  }
  // with a preconditioner, m more vectors keep P v_k of every Krylov vector
  // and one the step P z built from them
  SolveWork& W = solve_work(S, n, (size_t)m + 6 + (pc ? m + 1 : 0));
  VecOps& V = W.V;
  double* const wv = W.vec(m + 1);  // Arnoldi work vector
  double* const dx = W.vec(m + 2);
  double* const x0 = W.vec(m + 3);
  double* const r = S.d_f;  // the residual of the last evaluation (only eval() writes it)
  double* const pz = W.vec(m + 5);
  auto Zk = [&](int k) { return W.vec((size_t)m + 6 + k); };  // P v_k
  double* const dxp = pc ? W.vec((size_t)2 * m + 6) : nullptr;
  double* const gs = W.gs.p;  // projections of both Gram-Schmidt passes, then ||w||^2
  constexpr int kGsPass = 32;
  double* x = S.d_u;
  const bool ptc = o.ptc_dt0 > 0;
  double dt = ptc ? o.ptc_dt0 : 0.0;
  double sigma = ptc ? 1.0 / dt : 0.0;
  // the GMRES operator v -> (sigma I + J) P v (without a preconditioner P = I);
  // P v goes to pv
  auto op = [&](const double* v, double* out, double* pv) {
    const double* a = v;
    if (pc) {
      precond_apply(pc, S, v, pv, sigma);
      a = pv;
    }
    jvp_at_base(S, a, out, o.jvp_precision);
    if (sigma != 0.0) V.axpby(sigma, a, 1.0, out);
    st.jvp_evals++;
  };
  auto Vk = [&](int k) { return W.vec((size_t)k); };
  hipEvent_t e0, e1;
  SFX_HIP(hipEventCreate(&e0));
  SFX_HIP(hipEventCreate(&e1));
  SFX_HIP(hipEventRecord(e0, S.stream));

  int prec = o.fp64_switch > 0 ? SFX_FP32 : SFX_FP64;
  auto eval = [&](int p) {
    S.residual_replay(p);
    if (p == SFX_FP32)
      st.residual_evals_fp32++;
    else
      st.residual_evals++;
    return std::sqrt(V.dot(r, r));
  };
  double rn = eval(prec);
  int it = 0;
  bool converged = false;
  while (true) {
    if (!std::isfinite(rn)) throw Error{SFX_E_NONFINITE, "non-finite residual in Newton-Krylov"};
    if (prec == SFX_FP32 && rn < o.fp64_switch) {
      prec = SFX_FP64;
      rn = eval(prec);
    }
    if (o.verbose) std::fprintf(stderr, "nk it=%d |f|=%.6e prec=%s\n", it, rn, prec == SFX_FP64 ? "fp64" : "fp32");
    if (prec == SFX_FP64 && rn < o.tolerance) {
      converged = true;
      break;
    }
    if (it >= o.max_iter) break;
    it++;
    // ---- GMRES(m) for J dx = -r, dx0 = 0 ----
    double eta = o.gmres_rtol;
    // cleared by stores, not by scaling: the work vectors are shared with
    // DIIS and kept between solves, and 0 * NaN left there would stay NaN
    V.zero(dx);
    if (pc) V.zero(dxp);
    int total = 0;
    double beta0 = rn, lin = rn;  // lin: ||linear residual|| GMRES leaves
    for (int cycle = 0; cycle < 1 + o.gmres_max_iter / m && total < o.gmres_max_iter; cycle++) {
      // residual of the linear system: r_l = -r - J dx
      double beta;
      if (cycle == 0) {
        V.scale(-1.0 / rn, r, Vk(0));
        beta = rn;
      } else {
        op(dx, wv, pz);
        V.axpby(-1.0, r, -1.0, wv);  // w = -r - J P z
        beta = std::sqrt(V.dot(wv, wv));
        lin = beta;
        if (beta <= eta * beta0) break;
        V.scale(1.0 / beta, wv, Vk(0));
      }
      std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0);
      g[0] = beta;
      int j = 0;
      double res = beta;
      for (; j < m && total < o.gmres_max_iter; j++) {
        op(Vk(j), wv, pc ? Zk(j) : pz);
        total++;
        // classical Gram-Schmidt, twice, with the projections kept on the
        // device; one read-back per Arnoldi step brings both passes and ||w||^2
        std::vector<double> h(j + 1, 0.0), hb(2 * kGsPass + 1);
        for (int pass = 0; pass < 2; pass++) {
          for (int base = 0; base <= j; base += 16) {
            DotList dl{};
            dl.count = std::min(16, j + 1 - base);
            for (int i = 0; i < dl.count; i++) {
              dl.a[i] = Vk(base + i);
              dl.b[i] = wv;
            }
            V.dots_dev(dl, gs + pass * kGsPass + base, pass * kGsPass + base);
          }
          VecList vl{};
          vl.count = j + 2;
          for (int i = 0; i <= j; i++) vl.v[i] = Vk(i);
          vl.v[j + 1] = wv;
          V.subtract_projections(vl, gs + pass * kGsPass, wv);
        }
        {
          DotList dl{};
          dl.count = 1;
          dl.a[0] = dl.b[0] = wv;
          V.dots_dev(dl, gs + 2 * kGsPass, 2 * kGsPass);
        }
        V.wait_host(2 * kGsPass + 1, hb.data());
        for (int i = 0; i <= j; i++) h[i] = (0.0 + hb[i]) + hb[kGsPass + i];
        double hn = std::sqrt(hb[2 * kGsPass]);
        for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = h[i];
        H[(size_t)(j + 1) * m + j] = hn;
        // apply previous rotations, then a new one
        for (int i = 0; i < j; i++) {
          double a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
          H[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
          H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
        }
        double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
        double d = std::hypot(a, b);
        cs[j] = d > 0 ? a / d : 1.0;
        sn[j] = d > 0 ? b / d : 0.0;
        H[(size_t)j * m + j] = d;
        H[(size_t)(j + 1) * m + j] = 0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        res = std::fabs(g[j + 1]);
        if (hn > 0) V.scale(1.0 / hn, wv, Vk(j + 1));
        if (res <= eta * beta0 || hn == 0) {
          j++;
          break;
        }
      }
      // y = H^-1 g, dx += V y
      int k = j;
      std::vector<double> y(k, 0.0);
      for (int i = k - 1; i >= 0; i--) {
        double s = g[i];
        for (int l = i + 1; l < k; l++) s -= H[(size_t)i * m + l] * y[l];
        y[i] = s / H[(size_t)i * m + i];
      }
      VecList vl{};
      vl.count = k + 1;
      for (int i = 0; i < k; i++) {
        vl.v[i] = Vk(i);
        vl.c[i] = y[i];
      }
      vl.v[k] = dx;
      vl.c[k] = 1.0;
      V.combine(vl, dx);
      if (pc) {  // the same combination of the P v_k: P dx without applying P again
        for (int i = 0; i < k; i++) vl.v[i] = Zk(i);
        vl.v[k] = dxp;
        V.combine(vl, dxp);
      }
      lin = res;
      if (o.verbose) std::fprintf(stderr, "  gmres cycle=%d its=%d est=%.3e\n", cycle, j, res / beta0);
      if (res <= eta * beta0) break;
    }
    const double* const step = pc ? dxp : dx;  // P z
    // The linear model predicts ||f|| at the full step: f + J dx, whose norm
    // is the linear residual GMRES left plus, under continuation, at most
    // sigma ||dx||. When that is already below the switch the trial point is
    // evaluated in FP64 at once, rather than in FP32 first and then again in
    // FP64 once the FP32 norm trips the switch.
    if (prec == SFX_FP32 && lin < o.fp64_switch) {
      double pred = lin;
      if (sigma != 0.0) pred += sigma * std::sqrt(V.dot(step, step));
      if (pred < o.fp64_switch) prec = SFX_FP64;
    }
    if (ptc) {
      V.copy(x, x0);
      V.axpby(1.0, step, 1.0, x);
      const double rt = eval(prec);
      if (!std::isfinite(rt) || rt > 5.0 * rn) {
        V.copy(x0, x);
        rn = eval(prec);  // back to the last point (and its base state for the JVPs)
        dt *= 0.25;
        if (o.verbose) std::fprintf(stderr, "  ptc step rejected (|f|=%.3e), dt=%.3e\n", rt, dt);
        if (dt < 1e-8) break;
      } else {
        dt = std::min(dt * rn / rt, 1e12);
        rn = rt;
        if (o.verbose) std::fprintf(stderr, "  ptc dt=%.3e\n", dt);
      }
      sigma = 1.0 / dt;
      continue;
    }
    // ---- backtracking line search on ||f|| ----
    V.copy(x, x0);
    double alpha = 1.0, rt = 0, best = INFINITY, best_alpha = 0;
    bool accepted = false;
    for (int ls = 0; ls < 12; ls++) {
      V.copy(x0, x);
      V.axpby(alpha, step, 1.0, x);
      rt = eval(prec);
      if (std::isfinite(rt) && rt < best) {
        best = rt;
        best_alpha = alpha;
      }
      if (std::isfinite(rt) && rt <= (1 - 1e-4 * alpha) * rn) {
        accepted = true;
        break;
      }
      alpha *= 0.5;
    }
    if (!accepted) {
      // take the best finite trial if it lowers ||f|| at all, else stop at x0
      V.copy(x0, x);
      if (best < rn) {
        V.axpby(best_alpha, step, 1.0, x);
        rt = eval(prec);
      } else {
        rt = eval(prec);
        rn = rt;
        if (prec == SFX_FP32) {
          // FP32 residuals stop improving near their rounding floor; below it
          // only FP64 evaluations can still show descent
          prec = SFX_FP64;
          rn = eval(prec);
          if (o.verbose) std::fprintf(stderr, "  FP32 line search found no descent, continuing in FP64, |f|=%.3e\n", rn);
          continue;
        }
        if (o.verbose) std::fprintf(stderr, "  line search found no descent, |f|=%.3e\n", rt);
        break;
      }
      if (o.verbose) std::fprintf(stderr, "  line search fell back to alpha=%g, |f|=%.3e\n", best_alpha, rt);
    }
    rn = rt;
  }
  SFX_HIP(hipEventRecord(e1, S.stream));
  SFX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  SFX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  st.iterations = it;
  st.residual_norm = rn;
  st.seconds = ms * 1e-3;
  st.converged = converged;
  if (prec != SFX_FP64) {  // report the FP64 residual of the returned point
    S.residual_replay(SFX_FP64);
    st.residual_norm = std::sqrt(V.dot(S.d_f, S.d_f));
  }
}

}  // namespace sfx
