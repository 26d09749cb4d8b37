#!/usr/bin/env python3
"""SCF solves/s on the 3-D simple-cubic 64^3 periodic A12B28 diblock melt
(N = 40, f_A = 0.3, chi = 0.4, u ~ N(0, 0.1) from seed 3), MPS.cpp form.

One step = one SCF solve from the seeded initial field to ||f||_2 < 1e-7 of
the FP64 residual at chi = 0.4. chi = 0.4 lies 0.06% below this lattice's
spinodal (chi_s = 0.4002 for A12B28), where the seeded noise leaves every
local solver crawling through a near-singular region; both forms therefore
solve as a two-step `var` sweep of chi (Vary.cpp; SF_Box.cpp:145-176): chi =
0.38 from the seeded field, then chi = 0.4 from that solution, both inside
the timed region. The headline form is matrix-free Newton-Krylov whose
Jacobian products are FP32 tangent-propagator sweeps (per-rank log-scaling),
with the residual in FP64 near convergence. The plain FP64 form (the DIIS of
MPS.cpp, m = 10, eta = 0.1, all sweeps FP64) runs the same sweep on the same
input and is reported beside it. Inputs are resident on the GPU before the
timed region. --direct adds both forms solving at chi = 0.4 directly, with
their iteration caps.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCF solves/s to |res|<1e-7, 3D 64^3 AB N=40: adjoint+FP32 vs plain FP64 GPU"
PEAK_HBM_GBS = 8000.0


def nk_options(args):
    return dict(method="nk", tolerance=1e-7, max_iter=args.nk_max_iter, gmres_restart=30,
                gmres_max_iter=args.gmres_max_iter, gmres_rtol=args.gmres_rtol, jvp_precision="fp32",
                fp64_switch=args.fp64_switch, precondition=1, ptc_dt0=args.ptc_dt0)


def diis_options(args):
    return dict(method="diis", tolerance=1e-7, max_iter=args.diis_max_iter, diis_m=10, diis_eta=0.1)


def sweep(G, x, path, opts, continuation=None):
    """The chi sweep; returns (x, merged stats)."""
    xs, sts = G.solve_chi_sweep(x, 0, 1, path, check_converged=False, continuation=continuation, **opts)
    keys = ("iterations", "residual_evals", "residual_evals_fp32", "jvp_evals")
    st = {k: sum(s[k] for s in sts) for k in keys}
    st["converged"] = int(len(sts) == len(path) and sts[-1]["converged"])
    st["residual_norm"] = sts[-1]["residual_norm"]
    st["per_chi"] = [{k: s[k] for k in ("chi",) + keys + ("residual_norm",)} for s in sts]
    return xs, st


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "melt64_nk_sweep_hbm_traffic.md")
# the profile entry of the dominant kernel -> its name in the rocprofv3 counter summary
TRAFFIC_KERNEL = {("tangent propagate+compose", "fp32"): "k_block3<float, true, true, true>"}


def measured_traffic(d, n):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy of
    known size), when it covers this kernel and lattice; else None."""
    name = TRAFFIC_KERNEL.get((d["kernel"], d["precision"]))
    if n != 64 or name is None or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith(f"traffic_bytes_per_launch {name} = "):
            return float(line.rsplit("=", 1)[1])
    return None


def dominant(prof):
    """The kernel kind that moves the most algorithmic bytes in the profile
    (launches x bytes per launch). By time the tangent compose and tangent
    forward kernels are within a few percent of each other, so a choice by
    time could flip between runs; by bytes the compose kernel leads by 1.7x."""
    if not prof:
        return None
    d = max(prof, key=lambda e: e["launches"] * e["bytes_per_launch"])
    achieved = d["bytes_per_launch"] / (d["avg_us"] * 1e-6) / 1e9
    return d, achieved


def cpu_baseline(args, evals_needed):
    """The oracle's FP64 MPS.cpp restatement on this host's cores: time a bounded
    number of residual evaluations of the same 64^3 melt and scale by the number
    of residual evaluations the plain DIIS chi sweep needs."""
    from oracle import orc
    from field_amd import configs
    lat, sysd, x0 = configs.melt_3d(n=args.n)
    N = 40
    P = orc.MPS(args.n, 12, 28, 0.4, args.n ** 3 / N)
    P.residual(x0)
    reps, t0 = 0, time.perf_counter()
    while True:
        P.residual(x0)
        reps += 1
        if time.perf_counter() - t0 > args.cpu_seconds or reps >= args.cpu_max_reps:
            break
    t_res = (time.perf_counter() - t0) / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / (t_res * evals_needed), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{reps} FP64 residual evaluations of the {args.n}^3 melt by the oracle's MPS.cpp "
                      f"restatement (OpenMP, {cores} threads), {t_res * 1e3:.1f} ms each, scaled by the "
                      f"{evals_needed} residual evaluations of the plain DIIS chi sweep"}


def wall_adsorption_leg(args):
    """The 100-layer wall adsorption problem (A_100 at phi_b 0.01 in monomer
    solvent, chi_AS 0.5, chi_s -1, lambda 1/6, u = 0 start) solved to 1e-7 by
    the reference's default iteration (pseudo-Hessian Newton, its stop rule:
    accuracy <= 1e-7 and |g| <= 1e-6) and by FP64 DIIS (|g| < 1e-7): best of
    five wall-clock solves after one warm-up, residuals replayed from the
    system's captured graph."""
    from field_amd import configs, sfx
    lat, sysd, x0 = configs.adsorption_1d()
    G = sfx.System(lat, sysd)
    out = {"workload": "Lat1DFlat1stO 100 layers, surface wall chi_s=-1 / mirror1, A_100 phi_b=0.01 + S, "
                       "chi_AS=0.5, lambda=1/6, u=0, tolerance 1e-7"}
    for label, kw in (("newton", dict(method="newton", tolerance=1e-7, max_iter=1000)),
                      ("diis", dict(method="diis", tolerance=1e-7, max_iter=5000, diis_eta=0.1))):
        G.solve(x0, **kw)
        best, st = None, None
        for _ in range(5):
            t0 = time.perf_counter()
            _, st = G.solve(x0, **kw)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        out[label] = {"ms_per_solve": round(best * 1e3, 3), "iterations": st["iterations"],
                      "residual_evals": st["residual_evals"], "residual_norm": st["residual_norm"],
                      "graph_replays": st["graph_replays"]}
        if label == "newton":
            out[label]["accuracy"] = st["accuracy"]
    G.close()
    return out


def batch_leg(args):
    """The 512 uneven 1-D adsorption problems (layers 50..400, N 20..1000) in
    one persistent launch, DIIS with per-problem backtracking to |f| < 1e-7:
    time per batched solve (HIP events around the launch), convergence, and the
    algorithmic bytes the launch moves against the HBM roofline. Per residual a
    problem writes and reads back its N/2 stored ranks (16 B per site-rank) and
    reads x and writes f (32 B per layer); the rank loop itself runs in LDS."""
    from field_amd import configs, sfx
    probs = configs.batch_1d(512)
    B = sfx.Batch(probs)
    B.solve(tolerance=1e-7, max_iter=args.batch_max_iter, diis_backtrack=1, diis_eta=1.0)  # warm-up
    x, its, norms, st = B.solve(tolerance=1e-7, max_iter=args.batch_max_iter, diis_backtrack=1, diis_eta=1.0)
    evals = its.astype(float) + 1
    per_eval = [(p["N"] // 2) * p["n_layers"] * 16.0 + 32.0 * p["n_layers"] for p in probs]
    nbytes = float(np.dot(evals, per_eval))
    conv = norms <= 1e-7
    stable = np.array([configs.bulk_is_stable(p) for p in probs])
    threads, tuned = B.tuning()
    return {"workload": "512 x Lat1DFlat1stO adsorption, layers U{50..400}, N U{20..1000}, chi_AS U[0,1], "
                        "chi_s U[-2,0], phi_b 10^U[-4,-1] (seeds 4..515), FP64, DIIS m=10 eta=1 with backtracking",
            "seconds_per_batched_solve": round(st["seconds"], 4), "converged": int(conv.sum()),
            "stable_bulk": int(stable.sum()), "stable_bulk_converged": int(conv[stable].sum()),
            "max_iterations": int(its.max()), "residual_evaluations": int(evals.sum()),
            "threads_per_problem": threads, "tuned_residual_ms": [round(t, 3) for t in tuned],
            "algorithmic_bytes": nbytes,
            "achieved_GBs": round(nbytes / st["seconds"] / 1e9, 1),
            "frac_of_hbm_peak": round(nbytes / st["seconds"] / 1e9 / PEAK_HBM_GBS, 4),
            "bound": "latency: a chain of N barriered rank steps per residual in each workgroup"}


def transfer_compare(args):
    """The 128^3 semiflexible system (A_200, restricted phi = 0.3, chi = 0.6,
    stiffness 1 kT): FP32 residual + JVP with the second-order direction
    transfer as structured rows and on the matrix cores; per-kernel average
    launch times (HIP events) and wall time per residual + JVP."""
    import numpy as np
    from field_amd import configs, sfx
    lat, sysd, x = configs.semiflexible_3d()
    out = {}
    for tk in ("rows", "mfma", "rows", "mfma"):
        G = sfx.System(lat, sysd, transfer_kernel=tk)
        v = np.random.default_rng(1).normal(0, 1, G.nvar)
        G.residual(x, precision="fp32")
        G.jvp(x, v, precision="fp32")
        G.profile_start()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            G.residual(x, precision="fp32")
            G.jvp(x, v, precision="fp32")
        t1 = time.perf_counter()
        prof = G.profile_stop()
        rec = {f"{p['kernel']} {p['precision']} avg_us": round(p["avg_us"], 2) for p in prof}
        rec["ms_per_residual_plus_jvp"] = round((t1 - t0) * 1e3 / args.steps, 2)
        out.setdefault(tk, []).append(rec)
        G.close()
    print(json.dumps({"workload": "3-D SC 128^3 second order, A_200 restricted phi=0.3, chi=0.6, stiffness 1 kT; "
                                  "FP32 residual + FP32 JVP", "transfer": out}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--nk-max-iter", type=int, default=100)
    ap.add_argument("--gmres-max-iter", type=int, default=60)
    ap.add_argument("--gmres-rtol", type=float, default=0.3)
    ap.add_argument("--ptc-dt0", type=float, default=3.0)
    ap.add_argument("--later-ptc-dt0", type=float, default=0.0,
                    help="pseudo-transient start of the warm-started sweep steps (0: Newton with line search)")
    ap.add_argument("--fp64-switch", type=float, default=1e-3)
    ap.add_argument("--diis-max-iter", type=int, default=20000)
    ap.add_argument("--baseline-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-max-reps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--chi-path", default="0.38,0.4")
    ap.add_argument("--direct", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 512-problem and the 100-layer legs")
    ap.add_argument("--batch-max-iter", type=int, default=5000)
    ap.add_argument("--transfer-compare", action="store_true",
                    help="only time the second-order transfer as rows and on the matrix cores (128^3), then exit")
    ap.add_argument("--wall-only", action="store_true",
                    help="only solve the 100-layer wall adsorption problem (SFNewton and DIIS), then exit")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one untimed FP32 elementwise copy of this many MiB first (a known-traffic "
                         "kernel that calibrates hardware counters read in the same run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the solve is one lattice "
              f"on one device; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.transfer_compare:
        transfer_compare(args)
        return
    if args.wall_only:
        print(json.dumps({"wall_adsorption": wall_adsorption_leg(args)}))
        return

    import numpy as np
    from field_amd import configs, sfx

    lat, sysd, x0 = configs.melt_3d(n=args.n)
    G = sfx.System(lat, sysd)
    import torch
    xdev = torch.tensor(x0, dtype=torch.float64, device="cuda")  # resident before timing
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)  # reads and writes copy_mib MiB
        del src, dst
    torch.cuda.synchronize()

    path = [float(v) for v in args.chi_path.split(",")]
    if path[-1] != 0.4:
        sys.exit("--chi-path must end at the workload's chi = 0.4")
    opts = nk_options(args)
    cont = {"ptc_dt0": args.later_ptc_dt0}
    for _ in range(args.warmup):
        sweep(G, xdev, path, opts, cont)
    torch.cuda.synchronize()
    # one untimed sweep with every propagator launch bracketed by HIP events
    # gives the per-kernel table and names the dominant kernel; the first
    # timed sweep brackets every launch of that kernel alone, the others run
    # without events (each event pair costs queue time the solve would not
    # otherwise spend; bracketing only some launches of a sweep stamps the
    # bracketed ones early)
    G.profile_start()
    sweep(G, xdev, path, opts, cont)
    torch.cuda.synchronize()
    prof_all = G.profile_stop()
    dom_all = dominant(prof_all)
    G.profile_start(only=[(dom_all[0]["kind"], dom_all[0]["precision"])] if dom_all else [])
    stats = []
    prof = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        xs, st = sweep(G, xdev, path, opts, cont)
        stats.append(st)
        if prof is None:
            prof = G.profile_stop()  # waits for the sweep's last events
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if prof is None:
        prof = G.profile_stop()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    converged = all(s["converged"] for s in stats)
    value = 1e3 / ms_per_step if converged else 0.0

    # plain FP64 GPU form on the same input (--baseline-steps 0 skips it)
    dopts = diis_options(args)
    bstats = []
    xb = None
    tb0 = time.perf_counter()
    for _ in range(args.baseline_steps):
        xb, sb = sweep(G, xdev, path, dopts)
        bstats.append(sb)
    torch.cuda.synchronize()
    tb1 = time.perf_counter()
    b_ms = (tb1 - tb0) * 1e3 / max(1, args.baseline_steps)
    b_conv = bool(bstats) and all(s["converged"] for s in bstats)

    direct = None
    if args.direct:  # both forms at chi = 0.4 from the seeded field, no sweep
        direct = {}
        for name, o in (("nk", opts), ("plain_fp64_diis", dopts)):
            G.set_chi(0, 1, 0.4)
            td = time.perf_counter()
            xd, sd = G.solve(xdev, check_converged=False, **o)
            torch.cuda.synchronize()
            direct[name] = {"seconds": round(time.perf_counter() - td, 3), "converged": sd["converged"],
                            "iterations": sd["iterations"], "residual_norm": sd["residual_norm"]}
    G.set_chi(0, 1, 0.4)

    # solutions: what each form reached
    m = lat.interior_mask()
    xs_h = xs.cpu().numpy()
    G.residual(xs_h)
    F_nk = G.free_energy()
    phiA = G.phi(0)[m]
    plain = None
    if bstats:
        G.residual(xb.cpu().numpy())
        phiA_b = G.phi(0)[m]
        plain = {"ms_per_solve": round(b_ms, 3), "solves_per_s": round(1e3 / b_ms, 4) if b_conv else 0.0,
                 "converged": b_conv, "iterations": bstats[0]["iterations"], "per_chi": bstats[0]["per_chi"],
                 "residual_norm": bstats[0]["residual_norm"], "free_energy": G.free_energy(),
                 "phiA_min": float(phiA_b.min()), "phiA_max": float(phiA_b.max())}

    parity = None
    if not args.no_parity:
        from oracle import orc
        O = orc.System(lat, sysd)
        fo = O.residual(xs_h)
        Fo = O.free_energy()
        parity = {"oracle_residual_norm_at_gpu_solution": float(np.linalg.norm(fo)),
                  "free_energy_gpu": F_nk, "free_energy_oracle": Fo,
                  "free_energy_rel_diff": abs(F_nk - Fo) / abs(Fo)}

    dom = dominant(prof)
    roofline = None
    if dom:
        d, achieved = dom
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(d, args.n),
                    "kernel": f"{d['kernel']} ({d['precision']})", "avg_us": round(d["avg_us"], 3),
                    "launches": d["launches"], "ranks_per_launch": round(d["ranks"] / d["launches"], 3),
                    "bytes_per_launch": d["bytes_per_launch"]}

    batch = None if args.no_batch else batch_leg(args)
    wall = None if args.no_batch else wall_adsorption_leg(args)

    cpu = None
    if not args.no_cpu and bstats:
        evals = bstats[0]["residual_evals"]
        cpu = cpu_baseline(args, evals)

    s0 = stats[-1]
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"3-D SC {args.n}^3 periodic A12B28 melt (N=40, f_A=0.3), chi=0.4, MPS.cpp form, "
                        f"u~N(0,0.1) seed 3; chi sweep {path} (Vary-style, each step solved to |f|<1e-7); "
                        f"Newton-Krylov GMRES(30) with the Fourier-space preconditioner (forcing "
                        f"{args.gmres_rtol}), pseudo-transient continuation from dt={args.ptc_dt0} on the "
                        f"first step and dt={args.later_ptc_dt0} on the warm-started ones, FP32 tangent "
                        f"JVPs, FP64 residual",
            "chi_path": path,
            "lattice": f"{args.n}^3",
            "chain": "A12B28",
        },
        "solve": {"converged": converged, "newton_iterations": s0["iterations"], "per_chi": s0["per_chi"],
                  "jvp_evals": s0["jvp_evals"], "residual_evals_fp64": s0["residual_evals"],
                  "residual_evals_fp32": s0["residual_evals_fp32"], "residual_norm": s0["residual_norm"],
                  "free_energy": F_nk, "phiA_min": float(phiA.min()), "phiA_max": float(phiA.max())},
        "plain_fp64_diis": plain,
        "speedup_vs_plain_fp64": round(b_ms / ms_per_step, 3) if (b_conv and converged) else None,
        "roofline": roofline,
        "kernels": prof_all,
        "parity": parity,
        "cpu_baseline": cpu,
        "direct_at_chi_0.4": direct,
        "batched_adsorption": batch,
        "wall_adsorption": wall,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
