"""GPU (HIP, gfx950) against the CPU oracle on the same inputs, through the C ABI.

Tolerances: FP64 path 1e-10 relative (absolute floor 1e-12 for entries near
zero); FP32 propagators with per-rank log-scaling are reported against the
FP64 oracle and bounded at 2e-4 absolute on residual entries."""
import math

import numpy as np
import pytest

import cases
from field_amd import configs

pytestmark = pytest.mark.gpu

RTOL = 1e-10
ATOL = 1e-12


def _close(a, b, rtol=RTOL, atol=ATOL):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol)


@pytest.fixture(scope="module")
def built(sfx, orc):
    out = {}
    for name in cases.NAMES:
        lat, sysd = cases.case(name)
        out[name] = (lat, sysd, sfx.System(lat, sysd), orc.System(lat, sysd))
    return out


@pytest.mark.parametrize("name", cases.NAMES)
def test_residual_phi_lngn_fp64(built, name):
    lat, sysd, G, O = built[name]
    print(f"case {name} nvar={G.nvar}", flush=True)
    x = cases.start(lat, sysd)
    fg = G.residual(x, precision="fp64")
    fo = O.residual(x)
    _close(fg, fo)
    for t in range(len(sysd.states)):
        _close(G.phi(t), O.phi(t))
    for m in range(len(sysd.molecules)):
        assert abs(G.lngn(m) - O.lngn(m)) <= RTOL * max(1.0, abs(O.lngn(m)))
        assert abs(G.phibulk(m) - O.phibulk(m)) <= RTOL * max(1e-3, abs(O.phibulk(m)))


@pytest.mark.parametrize("name", cases.NAMES)
def test_free_energy_fp64(built, name):
    lat, sysd, G, O = built[name]
    x = cases.start(lat, sysd, seed=12)
    G.residual(x)
    O.residual(x)
    Fg, Fo = G.free_energy(), O.free_energy()
    assert abs(Fg - Fo) <= 1e-10 * max(1.0, abs(Fo)), (Fg, Fo)


@pytest.mark.parametrize("name", cases.NAMES)
def test_fp32_logscaled_against_fp64_oracle(built, name):
    lat, sysd, G, O = built[name]
    x = cases.start(lat, sysd, seed=13)
    f32 = G.residual(x, precision="fp32")
    fo = O.residual(x)
    err = np.abs(f32 - fo).max()
    print(f"{name}: FP32 propagators, max |f32 - f64| = {err:.3e}")
    assert err < 2e-4
    lg = G.lngn(0)
    assert abs(lg - O.lngn(0)) < 1e-4 * max(1, abs(O.lngn(0)))


def test_fp32_scaling_survives_underflow(sfx, orc):
    """A long free chain on a repulsive field: G_N ~ e^{-N u} drops below the
    FP32 range (1e-38) without the per-rank log-scale; with it, ln G_N and the
    residual stay accurate."""
    from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec
    lat = LatticeSpec(1, (40,), lowerbound=("mirror1",), upperbound=("mirror1",), lambda_=1 / 6)
    sysd = SystemSpec(("A", "S"), np.array([[0, 0.5], [0.5, 0]]),
                      [MoleculeSpec("p", "(A)300", "free", phibulk=0.01), MoleculeSpec("s", "(S)1", "solvent")])
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    rng = np.random.default_rng(3)
    x = rng.normal(0, 0.1, G.nvar)
    x[0::2] += 0.6  # u_A (iteration variables alternate A, S per site)
    f32 = G.residual(x, precision="fp32")
    fo = O.residual(x)
    assert O.lngn(0) < math.log(1e-38)  # the unscaled FP32 propagator would underflow
    assert abs(G.lngn(0) - O.lngn(0)) < 1e-5 * abs(O.lngn(0))
    err = np.abs(f32 - fo).max()
    print(f"ln G_N = {O.lngn(0):.2f}; FP32 residual error {err:.2e}")
    assert err < 1e-4


def test_negative_solvent_bulk_fraction(sfx, orc):
    """A restricted chain whose theta / G_N exceeds 1 leaves the solvent a
    negative bulk fraction (the closure 1 - sum phi_b); the reference
    multiplies by it as it is, and so must both precisions."""
    from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec
    lat = LatticeSpec(1, (40,), lowerbound=("mirror1",), upperbound=("mirror1",), lambda_=1 / 6)
    sysd = SystemSpec(("A", "S"), np.array([[0, 0.5], [0.5, 0]]),
                      [MoleculeSpec("p", "(A)20", "restricted", theta=48.0), MoleculeSpec("s", "(S)1", "solvent")])
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    x = np.random.default_rng(4).normal(0, 0.02, G.nvar)
    fo = O.residual(x)
    assert O.phibulk(1) < 0
    f64 = G.residual(x, precision="fp64")
    assert abs(G.phibulk(1) - O.phibulk(1)) < 1e-12
    np.testing.assert_allclose(f64, fo, rtol=0, atol=1e-10 * max(1, np.abs(fo).max()))
    f32 = G.residual(x, precision="fp32")
    assert np.abs(f32 - fo).max() < 2e-4
    v = np.random.default_rng(5).normal(0, 1, G.nvar)
    h = 1e-6
    fd = (O.residual(x + h * v) - O.residual(x - h * v)) / (2 * h)
    jv = G.jvp(x, v, precision="fp64")
    assert np.abs(jv - fd).max() / np.abs(fd).max() < 1e-6


@pytest.mark.parametrize("name", ["1d_flat_adsorption", "2d_cylinder", "3d_sc_melt_n24", "1d_sphere",
                                  "3d_sc_mirrors_blocked", "1d_order2_adsorption",
                                  "3d_sc_order2_homopolymer", "3d_sc_order2_copolymer", "2d_cyl_order2"])
def test_checkpointed_sweeps_match_full_store(built, name, sfx):
    """A memory budget below the full propagator store switches to sqrt(N)
    checkpoints with recomputation; residuals and JVPs are bit-identical to the
    full store swept in the same order (forward, then backward) in both
    precisions."""
    lat, sysd, _, O = built[name]
    G = sfx.System(lat, sysd, chain_schedule="sequential")
    comps = 6 if lat.order == 2 else 1
    full = cases.store_ranks(lat, sysd) * comps * G.M * 8
    C = sfx.System(lat, sysd, memory_budget=full - 1)
    x = cases.start(lat, sysd, seed=21, sigma=0.1)
    v = np.random.default_rng(22).normal(0, 1, G.nvar)
    for prec in ("fp64", "fp32"):
        np.testing.assert_array_equal(C.residual(x, precision=prec), G.residual(x, precision=prec))
        np.testing.assert_array_equal(C.jvp(x, v, precision=prec), G.jvp(x, v, precision=prec))
    with pytest.raises(sfx.SFError):
        sfx.System(lat, sysd, memory_budget=2 * comps * G.M * 8)


@pytest.mark.parametrize("name", ["3d_sc_melt_n24", "3d_sc_mirrors_blocked"])
def test_middle_out_matches_sequential(built, name, sfx):
    """The middle-out schedule (forward and backward chains in the same
    launches, each composing half the ranks) against the sequential one:
    residuals and JVPs agree to rounding in both precisions, and the FP64
    residual matches the oracle."""
    lat, sysd, G, O = built[name]
    Q = sfx.System(lat, sysd, chain_schedule="sequential")
    x = cases.start(lat, sysd, seed=23, sigma=0.1)
    v = np.random.default_rng(24).normal(0, 1, G.nvar)
    f = G.residual(x, precision="fp64")
    np.testing.assert_allclose(f, O.residual(x), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(f, Q.residual(x, precision="fp64"), rtol=1e-12, atol=1e-13)
    j = G.jvp(x, v, precision="fp64")
    np.testing.assert_allclose(j, Q.jvp(x, v, precision="fp64"), rtol=1e-11, atol=1e-12 * np.abs(j).max())
    f32, q32 = G.residual(x, precision="fp32"), Q.residual(x, precision="fp32")
    assert np.abs(f32 - q32).max() < 1e-5 * max(1.0, np.abs(f).max())
    j32, k32 = G.jvp(x, v, precision="fp32"), Q.jvp(x, v, precision="fp32")
    assert np.abs(j32 - k32).max() < 1e-4 * np.abs(j).max()


@pytest.mark.parametrize("name", cases.NAMES)
def test_jvp_matches_central_differences(built, name):
    """J v by tangent propagators against (f(x+hv) - f(x-hv)) / 2h, h = 1e-6, FP64."""
    lat, sysd, G, O = built[name]
    rng = np.random.default_rng(21)
    x = cases.start(lat, sysd, seed=14)
    for k in range(2):
        v = rng.normal(0, 1, G.nvar)
        jv = G.jvp(x, v, precision="fp64")
        h = 1e-6
        fd = (O.residual(x + h * v) - O.residual(x - h * v)) / (2 * h)
        err = np.abs(jv - fd).max() / max(1e-12, np.abs(fd).max())
        print(f"{name}: JVP vs central FD rel err {err:.2e}")
        assert err < 1e-6
        jv32 = G.jvp(x, v, precision="fp32")
        err32 = np.abs(jv32 - jv).max() / np.abs(jv).max()
        print(f"{name}: FP32 JVP vs FP64 JVP rel err {err32:.2e}")
        assert err32 < 1e-3


def test_diis_follows_the_oracle_trajectory(sfx, orc):
    from field_amd import configs
    lat, sysd, x = configs.melt_3d(n=6, n_a=3, n_b=5, chi=0.4, sigma=0.1, seed=7)
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    xg, st = G.solve(x, method="diis", tolerance=1e-8)
    xo, it, err, rc = O.solve_diis(x, tol=1e-8)
    assert rc == 0 and st["converged"]
    assert abs(st["iterations"] - it) <= 2, (st["iterations"], it)
    fo = O.residual(xg)
    assert np.linalg.norm(fo) < 1.5e-8


@pytest.mark.parametrize("name", ["1d_flat_adsorption", "2d_flat", "3d_sc_melt", "3d_hex_three_molecules"])
def test_newton_krylov_converges(built, name, orc):
    lat, sysd, G, O = built[name]
    x = cases.start(lat, sysd, seed=15, sigma=0.05)
    xs, st = G.solve(x, method="nk", tolerance=1e-9, max_iter=60, gmres_rtol=1e-4)
    print(f"{name}: NK {st}")
    assert st["converged"]
    f = O.residual(xs)  # the oracle's FP64 residual at the GPU solution
    assert np.linalg.norm(f) < 1e-8
    G.residual(xs)
    assert abs(G.free_energy() - O.free_energy()) <= 1e-10 * max(1, abs(O.free_energy()))


def test_lattice_primitives_match_oracle(sfx, orc):
    """Cuda_tools-style per-rank calls on device pointers."""
    import ctypes as C
    import torch
    from field_amd.spec import LatticeSpec
    lat = LatticeSpec(3, (5, 6, 7), lattice_type="fcc", lowerbound=("periodic", "mirror1", "mirror2"),
                      upperbound=("periodic", "mirror2", "mirror1"))
    h = C.c_void_p()
    d = lat.desc()
    sfx.check(sfx.lib().sfx_lattice_create(C.byref(d), C.byref(h)))
    rng = np.random.default_rng(5)
    G = rng.random(lat.M)
    g = rng.random(lat.M)
    OL = orc.Lattice(lat)
    Gs = OL.set_boundaries(G)
    ref = OL.propagate(Gs, OL.set_boundaries(g))
    tG = torch.tensor(Gs, device="cuda")
    tg = torch.tensor(OL.set_boundaries(g), device="cuda")
    tout = torch.zeros(lat.M, dtype=torch.float64, device="cuda")
    sfx.check(sfx.lib().sfx_lattice_propagate(h, C.c_void_p(tG.data_ptr()), C.c_void_p(tg.data_ptr()),
                                              C.c_void_p(tout.data_ptr()), None))
    torch.cuda.synchronize()
    _close(tout.cpu().numpy(), ref)
    ln = C.c_double()
    sfx.check(sfx.lib().sfx_lattice_lngn(h, C.c_void_p(tout.data_ptr()), C.byref(ln), None))
    assert abs(ln.value - OL.lngn(ref)) < 1e-12
    side = torch.zeros_like(tout)
    sfx.check(sfx.lib().sfx_lattice_sides(h, C.c_void_p(tG.data_ptr()), C.c_void_p(side.data_ptr()), None))
    torch.cuda.synchronize()
    m = lat.interior_mask()
    _close(side.cpu().numpy()[m], OL.sides(Gs)[m])
    sfx.lib().sfx_lattice_destroy(h)


def test_set_chi_equals_fresh_system(sfx, orc):
    """sfx_set_chi: residual, JVP and preconditioner after a chi change are
    those of a system created at the new chi; a change that would merge
    iteration variables is refused."""
    lat, sys_a, _ = configs.melt_3d(n=7, n_a=3, n_b=5, chi=0.3)
    _, sys_b, _ = configs.melt_3d(n=7, n_a=3, n_b=5, chi=0.45)
    Ga, Gb = sfx.System(lat, sys_a), sfx.System(lat, sys_b)
    mask = np.tile(lat.interior_mask().reshape(-1), 2)
    rng = np.random.default_rng(8)
    x = rng.normal(0, 0.2, Ga.nvar) * mask
    v = rng.normal(0, 1, Ga.nvar) * mask
    Ga.precondition(v)  # build the preconditioner at chi = 0.3 first
    Ga.set_chi(0, 1, 0.45)
    fa, fb = Ga.residual(x), Gb.residual(x)
    assert np.array_equal(fa, fb)
    assert np.abs(fa - orc.System(lat, sys_b).residual(x)).max() < 1e-12
    assert np.array_equal(Ga.jvp(x, v, precision="fp64"), Gb.jvp(x, v, precision="fp64"))
    assert np.array_equal(Ga.precondition(v), Gb.precondition(v))
    Ga.residual(x)
    Gb.residual(x)
    assert Ga.free_energy() == Gb.free_energy()
    with pytest.raises(sfx.SFError):
        Ga.set_chi(0, 1, 0.0)  # both chi rows would be zero: one iteration variable instead of two
    assert np.array_equal(Ga.residual(x), fb)  # a refused change leaves chi alone


def test_stream_set_gives_the_same_bits_on_a_caller_stream(sfx):
    """sfx_stream_set: on a torch stream the system reads inputs torch queued
    there and returns the residual, JVP and solve it returns on its own
    stream, bit for bit; None hands it back its own stream."""
    import torch

    lat, sysd, x0 = configs.melt_3d(n=16)
    G = sfx.System(lat, sysd, precision="fp32")
    mask = np.tile(lat.interior_mask().reshape(-1), 2)
    v = np.random.default_rng(11).normal(0, 1, G.nvar) * mask
    f_own, jv_own = G.residual(x0), G.jvp(x0, v)
    kw = dict(method="nk", tolerance=1e-8, max_iter=40, ptc_dt0=1.0)
    x_own, st_own = G.solve(x0, **kw)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        G.set_stream(s)
        xd = torch.from_numpy(x0).cuda() * 1.0  # queued on s
        vd = torch.from_numpy(v).cuda() * 1.0
        fd, jd = G.residual(xd), G.jvp(xd, vd)
        xs, st_s = G.solve(xd, **kw)
        s.synchronize()
    assert np.array_equal(fd.cpu().numpy(), f_own)
    assert np.array_equal(jd.cpu().numpy(), jv_own)
    assert np.array_equal(xs.cpu().numpy(), x_own) and st_s["iterations"] == st_own["iterations"]
    G.set_stream(None)
    assert np.array_equal(G.residual(x0), f_own)


def test_profile_select_times_only_the_chosen_kernel(sfx):
    """sfx_profile_start_select: the chosen (kind, precision) pair gets the
    launches and ranks the full profile counts for it, and nothing else is
    timed; results do not depend on profiling."""
    lat, sysd, x0 = configs.melt_3d(n=16)
    G = sfx.System(lat, sysd, precision="fp32")
    v = np.random.default_rng(12).normal(0, 1, G.nvar) * np.tile(lat.interior_mask().reshape(-1), 2)
    f0, j0 = G.residual(x0), G.jvp(x0, v)
    G.profile_start()
    f1, j1 = G.residual(x0), G.jvp(x0, v)
    full = {(e["kind"], e["precision"]): e for e in G.profile_stop()}
    assert (3, "fp32") in full and len(full) >= 3
    G.profile_start(only=[(3, "fp32")])
    f2, j2 = G.residual(x0), G.jvp(x0, v)
    sel = G.profile_stop()
    assert [(e["kind"], e["precision"]) for e in sel] == [(3, "fp32")]
    assert sel[0]["launches"] == full[(3, "fp32")]["launches"] and sel[0]["ranks"] == full[(3, "fp32")]["ranks"]
    assert sel[0]["total_ms"] > 0
    for a, b in ((f0, f1), (f0, f2), (j0, j1), (j0, j2)):
        assert np.array_equal(a, b)


def test_chi_sweep_reaches_melt_solution(sfx, orc):
    """A two-step `var` sweep of chi_AB (0.38, then 0.4) on the melt from its
    seeded start ends at a solution of the chi = 0.4 system (checked by the
    oracle's residual and free energy), with the system left at chi = 0.4."""
    lat, sysd, x0 = configs.melt_3d(n=16)
    G = sfx.System(lat, sysd)
    xs, sts = G.solve_chi_sweep(x0, 0, 1, [0.38, 0.4], method="nk", tolerance=1e-9, max_iter=60, ptc_dt0=1.0)
    print("sweep", sts)
    assert [s["chi"] for s in sts] == [0.38, 0.4] and all(s["converged"] for s in sts)
    O = orc.System(lat, sysd)
    assert np.linalg.norm(O.residual(xs)) < 1e-8
    G.residual(xs)
    assert abs(G.free_energy() - O.free_energy()) <= 1e-10 * abs(O.free_energy())


@pytest.mark.parametrize("which", ["diblock", "triblock_three_states"])
def test_fourier_preconditioner_inverts_homogeneous_jacobian(sfx, which):
    """At u = 0 the melt is homogeneous and its Jacobian is exactly the
    Fourier-space block operator the preconditioner inverts: P (J v) returns v
    up to the null direction (a uniform shift of every field)."""
    from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec
    if which == "diblock":
        # an odd side: even sides make J singular at the checkerboard q = (pi, pi, pi)
        # for an even chain, and the preconditioner leaves that direction alone
        lat, sysd, _ = configs.melt_3d(n=7, n_a=3, n_b=5, chi=0.4)
    else:
        lat = LatticeSpec(3, (6, 8, 5), lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
        chi = np.array([[0, 0.5, 1.0], [0.5, 0, 0.3], [1.0, 0.3, 0]])
        V = 6 * 8 * 5
        sysd = SystemSpec(("A", "B", "C"), chi, [MoleculeSpec("abc", "(A)2(B)4(C)3", "restricted", theta=0.7 * V),
                                                 MoleculeSpec("b", "(B)5", "restricted", theta=0.3 * V)],
                          form="melt", melt_phibulk=(0.0, 0.0, 0.0))
    G = sfx.System(lat, sysd)
    mask = np.tile(lat.interior_mask().reshape(-1), len(sysd.states))
    v = np.random.default_rng(3).normal(0, 1, G.nvar) * mask
    jv = G.jvp(np.zeros(G.nvar), v, precision="fp64")
    w = G.precondition(jv)
    expect = (v - v[mask].mean()) * mask
    err = np.abs(w - expect).max() / np.abs(expect).max()
    print(f"{which}: |P J v - v| / |v| = {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("name", ["3d_sc_mirrors_blocked", "3d_sc_melt_n24"])
def test_per_rank_and_blocked_sweeps_agree(sfx, orc, name):
    """sfx_options.sweep_kernel: the one-rank-per-launch and the blocked 3-D
    sweeps give the oracle's residual (FP64, 1e-10) and the same FP32 residual
    and JVP to FP32 rounding."""
    lat, sysd = cases.case(name)
    P = sfx.System(lat, sysd, sweep_kernel="per_rank")
    B = sfx.System(lat, sysd, sweep_kernel="blocked")
    x = cases.start(lat, sysd, seed=31)
    v = np.random.default_rng(32).normal(0, 1, P.nvar)
    fo = orc.System(lat, sysd).residual(x)
    for S in (P, B):
        _close(S.residual(x, precision="fp64"), fo)
    f32p, f32b = P.residual(x, precision="fp32"), B.residual(x, precision="fp32")
    assert np.abs(f32p - f32b).max() < 2e-5 * max(1.0, np.abs(fo).max())
    jp, jb = P.jvp(x, v, precision="fp32"), B.jvp(x, v, precision="fp32")
    assert np.abs(jp - jb).max() < 1e-4 * np.abs(jb).max()


@pytest.mark.parametrize("n, n_a, n_b, chi", [(16, 3, 4, 0.5), (64, 12, 28, 0.38)])
def test_fourier_preconditioner_matches_numpy(sfx, n, n_a, n_b, chi):
    """Power-of-two sides (the three-pass path: planes FFT'd in LDS, x-lines
    FFT'd with the block solve fused) against numpy: FFT of the residual,
    (A^T A + mu^2 I)^-1 A^T r per wave vector with A = J0(q) the lattice-RPA
    Jacobian (pinned against the oracle's finite differences in
    test_spinodal.py) and mu = 1e-6 max|A|, inverse FFT. FP32 transforms."""
    from test_spinodal import rpa_jacobian
    lat, sysd, _ = configs.melt_3d(n=n, n_a=n_a, n_b=n_b, chi=chi)
    G = sfx.System(lat, sysd)
    mask = lat.interior_mask().reshape(-1)
    r = np.random.default_rng(9).normal(0, 1, (2, G.M)) * mask
    w = G.precondition(r.reshape(-1)).reshape(2, G.M)
    inner = np.flatnonzero(mask)
    R = np.fft.fftn(r[:, inner].reshape(2, n, n, n), axes=(1, 2, 3))
    A = np.moveaxis(rpa_jacobian(n, n_a, n_b, chi), (0, 1), (-2, -1))  # (n, n, n, 2, 2)
    mu2 = (1e-6 * np.abs(A).max(axis=(-2, -1))) ** 2
    At = np.swapaxes(A, -1, -2)
    B = At @ A + mu2[..., None, None] * np.eye(2)
    rhs = (At @ np.moveaxis(R, 0, -1)[..., None])[..., 0]
    Wq = np.linalg.solve(B, rhs[..., None])[..., 0]
    wn = np.fft.ifftn(np.moveaxis(Wq, -1, 0), axes=(1, 2, 3)).real.reshape(2, -1)
    err = np.abs(w[:, inner] - wn).max() / np.abs(wn).max()
    print(f"{n}^3 preconditioner vs numpy: {err:.2e}")
    assert err < 1e-4
    assert np.all(w[:, ~mask] == 0)


@pytest.mark.parametrize("name", cases.NAMES)
def test_grand_and_chemical_potentials(built, name, sfx):
    """sfx_grand_potential and sfx_chemical_potential against the oracle at a
    seeded field (FP64, 1e-10 relative); the melt form refuses both."""
    lat, sysd, G, O = built[name]
    x = cases.start(lat, sysd, seed=41, sigma=0.1)
    G.residual(x, precision="fp64")
    if sysd.form == "melt":
        with pytest.raises(sfx.SFError):
            G.grand_potential()
        with pytest.raises(sfx.SFError):
            G.chemical_potential(0)
        return
    O.residual(x)
    og, oo = G.grand_potential(), O.grand_potential()
    print(f"{name}: grand potential GPU {og:.12e} oracle {oo:.12e}")
    assert abs(og - oo) <= 1e-10 * max(1.0, abs(oo))
    for m in range(len(sysd.molecules)):
        mg, mo = G.chemical_potential(m), O.chemical_potential(m)
        assert abs(mg - mo) <= 1e-10 * max(1.0, abs(mo)), (m, mg, mo)


@pytest.mark.parametrize("name", ["3d_sc_order2_homopolymer", "3d_sc_order2_copolymer", "1d_order2_adsorption",
                                  "1d_cyl_order2", "2d_cyl_order2"])
def test_second_order_transfer_on_matrix_cores(built, name, sfx):
    """The 6x6 direction transfer on the matrix cores (v_mfma_f32_4x4x1f32)
    against the structured rows: FP32 residuals and JVPs agree to FP32
    rounding, and both stay within the FP32 bound of the FP64 oracle."""
    lat, sysd, G, O = built[name]
    Gm = sfx.System(lat, sysd, transfer_kernel="mfma")
    Gr = sfx.System(lat, sysd, transfer_kernel="rows")
    x = cases.start(lat, sysd, seed=21)
    fo = O.residual(x)
    fm, fr = Gm.residual(x, precision="fp32"), Gr.residual(x, precision="fp32")
    print(f"{name}: |f_mfma - f_rows| = {np.abs(fm - fr).max():.2e}, |f_mfma - f_oracle| = {np.abs(fm - fo).max():.2e}")
    assert np.abs(fm - fr).max() < 1e-5
    assert np.abs(fm - fo).max() < 2e-4
    v = np.random.default_rng(3).normal(0, 1, G.nvar)
    jm, jr = Gm.jvp(x, v, precision="fp32"), Gr.jvp(x, v, precision="fp32")
    assert np.abs(jm - jr).max() <= 1e-4 * np.abs(jr).max()
    # FP64 sweeps keep the rows: the option changes nothing there
    np.testing.assert_array_equal(Gm.residual(x, precision="fp64"), Gr.residual(x, precision="fp64"))
