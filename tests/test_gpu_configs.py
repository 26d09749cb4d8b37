"""The five workloads at their full sizes, on the GPU through the C ABI.

Where the oracle finishes in seconds (the 64^3 melt, the 1-D wall, the 2-D
cylindrical pore) it checks the GPU directly at full size; for the 128^3 second-order system it
would take minutes, so that one is held to properties that do not depend on
size: FP32 against FP64, sqrt(N) checkpointing against the full store, and
the Jacobian product against central differences of the GPU's FP64
residual."""
import numpy as np
import pytest

from field_amd import configs

pytestmark = pytest.mark.gpu


def test_melt64_against_oracle_at_the_seeded_field(sfx, orc):
    """The 64^3 A12B28 melt (chi = 0.4, u ~ N(0, 0.1), seed 3) at its seeded
    field, where the profile is far from homogeneous: FP64 residual, phi per
    state, ln G_N and free energy against the oracle to 1e-10, and J v
    against the oracle's own central differences (h = 1e-6, FP64) on 16
    random directions; the FP32 path's errors against the FP64 oracle are
    reported."""
    lat, sysd, x = configs.melt_3d()
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    print("64^3 melt: FP64 residual, GPU and oracle", flush=True)
    fg = G.residual(x, precision="fp64")
    fo = O.residual(x)
    np.testing.assert_allclose(fg, fo, rtol=1e-10, atol=1e-12)
    assert np.abs(fo).max() > 1e-2  # far from a solution: the field is inhomogeneous
    for t in range(len(sysd.states)):
        np.testing.assert_allclose(G.phi(t), O.phi(t), rtol=1e-10, atol=1e-13)
    assert abs(G.lngn(0) - O.lngn(0)) <= 1e-10 * abs(O.lngn(0))
    Fg, Fo = G.free_energy(), O.free_energy()
    assert abs(Fg - Fo) <= 1e-10 * abs(Fo), (Fg, Fo)
    f32 = G.residual(x, precision="fp32")
    F32 = G.free_energy()
    print(f"FP32 path vs FP64 oracle: residual max abs err {np.abs(f32 - fo).max():.2e}, "
          f"free energy rel err {abs(F32 - Fo) / abs(Fo):.2e}", flush=True)
    assert np.abs(f32 - fo).max() < 2e-4
    mask = np.tile(lat.interior_mask().reshape(-1), 2)
    rng = np.random.default_rng(33)
    h = 1e-6
    worst, worst32 = 0.0, 0.0
    for k in range(16):
        v = rng.normal(0, 1, G.nvar) * mask
        fd = (O.residual(x + h * v) - O.residual(x - h * v)) / (2 * h)
        jv = G.jvp(x, v, precision="fp64")
        j32 = G.jvp(x, v, precision="fp32")
        scale = np.abs(fd).max()
        worst = max(worst, np.abs(jv - fd).max() / scale)
        worst32 = max(worst32, np.abs(j32 - fd).max() / scale)
        print(f"direction {k}: FP64 J v vs oracle central differences {np.abs(jv - fd).max() / scale:.2e}",
              flush=True)
    print(f"64^3 melt: worst FP64 {worst:.2e}, FP32 {worst32:.2e}")
    assert worst < 1e-6
    assert worst32 < 1e-3


def test_wall_adsorption_full_size(sfx, orc):
    """100 layers, A_100 at phi_b = 0.01 next to a wall with chi_s = -1, u = 0
    start, solved to 1e-7: residual parity at the start and at the solution,
    free energy to 1e-8."""
    lat, sysd, x0 = configs.adsorption_1d()
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    x = np.random.default_rng(1).normal(0, 0.05, G.nvar)
    np.testing.assert_allclose(G.residual(x, precision="fp64"), O.residual(x), rtol=1e-10, atol=1e-12)
    xs, st = G.solve(x0, method="nk", tolerance=1e-7, max_iter=60, gmres_rtol=1e-4, ptc_dt0=0.0)
    print("wall adsorption:", st)
    assert st["converged"]
    assert np.linalg.norm(O.residual(xs)) < 1e-7
    G.residual(xs)
    assert abs(G.free_energy() - O.free_energy()) <= 1e-8 * abs(O.free_energy())


def test_cylindrical_pore_full_size(sfx, orc):
    """64 x 128 cylindrical pore, A20B20 at phi_b = 0.05 with an attractive
    wall: residual parity (FP64 1e-10, FP32 reported and bounded) and JVP
    against the oracle's central differences at the seeded start."""
    lat, sysd, x = configs.cylinder_2d()
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    fo = O.residual(x)
    np.testing.assert_allclose(G.residual(x, precision="fp64"), fo, rtol=1e-10, atol=1e-12)
    e32 = np.abs(G.residual(x, precision="fp32") - fo).max()
    print(f"pore: FP32 residual error {e32:.2e}")
    assert e32 < 2e-4
    rng = np.random.default_rng(3)
    for _ in range(2):
        v = rng.normal(0, 1, G.nvar)
        h = 1e-6
        fd = (O.residual(x + h * v) - O.residual(x - h * v)) / (2 * h)
        jv = G.jvp(x, v, precision="fp64")
        assert np.abs(jv - fd).max() / np.abs(fd).max() < 1e-6
        j32 = G.jvp(x, v, precision="fp32")
        assert np.abs(j32 - jv).max() / np.abs(jv).max() < 1e-3


def test_cylindrical_pore_fp32_solve_matches_fp64(sfx, orc):
    """The same pore solved with FP32 Jacobian products and with FP64 ones
    reaches the same solution: the oracle's residual there is below the
    tolerance and the free energies agree to 1e-8."""
    lat, sysd, x = configs.cylinder_2d()
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    F = {}
    for prec in ("fp32", "fp64"):
        xs, st = G.solve(x, method="nk", tolerance=1e-8, max_iter=80, gmres_rtol=1e-3, jvp_precision=prec)
        print(f"pore NK jvp {prec}:", st)
        assert st["converged"]
        assert np.linalg.norm(O.residual(xs)) < 1e-8
        G.residual(xs)
        F[prec] = G.free_energy()
    assert abs(F["fp32"] - F["fp64"]) <= 1e-8 * abs(F["fp64"])


def test_semiflexible_full_size_properties(sfx):
    """128^3 periodic box, second-order A_200 at phi = 0.3 in solvent:
    FP32 propagators against FP64, a checkpointed store against the full one
    (bit for bit), and J v against central differences of the FP64 residual."""
    import torch
    lat, sysd, x = configs.semiflexible_3d()
    G = sfx.System(lat, sysd)
    xd = torch.tensor(x, device="cuda")
    f64 = G.residual(xd, precision="fp64")
    f32 = G.residual(xd, precision="fp32")
    e32 = (f32 - f64).abs().max().item()
    print(f"semiflexible 128^3: FP32 residual error {e32:.2e}")
    assert e32 < 2e-4
    full = 200 * 6 * G.M * 8  # bytes of the full FP64 store
    C = sfx.System(lat, sysd, memory_budget=full // 4)  # sqrt(N) checkpoints need 2 sqrt(N) of 200 ranks
    assert torch.equal(C.residual(xd, precision="fp32"), f32)
    v = torch.tensor(np.random.default_rng(4).normal(0, 1, G.nvar), device="cuda")
    h = 1e-6
    fd = (G.residual(xd + h * v, precision="fp64") - G.residual(xd - h * v, precision="fp64")) / (2 * h)
    jv = G.jvp(xd, v, precision="fp64")
    rel = ((jv - fd).abs().max() / fd.abs().max()).item()
    print(f"semiflexible 128^3: JVP vs central differences {rel:.2e}")
    assert rel < 1e-5
    j32 = G.jvp(xd, v, precision="fp32")
    assert ((j32 - jv).abs().max() / jv.abs().max()).item() < 1e-3


def test_semiflexible_solve_is_homogeneous(sfx):
    """The 128^3 semiflexible system solved by Newton-Krylov with sqrt(N)
    checkpoints forced by a memory budget, to 1e-7 in the FP64 residual.
    chi = 0.6 lies below the spinodal (1/(N phi) + 1/(1 - phi) = 2 chi gives
    chi_s = 0.72 at phi = 0.3, N = 200), so the solution is the homogeneous
    phi_A = 0.3 at every site: a closed form at full size."""
    import torch
    lat, sysd, x0 = configs.semiflexible_3d()
    full = 200 * 6 * lat.M * 8
    G = sfx.System(lat, sysd, memory_budget=full // 4)
    xs, st = G.solve(torch.tensor(x0, device="cuda"), method="nk", tolerance=1e-7, max_iter=60, gmres_rtol=0.3,
                     jvp_precision="fp32", ptc_dt0=3.0)
    print("semiflexible 128^3 solve:", st)
    assert st["converged"]
    G.residual(xs)
    inner = lat.interior_mask().reshape(-1)
    phi_a = G.phi(0)[inner]
    assert np.abs(phi_a - 0.3).max() < 1e-6
    assert np.abs(G.phi(1)[inner] - 0.7).max() < 1e-6
