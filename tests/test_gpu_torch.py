"""The PyTorch wrapper (field_amd.torch_ops) on the GPU: forward-mode AD
through torch.autograd.Function gives the hand-written tangent sweep's J v,
checked against the reference gradient (torch autodiff of a plain-torch
restatement, tests/torch_ref.py) and against central differences."""
import numpy as np
import pytest

from field_amd import configs

pytestmark = pytest.mark.gpu


def test_forward_ad_matches_reference_gradient(sfx):
    import torch

    import torch_ref
    from field_amd import torch_ops

    n, na, nb, chi = 6, 3, 5, 0.4
    lat, sysd, x = configs.melt_3d(n=n, n_a=na, n_b=nb, chi=chi, sigma=0.2, seed=9)
    G = sfx.System(lat, sysd)
    mask = np.tile(lat.interior_mask().reshape(-1), 2)
    seq, chim, theta = [0] * na + [1] * nb, [[0, chi], [chi, 0]], float(n ** 3)
    u = torch.from_numpy(x).cuda()
    ui = torch_ref.to_interior(torch.from_numpy(x), n)
    ref = lambda w: torch_ref.melt_residual(w, seq, chim, theta, (0.0, 0.0))  # noqa: E731
    rng = np.random.default_rng(1)
    worst = 0.0
    for k in range(16):
        vh = rng.normal(0, 1, G.nvar) * mask
        f, jv = torch_ops.jvp(u, torch.from_numpy(vh).cuda(), G)
        fr, jr = torch.func.jvp(ref, (ui,), (torch_ref.to_interior(torch.from_numpy(vh), n),))
        np.testing.assert_allclose(f.cpu().numpy(), torch_ref.to_padded(fr, n).numpy(), rtol=0, atol=1e-11)
        jr = torch_ref.to_padded(jr, n).numpy()
        err = np.abs(jv.cpu().numpy() - jr).max() / np.abs(jr).max()
        worst = max(worst, err)
    print(f"forward-mode J v vs torch autodiff reference: max rel err {worst:.2e}")
    assert worst < 1e-10


def test_forward_ad_fp32_and_errors(sfx):
    import torch

    from field_amd import torch_ops

    lat, sysd, x = configs.melt_3d(n=6, n_a=3, n_b=5, chi=0.4, sigma=0.2, seed=9)
    G = sfx.System(lat, sysd)
    u = torch.from_numpy(x).cuda()
    v = torch.from_numpy(np.random.default_rng(2).normal(0, 1, G.nvar)
                         * np.tile(lat.interior_mask().reshape(-1), 2)).cuda()
    _, j64 = torch_ops.jvp(u, v, G, "fp64")
    _, j32 = torch_ops.jvp(u, v, G, "fp32")
    err = (j32 - j64).abs().max().item() / j64.abs().max().item()
    print(f"FP32 tangent sweep vs FP64: rel err {err:.2e}")
    assert err < 1e-3
    uu = u.clone().requires_grad_(True)
    torch_ops.residual(uu, G).sum().backward()  # reverse mode: J^T 1 (sfx_vjp)
    ref = torch.from_numpy(G.vjp(x, np.ones(G.nvar))).cuda()
    assert torch.allclose(uu.grad, ref, rtol=1e-12, atol=1e-14)
    with pytest.raises(TypeError):
        torch_ops.residual(u.cpu(), G)


def test_config_melt_jvp_against_central_differences(sfx):
    """64^3 AB diblock melt (N = 40): J v against (f(x+hv) - f(x-hv)) / 2h with
    h = 1e-6 in FP64 on 16 random directions; the FP32 tangent sweep's error
    against the FP64 one is reported."""
    import torch

    lat, sysd, x = configs.melt_3d()
    G = sfx.System(lat, sysd)
    mask = np.tile(lat.interior_mask().reshape(-1), 2)
    xd = torch.from_numpy(x).cuda()
    rng = np.random.default_rng(16)
    h = 1e-6
    worst, worst32 = 0.0, 0.0
    for k in range(16):
        v = torch.from_numpy(rng.normal(0, 1, G.nvar) * mask).cuda()
        fd = (G.residual(xd + h * v, precision="fp64") - G.residual(xd - h * v, precision="fp64")) / (2 * h)
        jv = G.jvp(xd, v, precision="fp64")
        j32 = G.jvp(xd, v, precision="fp32")
        scale = fd.abs().max().item()
        worst = max(worst, (jv - fd).abs().max().item() / scale)
        worst32 = max(worst32, (j32 - jv).abs().max().item() / scale)
    print(f"64^3 melt: J v vs central differences max rel err {worst:.2e}; FP32 J v vs FP64 {worst32:.2e}")
    assert worst < 1e-6
    assert worst32 < 1e-3
