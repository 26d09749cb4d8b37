"""Transposed Jacobian products J^T w (sfx_vjp) through the C ABI.

Pinned by the transpose identity <w, J v> = <J^T w, v> against the tangent
product J v (itself checked against the oracle's central differences in the
parity suite) on every parity case the transposed sweep supports, in FP64 to
1e-10 relative; by torch.autograd.gradcheck on a small melt (finite
differences of the GPU residual); and by the explicit refusal on the lattices
whose backward sweep is not the forward sweep's adjoint."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

# lattices whose backward sweep is not the adjoint of the forward sweep: the
# second-order 2-D cylinder (the reference's backward sweep), HEX with mirror
# sides (a neighbour set that is not reflection symmetric), a second-order 2-D
# mirror2 lower bound (components copied without reversal)
UNSUPPORTED = {"2d_cyl_order2", "3d_hex_three_molecules", "2d_order2_copolymer"}


def _hex_periodic():
    """The three-molecule HEX system (a free chain, a restricted monomer, the
    solvent) on an all-periodic HEX lattice, where the transposed sweep applies."""
    from field_amd.spec import LatticeSpec
    _, sysd = cases.case("3d_hex_three_molecules")
    return LatticeSpec(3, (5, 6, 4), lattice_type="hex", lowerbound=("periodic",) * 3,
                       upperbound=("periodic",) * 3), sysd


EXTRA = {"3d_hex_periodic": _hex_periodic}


def _case(name):
    return EXTRA[name]() if name in EXTRA else cases.case(name)


@pytest.fixture(scope="module")
def systems(sfx):
    return {name: sfx.System(*_case(name)) for name in list(cases.NAMES) + list(EXTRA)}


def _dirs(G, lat, sysd, seed):
    rng = np.random.default_rng(seed)
    x = cases.start(lat, sysd, seed=seed)
    # directions on the iteration variables only (the melt layout keeps its halos at zero)
    ref = G.residual(x, precision="fp64")
    mask = np.ones(G.nvar) if sysd.form == "sfbox" else np.tile(lat.interior_mask().reshape(-1), G.nvar // lat.M)
    return x, rng.normal(0, 1, G.nvar) * mask, rng.normal(0, 1, G.nvar) * mask


@pytest.mark.parametrize("name", [n for n in cases.NAMES if n not in UNSUPPORTED] + list(EXTRA))
def test_transpose_identity(systems, name):
    lat, sysd = _case(name)
    G = systems[name]
    print(f"case {name} nvar={G.nvar}", flush=True)
    worst32 = 0.0
    for seed in (5, 6):
        x, v, w = _dirs(G, lat, sysd, seed)
        jv = G.jvp(x, v, precision="fp64")
        jtw = G.vjp(x, w, precision="fp64")
        lhs, rhs = float(w @ jv), float(jtw @ v)
        scale = np.linalg.norm(w) * np.linalg.norm(jv)
        print(f"  <w, Jv> = {lhs:.15e}, <J^T w, v> = {rhs:.15e}", flush=True)
        assert abs(lhs - rhs) <= 1e-10 * scale, (lhs, rhs, scale)
        # FP32 sweeps: the same identity to FP32 accuracy (reported, loosely bounded)
        jtw32 = G.vjp(x, w, precision="fp32")
        worst32 = max(worst32, abs(float(jtw32 @ v) - lhs) / scale)
    print(f"  FP32 transposed product: identity holds to {worst32:.2e}")
    assert worst32 < 1e-4


@pytest.mark.parametrize("name", sorted(UNSUPPORTED))
def test_unsupported_lattice_is_refused(systems, sfx, name):
    lat, sysd = cases.case(name)
    G = systems[name]
    x = cases.start(lat, sysd)
    with pytest.raises(sfx.SFError) as e:
        G.vjp(x, np.ones(G.nvar))
    assert e.value.code == -3


def test_torch_gradcheck_on_small_melt(sfx):
    """torch.autograd.gradcheck of the residual's backward (J^T w) against
    finite differences of the GPU residual, 6^3 periodic A3B5 melt."""
    import torch

    from field_amd import configs, torch_ops

    lat, sysd, x = configs.melt_3d(n=6, n_a=3, n_b=5, chi=0.5, sigma=0.1, seed=9)
    G = sfx.System(lat, sysd)
    mask = torch.from_numpy(np.tile(lat.interior_mask().reshape(-1), 2)).cuda()
    u0 = torch.from_numpy(x).cuda()

    def f(u):  # the residual as a function of the interior fields (halos stay 0)
        return torch_ops.residual(u * mask, G)

    u = u0.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(f, (u,), eps=1e-6, atol=1e-6, rtol=1e-5, fast_mode=True)
    # a plain backward gives J^T w
    w = torch.randn(G.nvar, dtype=torch.float64, device="cuda") * mask
    out = f(u)
    (gu,) = torch.autograd.grad(out, u, w)
    ref = torch.from_numpy(G.vjp(x, w.cpu().numpy())).cuda() * mask
    assert torch.allclose(gu, ref, rtol=1e-12, atol=1e-14)
