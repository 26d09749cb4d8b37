"""range_restricted molecules on the GPU (SF_Copol1stO.cpp:1246-1280) against
the oracle: residual, profiles, phi_b, ln G_N and free energy in FP64 and
FP32, a Newton solve to the oracle's fixed point, and the refusal of Jacobian
products."""
import numpy as np
import pytest

import range_cases
from field_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", range_cases.NAMES)
def test_range_restricted_parity(sfx, orc, name):
    lat, sysd = range_cases.case(name)
    print(name, flush=True)
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    x = np.random.default_rng(4).normal(0, 0.1, G.nvar)
    fo = O.residual(x)
    np.testing.assert_allclose(G.residual(x, precision="fp64"), fo, rtol=1e-10, atol=1e-12)
    for t in range(len(sysd.states)):
        np.testing.assert_allclose(G.phi(t), O.phi(t), rtol=1e-10, atol=1e-13)
    assert G.phibulk(0) == pytest.approx(O.phibulk(0), rel=1e-10)
    assert G.phibulk(1) == pytest.approx(O.phibulk(1), rel=1e-10)
    assert G.lngn(0) == pytest.approx(O.lngn(0), rel=1e-10, abs=1e-12)
    assert G.free_energy() == pytest.approx(O.free_energy(), rel=1e-10, abs=1e-10)
    mol = sysd.molecules[0]
    box = range_cases.box_mask(lat, mol.range_lo, mol.range_hi)
    amount = sum(G.phi(t)[box].sum() for t, s in enumerate(sysd.states) if s != "S")
    assert amount == pytest.approx(mol.phi_range, rel=1e-12)
    f32 = G.residual(x, precision="fp32")
    err = np.abs(f32 - fo).max()
    print(name, "fp32 max err", err, flush=True)
    assert err < 2e-4


def test_range_restricted_newton_solve(sfx, orc):
    lat, sysd = range_cases.case("1d_homopolymer")
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    x0 = np.zeros(G.nvar)
    xo, so = O.solve_newton(x0, tolerance=1e-9, iterationlimit=500)
    x, st = G.solve(x0, method="newton", tolerance=1e-9, max_iter=500)
    print("gpu", st, "oracle", so, flush=True)
    assert st["converged"] and abs(st["iterations"] - so["iterations"]) <= max(3, so["iterations"] // 10)
    np.testing.assert_allclose(x, xo, atol=1e-7)
    assert np.linalg.norm(O.residual(x)) < 1e-8


def test_range_restricted_refuses_jacobian_products(sfx):
    lat, sysd = range_cases.case("1d_copolymer")
    G = sfx.System(lat, sysd)
    x = np.zeros(G.nvar)
    for call in (lambda: G.jvp(x, x + 1), lambda: G.vjp(x, x + 1), lambda: G.solve(x, method="nk")):
        with pytest.raises(sfx.SFError) as e:
            call()
        assert e.value.code == _abi.SFX_E_UNSUPPORTED
