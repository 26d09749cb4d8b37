"""SFNewton's default iteration on the GPU residual (sfx_solve with
SFX_SOLVER_NEWTON): the host pseudo-Hessian Newton against the oracle's
restatement on the same systems, and the captured residual graph it replays
against plain launches."""
import numpy as np
import pytest

import cases
from field_amd import _abi, configs

pytestmark = pytest.mark.gpu


def test_newton_wall_adsorption_reaches_the_oracle_fixed_point(sfx, orc):
    """100 layers, A_100 at phi_b 0.01, chi_s = -1, u = 0, tolerance 1e-7: the
    reference's stop rule (accuracy <= 1e-7, |g| <= 1e-6) on the GPU residual;
    the same iteration as the oracle's Newton up to the residual's rounding."""
    lat, sysd, x0 = configs.adsorption_1d()
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    xo, so = O.solve_newton(x0, tolerance=1e-7, iterationlimit=1000)
    x, st = G.solve(x0, method="newton", tolerance=1e-7, max_iter=1000)
    print("wall adsorption: gpu", st, "oracle", so, flush=True)
    assert st["converged"] and st["accuracy"] <= 1e-7 and st["residual_norm"] <= 1e-6
    assert abs(st["iterations"] - so["iterations"]) <= 2
    np.testing.assert_allclose(x, xo, atol=1e-7)
    assert np.linalg.norm(O.residual(x)) <= 1e-6
    O.residual(xo)
    G.residual(x)
    assert abs(G.free_energy() - O.free_energy()) <= 1e-8 * abs(O.free_energy())


@pytest.mark.parametrize("name", cases.NAMES)
def test_newton_parity_cases(sfx, orc, name):
    """Every parity system from its seeded start, tolerance 1e-8."""
    lat, sysd = cases.case(name)
    G, O = sfx.System(lat, sysd), orc.System(lat, sysd)
    x0 = cases.start(lat, sysd, seed=15, sigma=0.05)
    print(name, flush=True)
    xo, so = O.solve_newton(x0, tolerance=1e-8, iterationlimit=2000)
    x, st = G.solve(x0, method="newton", tolerance=1e-8, max_iter=2000)
    print(name, "gpu", st["iterations"], st["residual_evals"], "oracle", so["iterations"], so["functioncalls"],
          flush=True)
    assert st["converged"]
    assert abs(st["iterations"] - so["iterations"]) <= max(3, so["iterations"] // 10)
    assert np.linalg.norm(O.residual(x)) <= 1e-7
    # the melt form's iteration vector carries its halo sites; the library
    # returns them as 0, the oracle leaves the start's values there (the
    # residual never sees them): compare where the residual is defined
    live = O.residual(x0) != 0
    np.testing.assert_allclose(x[live], xo[live], atol=1e-6)


def test_newton_graph_replays_match_plain_launches(sfx):
    """The solve replays the residual from a captured HIP graph; with kernel
    timing on it launches every kernel instead. Both give the same bits, over
    every replay of the solve, and again after a chi change drops the graph
    (against a system created at the new chi)."""
    lat, sysd, x0 = configs.adsorption_1d()
    G = sfx.System(lat, sysd)
    xg, sg = G.solve(x0, method="newton", tolerance=1e-9, max_iter=1000)
    G.profile_start()
    xe, se = G.solve(x0, method="newton", tolerance=1e-9, max_iter=1000)
    G.profile_stop()
    print("graph", sg, "plain", se, flush=True)
    assert sg["residual_evals"] == se["residual_evals"] and sg["residual_evals"] > 10
    # the first solve captures on its first residual; plain launches with timing on
    assert sg["graph_replays"] == sg["residual_evals"] - 1 and se["graph_replays"] == 0
    assert np.array_equal(xg, xe)
    # a chi change between solves: graph dropped, captured again
    G.set_chi(0, 1, 0.7)
    x2, s2 = G.solve(x0, method="newton", tolerance=1e-9, max_iter=1000)
    assert s2["graph_replays"] == s2["residual_evals"] - 1
    lat3, sysd3, _ = configs.adsorption_1d(chi_as=0.7)
    H = sfx.System(lat3, sysd3)
    H.profile_start()
    x3, s3 = H.solve(x0, method="newton", tolerance=1e-9, max_iter=1000)
    H.profile_stop()
    assert s2["residual_evals"] == s3["residual_evals"]
    assert np.array_equal(x2, x3)
    assert not np.array_equal(x2, xg)


def test_newton_refuses_large_systems(sfx):
    lat, sysd, x0 = configs.adsorption_1d(n_layers=9000)
    G = sfx.System(lat, sysd)
    assert G.nvar > 16384
    with pytest.raises(sfx.SFError) as e:
        G.solve(x0, method="newton")
    assert e.value.code == _abi.SFX_E_UNSUPPORTED
