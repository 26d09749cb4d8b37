"""Solver bookkeeping through the C ABI: what DIIS and Newton-Krylov return
and report, and that one solver's leftovers in the shared work vectors do
not leak into the next solve."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _melt(sfx):
    """The small periodic melt of the parity cases, with the start the
    Newton-Krylov parity test solves from."""
    lat, sysd = cases.case("3d_sc_melt")
    return lat, sysd, sfx.System(lat, sysd), cases.start(lat, sysd, seed=15, sigma=0.05)


def test_nk_after_nonfinite_diis_converges(sfx, orc):
    """A DIIS solve that blows up leaves NaN/Inf in the solver's work vectors
    (they are kept per system). A Newton-Krylov solve on the same system from a
    good start must still converge: its vectors are cleared by stores, not by
    multiplying old contents by zero."""
    lat, sysd, G, x0 = _melt(sfx)
    print("DIIS with eta = 1e3 (diverges)", flush=True)
    with pytest.raises(sfx.SFError) as e:
        G.solve(x0, method="diis", max_iter=200, diis_eta=1e3)
    assert e.value.code in (-4, -6)  # non-finite, or the iteration limit on a huge |f|
    print("NK from the seeded start", flush=True)
    x, st = G.solve(x0, method="nk", tolerance=1e-9, max_iter=60, gmres_rtol=1e-4)
    assert st["converged"] and st["residual_norm"] < 1e-9
    assert np.linalg.norm(orc.System(lat, sysd).residual(x)) < 1e-8


def test_diis_reports_the_norm_of_the_returned_point(sfx):
    """At the iteration limit DIIS returns its last combination; the reported
    norm is that point's FP64 residual, not the previous evaluation's."""
    lat, sysd, G, x0 = _melt(sfx)
    x, st = G.solve(x0, method="diis", max_iter=7, check_converged=False)
    assert not st["converged"]
    f = G.residual(x, precision="fp64")
    assert abs(np.linalg.norm(f) - st["residual_norm"]) <= 1e-12 * np.linalg.norm(f)


def test_diis_returns_a_converged_start_unmoved(sfx):
    """A start that already meets the tolerance comes back bit for bit, as converged."""
    lat, sysd, G, x0 = _melt(sfx)
    xs, st = G.solve(x0, method="nk", tolerance=1e-9, max_iter=60, gmres_rtol=1e-4)
    assert st["converged"]
    x2, st2 = G.solve(xs, method="diis", tolerance=1e-7)
    assert st2["converged"] and st2["iterations"] == 0
    np.testing.assert_array_equal(x2, xs)
    assert st2["residual_norm"] == pytest.approx(np.linalg.norm(G.residual(xs, precision="fp64")), rel=1e-12)


def test_nk_rejects_a_bad_fp64_switch(sfx):
    lat, sysd, G, x0 = _melt(sfx)
    with pytest.raises(sfx.SFError) as e:
        G.solve(x0, method="nk", fp64_switch=-1.0)
    assert e.value.code == -1


def test_nk_low_fp64_switch_falls_back_to_fp64(sfx, orc):
    """With the FP64 switch below the FP32 residual's rounding floor the FP32
    line search stalls; the solve continues in FP64 and converges."""
    lat, sysd, G, x0 = _melt(sfx)
    x, st = G.solve(x0, method="nk", tolerance=1e-9, max_iter=60, gmres_rtol=1e-4, jvp_precision="fp32",
                    fp64_switch=1e-12, ptc_dt0=0.0)
    print("NK with fp64_switch 1e-12:", st)
    assert st["converged"] and st["residual_evals"] >= 1
    assert np.linalg.norm(orc.System(lat, sysd).residual(x)) < 1e-8
