"""Batched 1-D adsorption problems (sfx_batch_*) against the oracle.

Each problem of a batch is also described as a single SF system
(configs.adsorption_problem) and evaluated by the oracle; the batch must
reproduce its residual, ln G_N and phi to 1e-10 and its DIIS solves."""
import numpy as np
import pytest

from field_amd import configs
from field_amd.configs import normal

pytestmark = pytest.mark.gpu

# uneven sizes, odd and even chains, the smallest legal problem, lambda 1/4,
# chi_AS = 0 and chi_s = 0
PROBLEMS = [
    {"n_layers": 2, "N": 2, "chi_as": 0.5, "chi_s": -1.0, "phibulk": 0.01},
    {"n_layers": 7, "N": 3, "chi_as": 0.3, "chi_s": -0.5, "phibulk": 0.2},
    {"n_layers": 40, "N": 25, "chi_as": 0.8, "chi_s": -2.0, "phibulk": 1e-3},
    {"n_layers": 33, "N": 64, "chi_as": 0.0, "chi_s": -1.5, "phibulk": 0.05},
    {"n_layers": 65, "N": 10, "chi_as": 0.5, "chi_s": 0.0, "phibulk": 0.1, "lambda_": 0.25},
    {"n_layers": 130, "N": 101, "chi_as": 0.4, "chi_s": -0.7, "phibulk": 1e-4},
    {"n_layers": 257, "N": 40, "chi_as": 0.2, "chi_s": -1.2, "phibulk": 0.03},
    {"n_layers": 100, "N": 100, "chi_as": 0.5, "chi_s": -1.0, "phibulk": 0.01},
]


def _start(seed, nv, sigma=0.1):
    return normal(seed, nv, sigma)


@pytest.fixture(scope="module")
def batch(sfx):
    return sfx.Batch(PROBLEMS)


def test_batch_layout_and_tuner(batch):
    assert batch.nvar == sum(2 * p["n_layers"] for p in PROBLEMS)
    assert list(np.diff(batch.offsets)) == [2 * p["n_layers"] for p in PROBLEMS]
    threads, ms = batch.tuning()
    print("tuner:", threads, ms)
    assert threads in (64, 128, 256, 512) and all(m > 0 for m in ms)


def test_batch_residual_matches_oracle(batch, orc):
    x = np.concatenate([_start(100 + i, 2 * p["n_layers"]) for i, p in enumerate(PROBLEMS)])
    f = batch.residual(x)
    lngn = batch.lngn()
    for i, (p, xi, fi) in enumerate(zip(PROBLEMS, batch.split(x), batch.split(f))):
        lat, sysd, _ = configs.adsorption_problem(p)
        O = orc.System(lat, sysd)
        fo = O.residual(xi)
        err = np.abs(fi - fo).max() / max(1.0, np.abs(fo).max())
        print(f"problem {i} {p}: residual rel err {err:.2e}")
        assert err < 1e-10
        assert abs(lngn[i] - O.lngn(0)) <= 1e-10 * max(1.0, abs(O.lngn(0)))
        for state in (0, 1):
            po = O.phi(state)
            pg = batch.phi(i, state)
            inner = slice(1, p["n_layers"] + 1)
            np.testing.assert_allclose(pg[inner], po[inner], rtol=1e-10, atol=1e-14)


def test_batch_residual_on_device(batch, orc):
    import torch

    x = np.concatenate([_start(200 + i, 2 * p["n_layers"], 0.05) for i, p in enumerate(PROBLEMS)])
    f_host = batch.residual(x)
    f_dev = batch.residual(torch.from_numpy(x).cuda()).cpu().numpy()
    np.testing.assert_array_equal(f_host, f_dev)


def test_batch_diis_matches_oracle(sfx, orc):
    B = sfx.Batch(PROBLEMS)
    x, its, norms, st = B.solve(tolerance=1e-7, max_iter=3000)
    print("batch DIIS:", list(its), st)
    for i, (p, xi) in enumerate(zip(PROBLEMS, B.split(x))):
        lat, sysd, x0 = configs.adsorption_problem(p)
        O = orc.System(lat, sysd)
        xo, ito, erro, rco = O.solve_diis(x0, tol=1e-7, max_iter=3000)
        print(f"problem {i}: gpu it {its[i]} |f| {norms[i]:.2e}; oracle it {ito} |f| {erro:.2e} rc {rco}")
        assert (norms[i] <= 1e-7) == (rco == 0)
        if rco == 0:
            # the same DIIS iteration up to the conditioning of its small solve
            assert abs(int(its[i]) - ito) <= max(3, ito // 20)
            assert np.linalg.norm(O.residual(xi)) <= 1e-7 * (1 + 1e-6)


def test_batch_warm_start_and_limits(sfx):
    B = sfx.Batch(PROBLEMS[:3])
    x, its, norms, st = B.solve(tolerance=1e-7, max_iter=3000)
    assert st["converged"]
    # from its own solution every problem stops at the first evaluation
    x2, its2, norms2, st2 = B.solve(x, tolerance=1e-7, max_iter=3000)
    assert list(its2) == [0, 0, 0] and st2["converged"]
    np.testing.assert_array_equal(x, x2)
    # an iteration limit ends every problem there and reports it
    x3, its3, norms3, st3 = B.solve(tolerance=1e-12, max_iter=5)
    assert not st3["converged"] and list(its3) == [5, 5, 5]
    with pytest.raises(sfx.SFError):
        B.solve(tolerance=1e-12, max_iter=5, check_converged=True)


def test_batch_rejects_bad_input(sfx):
    with pytest.raises(sfx.SFError):
        sfx.Batch([{"n_layers": 1, "N": 10, "chi_as": 0.5, "chi_s": -1.0, "phibulk": 0.01}])
    with pytest.raises(sfx.SFError):
        sfx.Batch([{"n_layers": 10, "N": 1, "chi_as": 0.5, "chi_s": -1.0, "phibulk": 0.01}])
    with pytest.raises(sfx.SFError):
        sfx.Batch([{"n_layers": 10, "N": 10, "chi_as": 0.5, "chi_s": -1.0, "phibulk": 1.5}])
    with pytest.raises(sfx.SFError):
        sfx.Batch([{"n_layers": 2000, "N": 10, "chi_as": 0.5, "chi_s": -1.0, "phibulk": 0.01}])


def test_uneven_sweep_sample(sfx, orc):
    """64 problems of the uneven sweep (layers 50..400, N 20..1000): every
    problem the batch reports converged has an oracle residual below the
    tolerance at the returned point."""
    probs = configs.batch_1d(64)
    B = sfx.Batch(probs)
    x, its, norms, st = B.solve(tolerance=1e-7, max_iter=3000)
    conv = norms <= 1e-7
    print(f"uneven sweep: {conv.sum()}/{len(probs)} converged, {st['seconds']:.3f} s, "
          f"max it {its.max()}, tuner {B.tuning()}")
    assert conv.sum() >= len(probs) // 2
    for i in np.flatnonzero(conv)[:16]:
        lat, sysd, _ = configs.adsorption_problem(probs[i])
        O = orc.System(lat, sysd)
        assert np.linalg.norm(O.residual(B.split(x)[i])) <= 1e-7 * (1 + 1e-6)


def test_uneven_sweep_backtracking(sfx, orc):
    """The same 64 problems with per-problem backtracking: no problem ends on a
    non-finite residual, every problem whose bulk is a stable Flory-Huggins
    solution converges, and every converged one passes the oracle."""
    probs = configs.batch_1d(64)
    B = sfx.Batch(probs)
    x, its, norms, st = B.solve(tolerance=1e-7, max_iter=3000, diis_backtrack=1, diis_eta=1.0)
    conv = norms <= 1e-7
    stable = np.array([configs.bulk_is_stable(p) for p in probs])
    print(f"uneven sweep, backtracking: {conv.sum()}/{len(probs)} converged, "
          f"{conv[stable].sum()}/{stable.sum()} with a stable bulk, {st['seconds']:.3f} s")
    assert np.isfinite(norms).all()
    assert conv[stable].all()
    for i in np.flatnonzero(conv)[:16]:
        lat, sysd, _ = configs.adsorption_problem(probs[i])
        O = orc.System(lat, sysd)
        assert np.linalg.norm(O.residual(B.split(x)[i])) <= 1e-7 * (1 + 1e-6)


def test_uneven_sweep_full_size(sfx, orc):
    """The whole 512-problem sweep (layers 50..400, N 20..1000, chi_AS 0..1,
    chi_s -2..0, phi_b 1e-4..1e-1) in one persistent launch with per-problem
    backtracking: no problem ends non-finite, every problem whose bulk is a
    stable Flory-Huggins solution converges, and every problem reported
    converged has an oracle residual below the tolerance at the returned
    point."""
    probs = configs.batch_1d(512)
    B = sfx.Batch(probs)
    print("512 problems, tuner", B.tuning(), flush=True)
    x, its, norms, st = B.solve(tolerance=1e-7, max_iter=5000, diis_backtrack=1, diis_eta=1.0)
    conv = norms <= 1e-7
    stable = np.array([configs.bulk_is_stable(p) for p in probs])
    print(f"512 problems: {conv.sum()} converged, {conv[stable].sum()}/{stable.sum()} with a stable bulk, "
          f"{st['seconds']:.3f} s, max iterations {its.max()}", flush=True)
    assert np.isfinite(norms).all()
    assert conv[stable].all()
    xs = B.split(x)
    worst = 0.0
    for i in np.flatnonzero(conv):
        lat, sysd, _ = configs.adsorption_problem(probs[i])
        r = np.linalg.norm(orc.System(lat, sysd).residual(xs[i]))
        worst = max(worst, r)
        assert r <= 1e-7 * (1 + 1e-6), (i, probs[i], r)
    print(f"largest oracle residual over the {conv.sum()} converged problems: {worst:.3e}")
