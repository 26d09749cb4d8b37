"""SFNewton's default iteration (pseudo-Hessian Newton, Scheutjens line search).

The oracle's factorisation routines are pinned against dense linear algebra,
its iteration against closed-form roots and an independent solver's fixed
point; the library's host iteration (sfx_newton_iterate, no GPU work) is then
held to the oracle bit for bit on the same residual functions, including runs
that reset the pseudo-Hessian, reverse directions, shorten steps five times in
one line search, and stop at the iteration limit.
"""

import numpy as np
import pytest

from field_amd import _abi, configs, sfx
from oracle import orc


def _random_ldu(n, rng):
    h = np.zeros((n, n), dtype=np.float32)
    a = rng.normal(0, 0.3, (n, n)).astype(np.float32)
    h[:] = np.tril(a, -1) + np.triu(a, 1)
    np.fill_diagonal(h, rng.uniform(0.5, 2.0, n).astype(np.float32) * rng.choice([-1, 1], n))
    # element (r, c) lives at flat index r + n c: a C-ordered array holds the transpose
    return np.ascontiguousarray(h.T)


def test_ldu_update_is_a_rank_one_update():
    """updatpos: L D U + alpha w v^T, refactorised in place (float storage)."""
    rng = np.random.default_rng(0)
    for n in (2, 5, 17):
        h = _random_ldu(n, rng)
        before = orc.ldu_dense(h)
        w, v = rng.normal(size=n), rng.normal(size=n)
        alpha = 0.37
        orc.ldu_update(h, w, v, alpha)
        after = orc.ldu_dense(h)
        np.testing.assert_allclose(after, before + alpha * np.outer(w, v), rtol=0, atol=2e-5 * np.abs(before).max())


def test_ldu_solve_multiply_sign():
    rng = np.random.default_rng(1)
    for n in (2, 6, 23):
        h = _random_ldu(n, rng)
        H = orc.ldu_dense(h)
        g = rng.normal(size=n)
        np.testing.assert_allclose(orc.ldu_solve(h, g), -np.linalg.solve(H, g), rtol=1e-9, atol=1e-9)
        w = rng.normal(size=n)
        np.testing.assert_allclose(orc.ldu_multiply(h, w, 1.5), 1.5 * H @ w, rtol=1e-12, atol=1e-12)
        assert orc.ldu_sign(h) == int(np.sign(np.linalg.det(H)))


def test_oracle_newton_start_and_accuracy():
    """Before the first step: H = (1 + |g|) I in float, p = -g / (1 + |g|) and
    the accuracy sqrt(|p| |g| / (1 + |x|)) (sfnewton.cpp:1523-1530, 2040)."""
    rng = np.random.default_rng(2)
    x0 = rng.normal(size=7)
    c = rng.normal(size=7)
    f = lambda x: np.tanh(x) - 0.5 * c
    _, st = orc.newton(f, x0, iterationlimit=0)
    g = f(x0)
    p = -g / np.float64(np.float32(1 + np.linalg.norm(g)))
    assert st["iterations"] == 0 and st["functioncalls"] == 1
    assert st["accuracy"] == pytest.approx(np.sqrt(np.linalg.norm(p) * np.linalg.norm(g) / (1 + np.linalg.norm(x0))),
                                           rel=1e-14)


@pytest.mark.parametrize("case", ["linear", "cubic", "exp"])
def test_oracle_newton_finds_closed_form_roots(case):
    rng = np.random.default_rng(3)
    n = 8
    c = rng.normal(size=n)
    if case == "linear":
        B = rng.normal(size=(n, n))
        A = B @ B.T + n * np.eye(n)
        f, root = (lambda x: A @ x - c), np.linalg.solve(A, c)
    elif case == "cubic":
        f = lambda x: x ** 3 + x - 5 * c
        root = np.array([np.real(r[np.abs(np.imag(r)) < 1e-9][0]) for r in (np.roots([1, 0, 1, -5 * ci]) for ci in c)])
    else:
        f, root = (lambda x: np.exp(x) - np.exp(c)), c
    x, st = orc.newton(f, np.zeros(n), tolerance=1e-12, iterationlimit=500)
    assert st["converged"]
    np.testing.assert_allclose(x, root, atol=1e-9)


def test_oracle_newton_reaches_the_diis_fixed_point():
    """100-layer wall adsorption (A_100, phi_b 0.01, chi_s -1): Newton's fixed
    point is the one the oracle's DIIS (MPS.cpp:228-284) finds independently."""
    lat, sysd, x0 = configs.adsorption_1d()
    O = orc.System(lat, sysd)
    xn, st = O.solve_newton(x0, tolerance=1e-11, iterationlimit=500)
    assert st["converged"]
    xd, it, err, rc = O.solve_diis(x0, tol=1e-11, max_iter=20000)
    assert rc == 0 and err < 1e-11
    np.testing.assert_allclose(xn, xd, atol=1e-8)


def _rosen(x):
    return np.concatenate([[-400 * x[0] * (x[1] - x[0] ** 2) - 2 * (1 - x[0])],
                           200 * (x[1:-1] - x[:-2] ** 2) - 400 * x[1:-1] * (x[2:] - x[1:-1] ** 2) - 2 * (1 - x[1:-1]),
                           [200 * (x[-1] - x[-2] ** 2)]])


def _counting(f):
    calls = [0]

    def g(x):
        calls[0] += 1
        return f(x) + 1e-13 * calls[0]
    return g


def _cases():
    rng = np.random.default_rng(1)
    n = 10
    c = rng.normal(size=n)
    B = rng.normal(size=(n, n))
    lat, sysd, x0 = configs.adsorption_1d()
    ads = orc.System(lat, sysd)
    lat2, sysd2, x02 = configs.adsorption_1d(n_layers=200, N=500, phibulk=1e-4, chi_s=-1.5)
    ads2 = orc.System(lat2, sysd2)
    return {
        "wall_adsorption": (lambda: ads.residual, x0, dict(tolerance=1e-7, max_iter=1000)),
        "adsorption_resets": (lambda: ads2.residual, x02, dict(tolerance=1e-7, max_iter=3000)),
        "rosenbrock_reversals": (lambda: _rosen, -np.ones(n), dict(tolerance=1e-8, max_iter=2000)),
        "indefinite_limit": (lambda: (lambda x: B @ x - c), np.zeros(n), dict(tolerance=1e-8, max_iter=300)),
        "five_step_line_search": (lambda: _counting(lambda x: np.exp(300 * x) - 1), np.full(2, -1e-3),
                                  dict(tolerance=1e-9, max_iter=300)),
        "steep_trust_region": (lambda: (lambda x: np.exp(3 * x) - 1), np.full(4, -3.0),
                               dict(tolerance=1e-9, max_iter=300, newton_deltamax=100.0)),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_library_newton_matches_oracle_bit_for_bit(name):
    make, x0, kw = _cases()[name]
    okw = dict(tolerance=kw["tolerance"], iterationlimit=kw["max_iter"])
    if "newton_deltamax" in kw:
        okw["deltamax"] = kw["newton_deltamax"]
    xo, so = orc.newton(make(), x0, **okw)
    xp, sp = sfx.newton_iterate(make(), x0, check_converged=False, **kw)
    print(name, so)
    assert sp["iterations"] == so["iterations"]
    assert sp["residual_evals"] == so["functioncalls"]
    assert sp["converged"] == so["converged"]
    assert np.array_equal(xp, xo)
    assert sp["accuracy"] == so["accuracy"] and sp["residual_norm"] == so["residual_norm"]
    if name == "adsorption_resets":
        assert so["hessian_resets"] >= 2
    if name == "rosenbrock_reversals":
        assert so["reversed_directions"] > 0
    if name == "indefinite_limit":
        assert not so["converged"] and so["iterations"] == 300
    if name == "five_step_line_search":
        assert so["time_dependent"]


def test_library_newton_reports_errors():
    with pytest.raises(sfx.SFError) as e:
        sfx.newton_iterate(lambda x: x, np.zeros(1))
    assert e.value.code == _abi.SFX_E_INVALID
    with pytest.raises(sfx.SFError) as e:
        sfx.newton_iterate(lambda x: x, np.zeros(3), newton_deltamax=0.0)
    assert e.value.code == _abi.SFX_E_INVALID
    with pytest.raises(sfx.SFError) as e:
        sfx.newton_iterate(lambda x: x - 1, np.zeros(3), max_iter=1)
    assert e.value.code == _abi.SFX_E_NOTCONVERGED
    x, st = sfx.newton_iterate(lambda x: x - 1, np.zeros(3), max_iter=1, check_converged=False)
    assert st["iterations"] == 1 and not st["converged"]
