"""The Flory-Huggins bulk classification of the batched adsorption problems
(configs.fh_binodal, configs.bulk_is_stable), pinned by what the mathematics
fixes: the symmetric blend's closed form, the critical point, and the common
tangent lying below f everywhere between the two phases."""
import math

import numpy as np
import pytest

from field_amd import configs


def _f(phi, N, chi):
    return phi / N * np.log(phi) + (1 - phi) * np.log1p(-phi) + chi * phi * (1 - phi)


def test_symmetric_blend_closed_form():
    """N = 1: the phases are mirror images, phi_b = 1 - phi_a, and each solves
    ln(phi / (1 - phi)) = chi (2 phi - 1)."""
    for chi in (2.05, 2.5, 3.0, 5.0):
        a, b = configs.fh_binodal(1, chi)
        assert abs(a + b - 1) < 1e-12
        assert abs(math.log(a / (1 - a)) - chi * (2 * a - 1)) < 1e-10


def test_below_and_near_critical_point():
    for N in (1, 20, 600):
        chic = 0.5 * (1 + N ** -0.5) ** 2
        assert configs.fh_binodal(N, chic * (1 - 1e-9)) is None
        a, b = configs.fh_binodal(N, chic * (1 + 1e-6))
        pc = 1 / (1 + math.sqrt(N))
        assert abs(a - pc) < 0.01 * pc and abs(b - pc) < 0.01 * pc


@pytest.mark.parametrize("N, chi", [(20, 0.8), (100, 0.7), (600, 0.9), (1000, 1.0)])
def test_common_tangent_is_the_lower_hull(N, chi):
    """The tangent through both phases touches f there and lies below it at every
    point in between (so it is the binodal and not another pair of roots)."""
    a, b = configs.fh_binodal(N, chi)
    assert 0 < a < b < 1
    slope = (_f(b, N, chi) - _f(a, N, chi)) / (b - a)
    d1 = lambda p: (math.log(p) + 1) / N - math.log1p(-p) - 1 + chi * (1 - 2 * p)
    assert abs(d1(b) - slope) < 1e-8 * max(1, abs(slope))
    if a > 1e-250:  # the dilute point is representable
        assert abs(d1(a) - slope) < 1e-8 * max(1, abs(slope))
    p = np.linspace(max(a, 1e-300), b, 2001)[1:-1]
    line = _f(b, N, chi) + slope * (p - b)
    assert (_f(p, N, chi) - line >= -1e-12).all()


def test_batch_sweep_classification():
    """Of the 512 uneven problems, 278 sit below the critical chi and 16 more
    have a bulk below the dilute binodal."""
    probs = configs.batch_1d(512)
    stable = np.array([configs.bulk_is_stable(p) for p in probs])
    below_c = np.array([p["chi_as"] <= 0.5 * (1 + p["N"] ** -0.5) ** 2 for p in probs])
    assert below_c.sum() == 278 and stable.sum() == 294
