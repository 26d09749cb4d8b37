"""1-D FP64 chains in one workgroup launch per stretch of ranks (the running
rank in LDS) against one launch per rank: the same bits for the residual, the
profiles and ln G_N, on every 1-D first-order system of the parity set, the
range_restricted set and long chains with odd and even lengths."""
import numpy as np
import pytest

import cases
import range_cases
from field_amd import configs

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    for name in cases.NAMES:
        lat, sysd = cases.case(name)
        if lat.gradients == 1 and lat.order == 1:
            out.append((name, lat, sysd))
    for name in range_cases.NAMES:
        lat, sysd = range_cases.case(name)
        if lat.gradients == 1 and lat.order == 1:
            out.append(("range_" + name, lat, sysd))
    for N, nl in ((1000, 400), (999, 250), (100, 100)):
        lat, sysd, _ = configs.adsorption_1d(n_layers=nl, N=N, phibulk=1e-3)
        out.append((f"adsorption_N{N}_layers{nl}", lat, sysd))
    return out


@pytest.mark.parametrize("name,lat,sysd", _systems(), ids=[s[0] for s in _systems()])
def test_chain_launch_matches_per_rank_launches(sfx, name, lat, sysd):
    print(name, flush=True)
    A = sfx.System(lat, sysd)
    B = sfx.System(lat, sysd, sweep_kernel="per_rank")
    x = np.random.default_rng(6).normal(0, 0.1, A.nvar)
    fa, fb = A.residual(x, precision="fp64"), B.residual(x, precision="fp64")
    assert np.array_equal(fa, fb)
    for t in range(len(sysd.states)):
        assert np.array_equal(A.phi(t), B.phi(t))
    for m in range(len(sysd.molecules)):
        assert A.lngn(m) == B.lngn(m)
