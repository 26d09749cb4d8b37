"""An SFbox input with a var sweep, solved on the GPU (field_amd.sfbox_input.run)
and checked step by step against the oracle."""
import numpy as np
import pytest

from field_amd.sfbox_input import run

pytestmark = pytest.mark.gpu

SWEEP = """
lat : flat : gradients : 1
lat : flat : geometry : flat
lat : flat : n_layers : 40
lat : flat : lowerbound : surface
lat : flat : upperbound : mirror1
lat : flat : lambda : 0.16666666666666666
mon : A : freedom : free
mon : S : freedom : free
mon : W : freedom : frozen
mon : W : frozen_range : lowerbound
mon : A : chi - S : 0.5
mon : A : chi - W : -1.0
mol : pol : composition : (A)20
mol : pol : freedom : free
mol : pol : phibulk : 0.001
mol : sol : composition : (S)1
mol : sol : freedom : solvent
newton : isaac : method : DIIS
newton : isaac : tolerance : 1e-9
newton : isaac : iterationlimit : 5000
var : mol : pol : phibulk
var : mol : step : 1
var : mol : end_value : 0.1
var : mol : scale : exponential
var : mol : type : real
start
"""


def test_var_sweep_input(orc, tmp_path):
    """phi_b from 1e-3 to 1e-1 in decades, each step warm-started from the
    last: every solution passes the oracle's residual, and the free energy,
    grand potential and chemical potentials match the oracle's there. The kal
    table holds a header and one row per step."""
    kal = tmp_path / "sweep.kal"
    res = run(SWEEP, kal=str(kal))
    assert [r["var"] for r in res] == pytest.approx([1e-3, 1e-2, 1e-1])
    rows = kal.read_text().splitlines()
    assert len(rows) == 4 and rows[0].startswith("sys : noname : free energy")
    assert [float(r.split("\t")[0]) for r in rows[1:]] == pytest.approx([r["free_energy"] for r in res], rel=1e-12)
    for r in res:
        assert r["converged"], r
        O = orc.System(r["lattice"], r["system"])
        assert np.linalg.norm(O.residual(r["x"])) < 1e-9 * (1 + 1e-6)
        assert abs(r["free_energy"] - O.free_energy()) <= 1e-10 * max(1, abs(O.free_energy()))
        assert abs(r["grand_potential"] - O.grand_potential()) <= 1e-10 * max(1, abs(O.grand_potential()))
        for m in range(2):
            assert abs(r["chemical_potential"][m] - O.chemical_potential(m)) <= 1e-10 * max(1, abs(O.chemical_potential(m)))
