"""range_restricted molecules in the oracle (SF_Copol1stO.cpp:1246-1280): the
range condition holds exactly, and the profile equals that of a restricted
molecule (fixed theta) at the theta the range condition produced -- the same
chain statistics reached by the other freedom's normalisation."""
import dataclasses

import numpy as np
import pytest

import range_cases
from oracle import orc


def _pol_states(sysd):
    return [i for i, s in enumerate(sysd.states) if s != "S"]


@pytest.mark.parametrize("name", range_cases.NAMES)
def test_range_condition_and_restricted_equivalence(name):
    lat, sysd = range_cases.case(name)
    O = orc.System(lat, sysd)
    x = np.random.default_rng(4).normal(0, 0.1, O.nvar)
    f = O.residual(x)
    mol = sysd.molecules[0]
    box = range_cases.box_mask(lat, mol.range_lo, mol.range_hi)
    phis = {t: O.phi(t) for t in _pol_states(sysd)}
    assert sum(p[box].sum() for p in phis.values()) == pytest.approx(mol.phi_range, rel=1e-13)
    theta = O.phibulk(0) * np.exp(O.lngn(0))
    rest = dataclasses.replace(sysd, molecules=[dataclasses.replace(mol, freedom="restricted", theta=theta),
                                                sysd.molecules[1]])
    R = orc.System(lat, rest)
    fr = R.residual(x)
    for t, p in phis.items():
        np.testing.assert_allclose(R.phi(t), p, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(fr, f, rtol=1e-11, atol=1e-13)
    assert R.phibulk(0) == pytest.approx(O.phibulk(0), rel=1e-12)


def test_range_restricted_solution_keeps_the_range_amount():
    lat, sysd = range_cases.case("1d_homopolymer")
    O = orc.System(lat, sysd)
    xs, st = O.solve_newton(np.zeros(O.nvar), tolerance=1e-9, iterationlimit=500)
    assert st["converged"]
    O.compute_phi(xs)
    mol = sysd.molecules[0]
    box = range_cases.box_mask(lat, mol.range_lo, mol.range_hi)
    assert O.phi(0)[box].sum() == pytest.approx(mol.phi_range, rel=1e-13)
    assert np.linalg.norm(O.residual(xs)) < 1e-8


def test_input_file_range_restricted_molecule():
    """`freedom : range_restricted` with restricted_range and phi_range
    (SF_Molecule.cpp:169-180) maps to the box and amount, and the solve runs
    SFNewton (no Jacobian products for this freedom)."""
    from field_amd import sfbox_input as si
    txt = "\n".join([
        "lat : flat : n_layers : 60", "lat : flat : geometry : flat", "lat : flat : gradients : 1",
        "lat : flat : lambda : 0.16666666666666667", "lat : flat : lowerbound : surface",
        "lat : flat : upperbound : mirror1", "mon : A : freedom : free", "mon : S : freedom : free",
        "mon : A : chi - S : 0.8", "mol : pol : composition : (A)20", "mol : pol : freedom : range_restricted",
        "mol : pol : restricted_range : 1;10", "mol : pol : phi_range : 3.0", "mol : sol : composition : (S)1",
        "mol : sol : freedom : solvent", "start", ""])
    lat, sysd, solver = si.build(si.Input(txt))
    mol = sysd.molecules[0]
    assert mol.freedom == "range_restricted" and mol.phi_range == 3.0
    assert tuple(mol.range_lo) == (1, 1, 1) and tuple(mol.range_hi) == (10, 1, 1)
    assert solver["method"] == "newton"
    ref_lat, ref_sys = range_cases.case("1d_homopolymer")
    x = np.random.default_rng(5).normal(0, 0.1, orc.System(lat, sysd).nvar)
    np.testing.assert_array_equal(orc.System(lat, sysd).residual(x), orc.System(ref_lat, ref_sys).residual(x))
