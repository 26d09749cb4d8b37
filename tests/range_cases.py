"""Systems with a range_restricted molecule (SF_Molecule.cpp:169-180): the
molecule's phi summed over a box of interior sites is fixed."""
import numpy as np

from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec


def _two_state(lat, comp, lo, hi, phi_range, chi=0.8):
    chi_m = np.array([[0.0, chi], [chi, 0.0]])
    if "B" in comp:
        chi_m = np.array([[0.0, 0.5, chi], [0.5, 0.0, 0.3], [chi, 0.3, 0.0]])
        states = ("A", "B", "S")
    else:
        states = ("A", "S")
    return lat, SystemSpec(states=states, chi=chi_m,
                           molecules=[MoleculeSpec("pol", comp, "range_restricted", range_lo=lo, range_hi=hi,
                                                   phi_range=phi_range),
                                      MoleculeSpec("sol", "(S)1", "solvent")])


def case(name):
    if name == "1d_homopolymer":
        lat = LatticeSpec(1, (60,), lowerbound=("surface",), upperbound=("mirror1",), lambda_=1 / 6)
        return _two_state(lat, "(A)20", (1, 1, 1), (10, 1, 1), 3.0)
    if name == "1d_copolymer":
        lat = LatticeSpec(1, (50,), lowerbound=("mirror1",), upperbound=("surface",), lambda_=1 / 6)
        return _two_state(lat, "(A)7(B)13", (30, 1, 1), (50, 1, 1), 4.0)
    if name == "1d_cylinder":
        lat = LatticeSpec(1, (40,), geometry="cylindrical", lowerbound=("mirror1",), upperbound=("surface",),
                          lambda_=1 / 6)
        return _two_state(lat, "(A)12", (1, 1, 1), (8, 1, 1), 2.5)
    if name == "2d_flat_box":
        lat = LatticeSpec(2, (14, 12), lowerbound=("surface", "mirror1"), upperbound=("mirror1", "periodic"),
                          lambda_=1 / 6)
        return _two_state(lat, "(A)5(B)5", (1, 3, 1), (4, 9, 1), 9.0)
    if name == "1d_order2":
        lat = LatticeSpec(1, (40,), lowerbound=("surface",), upperbound=("mirror1",), lambda_=1 / 6, order=2)
        return _two_state(lat, "(A)10", (1, 1, 1), (6, 1, 1), 2.0)
    raise KeyError(name)


NAMES = ["1d_homopolymer", "1d_copolymer", "1d_cylinder", "2d_flat_box", "1d_order2"]


def box_mask(lat, lo, hi):
    """Interior sites of the box, as a mask over the halo-padded lattice."""
    dims = lat.dims
    idx = np.indices(dims)
    m = np.ones(dims, dtype=bool)
    for a in range(lat.gradients):
        m &= (idx[a] >= lo[a]) & (idx[a] <= hi[a])
    return m.reshape(-1)
