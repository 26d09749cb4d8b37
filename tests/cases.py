"""Small systems covering every lattice, molecule architecture and freedom of
the hot path; shared by the GPU parity tests and smoke()."""
import math

import numpy as np

from field_amd import configs
from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec


def _chi(states, pairs):
    n = len(states)
    c = np.zeros((n, n))
    for a, b, v in pairs:
        i, j = states.index(a), states.index(b)
        c[i, j] = c[j, i] = v
    return c


def _wall(lat, where):
    MX, MY, MZ = lat.dims
    w = np.zeros((MX, MY, MZ))
    if where == "x0":
        w[0] = 1
    elif where == "x1":
        w[MX - 1] = 1
    elif where == "y1":
        w[:, MY - 1] = 1
    return w.reshape(-1)


def case(name):
    if name == "1d_flat_adsorption":
        lat, sysd, _ = configs.adsorption_1d(n_layers=30, N=20, phibulk=0.01, chi_as=0.5, chi_s=-1.0)
        return lat, sysd
    if name == "1d_flat_odd_mirror2":
        lat = LatticeSpec(1, (25,), lowerbound=("surface",), upperbound=("mirror2",), lambda_=0.25,
                          layer_adjustment=1.0)
        st = ("A", "B", "S", "W")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.8), ("A", "S", 0.3), ("B", "S", 1.1), ("B", "W", -0.6)]),
                          [MoleculeSpec("p", "(A)3(B)5(A)3", "restricted", theta=2.5),
                           MoleculeSpec("s", "(S)1", "solvent")], frozen={"W": _wall(lat, "x0")})
        return lat, sysd
    if name == "1d_cylinder":
        lat = LatticeSpec(1, (20,), geometry="cylindrical", lowerbound=("mirror1",), upperbound=("surface",),
                          lambda_=1 / 6)
        st = ("A", "S", "W")
        sysd = SystemSpec(st, _chi(st, [("A", "S", 0.6), ("A", "W", -1.2)]),
                          [MoleculeSpec("p", "(A)15", "free", phibulk=0.02), MoleculeSpec("s", "(S)1", "solvent")],
                          frozen={"W": _wall(lat, "x1")})
        return lat, sysd
    if name == "1d_sphere":
        lat = LatticeSpec(1, (18,), geometry="spherical", lowerbound=("surface",), upperbound=("mirror1",),
                          lambda_=0.2, offset_first_layer=2.0)
        st = ("A", "B", "S", "W")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.9), ("A", "S", 0.1), ("B", "S", 0.9), ("A", "W", -1.0)]),
                          [MoleculeSpec("p", "(A)4(B)8", "free", phibulk=0.03), MoleculeSpec("s", "(S)1", "solvent")],
                          frozen={"W": _wall(lat, "x0")})
        return lat, sysd
    if name == "2d_flat":
        lat = LatticeSpec(2, (9, 11), lowerbound=("mirror1", "periodic"), upperbound=("mirror2", "periodic"),
                          lambda_=1 / 6)
        st = ("A", "B", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 1.0), ("A", "S", 0.4), ("B", "S", 0.7)]),
                          [MoleculeSpec("p", "(A)5(B)3", "free", phibulk=0.1), MoleculeSpec("s", "(S)1", "solvent")])
        return lat, sysd
    if name == "2d_cylinder":
        lat, sysd, _ = configs.cylinder_2d(nr=8, nz=12)
        return lat, sysd
    if name == "3d_sc_melt":
        lat, sysd, _ = configs.melt_3d(n=8, n_a=3, n_b=5, chi=0.4)
        return lat, sysd
    if name == "3d_fcc_restricted":
        lat = LatticeSpec(3, (6, 5, 7), lattice_type="fcc", lowerbound=("periodic", "mirror1", "mirror2"),
                          upperbound=("periodic", "mirror1", "mirror1"))
        st = ("A", "B", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.7), ("A", "S", 0.2), ("B", "S", 0.9)]),
                          [MoleculeSpec("p", "(A)2(B)3(A)2", "restricted", theta=20.0),
                           MoleculeSpec("s", "(S)1", "solvent")])
        return lat, sysd
    if name == "3d_hex_three_molecules":
        lat = LatticeSpec(3, (5, 6, 4), lattice_type="hex", lowerbound=("mirror1", "periodic", "mirror2"),
                          upperbound=("mirror2", "periodic", "mirror1"))
        st = ("A", "C", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "C", 0.5), ("A", "S", 0.8), ("C", "S", -0.3)]),
                          [MoleculeSpec("p", "(A)6", "free", phibulk=0.05),
                           MoleculeSpec("c", "(C)1", "restricted", theta=10.0),
                           MoleculeSpec("s", "(S)1", "solvent")])
        return lat, sysd
    if name == "3d_sc_order2_homopolymer":
        lat = LatticeSpec(3, (6, 5, 7), order=2, lowerbound=("periodic", "mirror1", "mirror2"),
                          upperbound=("periodic", "mirror2", "mirror1"))
        st = ("A", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "S", 0.6)]),
                          [MoleculeSpec("p", "(A)12", "restricted", theta=60.0), MoleculeSpec("s", "(S)1", "solvent")],
                          stiffness=1.0 / (1.0 + 4.0 * math.exp(-1.0)))
        return lat, sysd
    if name == "3d_sc_order2_copolymer":
        lat = LatticeSpec(3, (5, 6, 4), order=2, lowerbound=("mirror1", "periodic", "mirror1"),
                          upperbound=("mirror2", "periodic", "mirror1"))
        st = ("A", "B", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.9), ("A", "S", 0.3), ("B", "S", 0.5)]),
                          [MoleculeSpec("p", "(A)3(B)5", "free", phibulk=0.1), MoleculeSpec("s", "(S)1", "solvent")],
                          stiffness=0.3)
        return lat, sysd
    if name == "1d_order2_adsorption":
        lat = LatticeSpec(1, (24,), order=2, lowerbound=("surface",), upperbound=("mirror1",), lambda_=1 / 6)
        st = ("A", "S", "W")
        sysd = SystemSpec(st, _chi(st, [("A", "S", 0.5), ("A", "W", -1.0)]),
                          [MoleculeSpec("p", "(A)15", "free", phibulk=0.02), MoleculeSpec("s", "(S)1", "solvent")],
                          frozen={"W": _wall(lat, "x0")}, stiffness=0.5)
        return lat, sysd
    if name == "2d_order2_copolymer":
        # mirror2 at a lower bound copies components plainly, at an upper bound it swaps them
        lat = LatticeSpec(2, (7, 9), order=2, lowerbound=("mirror2", "mirror1"), upperbound=("mirror1", "mirror2"),
                          lambda_=1 / 6)
        st = ("A", "B", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.8), ("A", "S", 0.2), ("B", "S", 0.6)]),
                          [MoleculeSpec("p", "(A)4(B)6", "restricted", theta=12.0), MoleculeSpec("s", "(S)1", "solvent")],
                          stiffness=0.6)
        return lat, sysd
    if name == "1d_cyl_order2":
        # the axis (mirror1) at the lower bound, an adsorbing outer wall
        lat = LatticeSpec(1, (22,), geometry="cylindrical", order=2, lowerbound=("mirror1",),
                          upperbound=("surface",), lambda_=1 / 6)
        st = ("A", "S", "W")
        sysd = SystemSpec(st, _chi(st, [("A", "S", 0.5), ("A", "W", -1.0)]),
                          [MoleculeSpec("p", "(A)14", "free", phibulk=0.02), MoleculeSpec("s", "(S)1", "solvent")],
                          frozen={"W": _wall(lat, "x1")}, stiffness=0.4)
        return lat, sysd
    if name == "1d_sphere_order2":
        lat = LatticeSpec(1, (16,), geometry="spherical", order=2, lowerbound=("surface",), upperbound=("mirror1",),
                          lambda_=1 / 6, offset_first_layer=2.0)
        st = ("A", "B", "S", "W")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.9), ("A", "S", 0.1), ("B", "S", 0.9), ("A", "W", -1.0)]),
                          [MoleculeSpec("p", "(A)4(B)8", "free", phibulk=0.03), MoleculeSpec("s", "(S)1", "solvent")],
                          frozen={"W": _wall(lat, "x0")}, stiffness=0.3)
        return lat, sysd
    if name == "2d_cyl_order2":
        lat = LatticeSpec(2, (8, 9), geometry="cylindrical", order=2, lowerbound=("mirror1", "mirror2"),
                          upperbound=("mirror1", "mirror1"), lambda_=1 / 6)
        st = ("A", "B", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.8), ("A", "S", 0.2), ("B", "S", 0.6)]),
                          [MoleculeSpec("p", "(A)4(B)6", "restricted", theta=40.0), MoleculeSpec("s", "(S)1", "solvent")],
                          stiffness=0.6)
        return lat, sysd
    if name == "3d_sc_melt_n24":
        # long enough for a checkpointed store smaller than the full one
        lat, sysd, _ = configs.melt_3d(n=8, n_a=10, n_b=14, chi=0.4)
        return lat, sysd
    if name == "3d_sc_mirrors_blocked":
        # large enough for the several-ranks-per-launch kernels; mirror1, mirror2 and
        # periodic sides, a full-store copolymer, an odd half-stored homopolymer
        lat = LatticeSpec(3, (9, 7, 6), lowerbound=("mirror1", "mirror2", "periodic"),
                          upperbound=("mirror2", "mirror1", "periodic"))
        st = ("A", "B", "S")
        sysd = SystemSpec(st, _chi(st, [("A", "B", 0.8), ("A", "S", 0.4), ("B", "S", 0.6)]),
                          [MoleculeSpec("p", "(A)3(B)6(A)2", "restricted", theta=40.0),
                           MoleculeSpec("h", "(B)9", "free", phibulk=0.1), MoleculeSpec("s", "(S)1", "solvent")])
        return lat, sysd
    raise KeyError(name)


NAMES = ["1d_flat_adsorption", "1d_flat_odd_mirror2", "1d_cylinder", "1d_sphere", "2d_flat", "2d_cylinder",
         "3d_sc_melt", "3d_fcc_restricted", "3d_hex_three_molecules", "3d_sc_order2_homopolymer",
         "3d_sc_order2_copolymer", "1d_order2_adsorption", "2d_order2_copolymer", "3d_sc_mirrors_blocked",
         "3d_sc_melt_n24", "1d_cyl_order2", "1d_sphere_order2", "2d_cyl_order2"]


def start(lat, sysd, seed=11, sigma=0.2):
    return np.random.default_rng(seed).normal(0.0, sigma, sysd.nvar(lat))


def store_ranks(lat, sysd):
    """Forward ranks the backward pass needs for the largest molecule: N for a
    second-order or asymmetric chain, N/2 for a palindrome, 0 for a monomer
    (SF_Homopol1stO.cpp:555, SF_Copol1stO.cpp:49-57)."""
    best = 0
    for seq in sysd.sequences():
        n = len(seq)
        if n == 1:
            r = 0
        elif lat.order == 2 or seq != seq[::-1]:
            r = n
        else:
            r = n // 2
        best = max(best, r)
    return best
