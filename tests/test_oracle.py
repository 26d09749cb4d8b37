"""Pins the CPU oracle (the restatement of the reference's algorithms) against
independent constructions, invariants, closed forms and the reference's own
data file. CPU only."""
import math
import os

import numpy as np
import pytest

from field_amd import configs
from field_amd.output import read_vtk, vtk_text
from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
_PAD = {"mirror1": "edge", "mirror2": "reflect", "periodic": "wrap", "surface": "constant"}


def _padded(G, lat):
    """Halo values by numpy padding: edge = mirror1, reflect = mirror2, wrap = periodic,
    zero = surface (a propagator vanishes on a frozen surface)."""
    a = G
    for ax in range(lat.gradients):
        lo, up = lat.lowerbound[ax], lat.upperbound[ax]
        width = [(0, 0)] * a.ndim
        if "periodic" in (lo, up):
            width[ax] = (1, 1)
            a = np.pad(a, width, mode="wrap")
            continue
        width[ax] = (1, 0)
        a = np.pad(a, width, mode=_PAD[lo]) if _PAD[lo] != "reflect" else _reflect(a, ax, 0)
        width[ax] = (0, 1)
        a = np.pad(a, width, mode=_PAD[up]) if _PAD[up] != "reflect" else _reflect(a, ax, 1)
    return a


def _reflect(a, ax, side):
    idx = [slice(None)] * a.ndim
    if side == 0:
        idx[ax] = slice(1, 2)
        return np.concatenate([a[tuple(idx)], a], axis=ax)
    n = a.shape[ax]
    idx[ax] = slice(n - 2, n - 1)
    return np.concatenate([a, a[tuple(idx)]], axis=ax)


def _shift(P, d):
    """interior view of the padded array P shifted by the offset vector d."""
    idx = tuple(slice(1 + k, P.shape[i] - 1 + k) for i, k in enumerate(d))
    return P[idx]


def numpy_propagate(lat, G, g):
    """Independent stencil: Lat1DFlat1stO.cpp:46, Lat2DFlat1stO.cpp:44 and the
    3-D neighbour lists (1/6 on the six faces, 1/12 on the FCC / HEX bonds)."""
    n = tuple(lat.n_layers[: lat.gradients])
    P = _padded(G.reshape(n), lat)
    l1 = lat.lambda_
    if lat.gradients == 1:
        b = _shift(P, (0,))
        out = b + l1 * (_shift(P, (-1,)) - 2 * b + _shift(P, (1,)))
    elif lat.gradients == 2:
        out = (1 - 4 * l1) * _shift(P, (0, 0)) + l1 * sum(
            _shift(P, d) for d in ((1, 0), (-1, 0), (0, 1), (0, -1)))
    else:
        if lat.lattice_type == "fcc":
            offs = [(a, b, 0) for a in (-1, 1) for b in (-1, 1)] + \
                   [(0, a, b) for a in (-1, 1) for b in (-1, 1)] + \
                   [(a, 0, b) for a in (-1, 1) for b in (-1, 1)]
            w = 1.0 / 12
        elif lat.lattice_type == "hex":
            offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                    (-1, 1, 0), (1, -1, 0), (0, 1, -1), (1, 0, -1), (-1, 0, 1), (0, -1, 1)]
            w = 1.0 / 12
        else:
            offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
            w = 1.0 / 6
        out = w * sum(_shift(P, d) for d in offs)
    return (g.reshape(n) * out).reshape(-1)


def _embed(lat, interior):
    full = np.zeros(lat.M)
    full[lat.interior_mask()] = interior
    return full


CASES = [
    LatticeSpec(1, (9,), lowerbound=("mirror1",), upperbound=("mirror2",), lambda_=0.25),
    LatticeSpec(1, (9,), lowerbound=("surface",), upperbound=("mirror1",), lambda_=1 / 6),
    LatticeSpec(1, (8,), lowerbound=("periodic",), upperbound=("periodic",), lambda_=0.3),
    LatticeSpec(2, (5, 7), lowerbound=("mirror1", "periodic"), upperbound=("surface", "periodic"), lambda_=0.2),
    LatticeSpec(2, (6, 4), lowerbound=("mirror2", "mirror1"), upperbound=("mirror1", "mirror2"), lambda_=1 / 6),
    LatticeSpec(3, (4, 5, 6), lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3),
    LatticeSpec(3, (4, 5, 6), lowerbound=("mirror1", "mirror2", "periodic"),
                upperbound=("mirror2", "mirror1", "periodic")),
    LatticeSpec(3, (5, 4, 6), lattice_type="fcc", lowerbound=("periodic", "mirror1", "mirror2"),
                upperbound=("periodic", "mirror1", "mirror1")),
    LatticeSpec(3, (5, 4, 6), lattice_type="hex", lowerbound=("mirror1", "periodic", "mirror2"),
                upperbound=("mirror2", "periodic", "mirror1")),
]


@pytest.mark.parametrize("lat", CASES, ids=lambda l: f"{l.gradients}d-{l.lattice_type}-{l.lowerbound[0]}")
def test_propagator_matches_independent_stencil(orc, lat):
    rng = np.random.default_rng(1)
    V = int(np.prod(lat.n_layers[: lat.gradients]))
    G = rng.random(V) + 0.1
    g = rng.random(V) + 0.5
    L = orc.Lattice(lat)
    out = L.propagate(_embed(lat, G), _embed(lat, g))[lat.interior_mask()]
    ref = numpy_propagate(lat, G, g)
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=0)


def _shell_volumes(geometry, n):
    r = np.arange(0, n + 2, dtype=float)  # outer radius of layer z (offset 0, adjustment 1)
    if geometry == "cylindrical":
        V = math.pi * r ** 2
    else:
        V = 4.0 / 3.0 * math.pi * r ** 3
    return np.diff(V)  # layers 1..n+1


@pytest.mark.parametrize("geometry", ["cylindrical", "spherical"])
def test_curved_walks_conserve_weight(orc, geometry):
    """With g = 1 and mirror bounds the site-weighted sum of a propagator is
    conserved (L(z) lambda+(z) = L(z+1) lambda-(z+1), Lat1DCylinder.cpp:380-390)."""
    n = 12
    lat = LatticeSpec(1, (n,), geometry=geometry, lowerbound=("mirror1",), upperbound=("mirror1",),
                      lambda_=1 / 6)
    L = orc.Lattice(lat)
    w = _shell_volumes(geometry, n)[:n]
    G = _embed(lat, np.random.default_rng(2).random(n))
    one = _embed(lat, np.ones(n))
    before = np.sum(w * G[1:-1])
    for _ in range(5):
        G = L.propagate(G, one)
    assert abs(np.sum(w * G[1:-1]) - before) < 1e-12 * before


def test_cylinder_2d_conserves_weight(orc):
    nx, ny = 6, 5
    lat = LatticeSpec(2, (nx, ny), geometry="cylindrical", lowerbound=("mirror1", "periodic"),
                      upperbound=("mirror1", "periodic"), lambda_=1 / 6)
    L = orc.Lattice(lat)
    w = np.repeat(_shell_volumes("cylindrical", nx)[:nx], ny)
    G = _embed(lat, np.random.default_rng(3).random(nx * ny))
    one = _embed(lat, np.ones(nx * ny))
    m = lat.interior_mask()
    before = np.sum(w * G[m])
    for _ in range(4):
        G = L.propagate(G, one)
    assert abs(np.sum(w * G[m]) - before) < 1e-12 * before


def _periodic_system(lat, g_value, N=7):
    sysd = SystemSpec(states=("A", "S"), chi=np.zeros((2, 2)),
                      molecules=[MoleculeSpec("p", f"(A){N}", "restricted", theta=3.0),
                                 MoleculeSpec("s", "(S)1", "solvent")])
    return sysd


PERIODIC = [c for c in CASES if set(c.lowerbound[: c.gradients]) == {"periodic"}] + [
    LatticeSpec(2, (5, 7), lowerbound=("periodic", "periodic"), upperbound=("periodic", "periodic"), lambda_=0.2),
    LatticeSpec(3, (4, 3, 5), lattice_type="fcc", lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3),
    LatticeSpec(3, (4, 3, 5), lattice_type="hex", lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3),
]


@pytest.mark.parametrize("lat", PERIODIC, ids=lambda l: f"{l.gradients}d-{l.lattice_type}")
def test_constant_field_closed_form(orc, lat):
    """g = e^-u constant on a periodic lattice: G_s = g^s, ln G_N = ln V + N ln g."""
    N, u = 7, 0.3
    sysd = _periodic_system(lat, math.exp(-u), N)
    S = orc.System(lat, sysd)
    x = np.full(S.nvar, u)
    S.compute_phi(x)
    V = int(np.prod(lat.n_layers[: lat.gradients]))
    assert abs(S.lngn(0) - (math.log(V) - N * u)) < 1e-12
    # restricted chain spreads uniformly: phi = theta / V
    phiA = S.phi(0)[lat.interior_mask()]
    np.testing.assert_allclose(phiA, 3.0 / V, rtol=1e-12)


def _two_state_system(lat, comp_states, comp, freedom="free", **kw):
    chi = np.zeros((3, 3))
    chi[0, 2] = chi[2, 0] = chi[1, 2] = chi[2, 1] = 0.7  # A and A' identical contacts
    mols = [MoleculeSpec("p", comp, freedom, **kw), MoleculeSpec("s", "(S)1", "solvent")]
    return SystemSpec(states=comp_states, chi=chi, molecules=mols)


@pytest.mark.parametrize("N", [10, 11])
def test_half_store_equals_full_sweep(orc, N):
    """A homopolymer takes the half-stored path (SF_Homopol1stO.cpp:552-622); the
    same chain written as two chemically identical blocks takes the full sweep
    (SF_Copol1stO.cpp:1186-1281). Their densities must agree."""
    lat = LatticeSpec(1, (15,), lowerbound=("mirror1",), upperbound=("mirror1",), lambda_=1 / 6)
    rng = np.random.default_rng(4)
    half = N // 2
    homo = _two_state_system(lat, ("A", "B", "S"), f"(A){N}", phibulk=0.05)
    split = _two_state_system(lat, ("A", "B", "S"), f"(A){half}(B){N - half}", phibulk=0.05)
    S1, S2 = orc.System(lat, homo), orc.System(lat, split)
    nsite = 15
    # one iteration variable per distinct chi row: A and B share one
    x = rng.normal(0, 0.3, S1.nvar)
    f1 = S1.residual(x)
    f2 = S2.residual(x)
    np.testing.assert_allclose(S1.phi(0), S2.phi(0) + S2.phi(1), rtol=1e-12, atol=1e-15)
    assert abs(S1.lngn(0) - S2.lngn(0)) < 1e-12
    np.testing.assert_allclose(f1, f2, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("geometry,upper", [("cylindrical", "mirror1"), ("spherical", "mirror1"),
                                             ("flat", "mirror2")])
def test_restricted_mass_balance(orc, geometry, upper):
    """sum_z w(z) phi(z) = theta for a restricted chain, with w the ln G_N site
    weights. Exact wherever diag(w) times the transfer matrix is symmetric
    (not for a curved lattice with a mirror2 bound, where the reference's
    lambda+(n) != lambda-(n) at the mirror layer)."""
    n = 20
    lat = LatticeSpec(1, (n,), geometry=geometry, lowerbound=("mirror1",), upperbound=(upper,),
                      lambda_=1 / 6)
    sysd = SystemSpec(states=("A", "B", "S"), chi=np.array([[0, 1, .5], [1, 0, .2], [.5, .2, 0]]),
                      molecules=[MoleculeSpec("p", "(A)4(B)6", "restricted", theta=25.0),
                                 MoleculeSpec("s", "(S)1", "solvent")])
    S = orc.System(lat, sysd)
    x = np.random.default_rng(5).normal(0, 0.2, S.nvar)
    S.compute_phi(x)
    w = _shell_volumes(geometry, n)[:n].copy() if geometry != "flat" else np.ones(n)
    if upper == "mirror2":
        w[-1] *= 0.5  # the mirror plane halves the last layer
    tot = np.sum(w * (S.phi(0) + S.phi(1))[1:-1])
    assert abs(tot - 25.0) < 1e-11 * 25


def test_monomer_is_phibulk_times_g(orc):
    lat, sysd, x = configs.adsorption_1d(n_layers=12, N=6)
    S = orc.System(lat, sysd)
    x = np.random.default_rng(6).normal(0, 0.2, S.nvar)
    S.compute_phi(x)
    m = lat.interior_mask()
    np.testing.assert_allclose(S.phi(1)[m], S.phibulk(1) * S.swf(1)[m], rtol=1e-14)
    assert abs(S.phibulk(1) - (1 - 0.01)) < 1e-15


def test_homogeneous_bulk_is_a_solution(orc):
    """No surfaces, periodic: x = 0 gives SWF = 1, phi = phi_b, phi_tot = 1 and
    ChemInt = ChemIntBulk, so every residual entry vanishes (SF_Solve.cpp:583)."""
    lat = LatticeSpec(3, (4, 4, 4), lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    chi = np.array([[0, 1.1, 0.4], [1.1, 0, 0.9], [0.4, 0.9, 0]])
    sysd = SystemSpec(states=("A", "B", "S"), chi=chi,
                      molecules=[MoleculeSpec("p", "(A)3(B)5", "free", phibulk=0.2),
                                 MoleculeSpec("s", "(S)1", "solvent")])
    S = orc.System(lat, sysd)
    f = S.residual(np.zeros(S.nvar))
    assert np.abs(f).max() < 1e-14
    m = lat.interior_mask()
    np.testing.assert_allclose(S.phi(0)[m] + S.phi(1)[m], 0.2, rtol=1e-14)


def test_free_energy_flory_huggins(orc):
    """Homogeneous bulk: F/V = sum_m (phi_m/N_m) ln phi_m + chi phi_A phi_S."""
    lat = LatticeSpec(1, (10,), lowerbound=("periodic",), upperbound=("periodic",), lambda_=1 / 6)
    chi = 0.8
    sysd = SystemSpec(states=("A", "S"), chi=np.array([[0, chi], [chi, 0]]),
                      molecules=[MoleculeSpec("p", "(A)20", "free", phibulk=0.3),
                                 MoleculeSpec("s", "(S)1", "solvent")])
    S = orc.System(lat, sysd)
    S.residual(np.zeros(S.nvar))
    F = S.free_energy()
    fh = 0.3 / 20 * math.log(0.3) + 0.7 * math.log(0.7) + chi * 0.3 * 0.7
    assert abs(F - 10 * fh) < 1e-12


def test_melt_free_energy_at_zero_field(orc):
    """u = 0, n N = V: F = chi V f (1 - f) (MPS.cpp:253-261)."""
    lat, sysd, _ = configs.melt_3d(n=6, n_a=3, n_b=5, chi=0.4)
    V = 216
    sysd.molecules[0].theta = float(V)
    S = orc.System(lat, sysd)
    S.residual(np.zeros(S.nvar))
    assert abs(S.free_energy() - 0.4 * V * (3 / 8) * (5 / 8)) < 1e-10


def test_mps_restatement_matches_general_system(orc):
    """Two independent restatements (MPS.cpp literal, and the SF molecule /
    residual chain in melt form) give the same residual, G_N and F."""
    n, na, nb = 6, 3, 5
    lat, sysd, x = configs.melt_3d(n=n, n_a=na, n_b=nb, chi=0.4, sigma=0.1, seed=7)
    S = orc.System(lat, sysd)
    P = orc.MPS(n, na, nb, 0.4, n ** 3 / (na + nb))
    f1, f2 = S.residual(x), P.residual(x)
    np.testing.assert_allclose(f1, f2, rtol=0, atol=1e-13)
    assert abs(S.lngn(0) - math.log(P.gn())) < 1e-13
    assert abs(S.free_energy() - P.helmholtz(x)) < 1e-10 * abs(P.helmholtz(x))


def test_mps_diis_trajectory_matches_general(orc):
    n, na, nb = 6, 3, 5
    lat, sysd, x = configs.melt_3d(n=n, n_a=na, n_b=nb, chi=0.4, sigma=0.1, seed=7)
    S = orc.System(lat, sysd)
    P = orc.MPS(n, na, nb, 0.4, n ** 3 / (na + nb))
    x1, it1, e1, rc1 = P.scf(x, tol=1e-8)
    x2, it2, e2, rc2 = S.solve_diis(x, tol=1e-8)
    assert rc1 == 0 and rc2 == 0 and it1 == it2
    np.testing.assert_allclose(x1, x2, atol=1e-10)


def test_vtk_fixture_symmetric_melt(orc):
    """tests/golden/test.vtk is the reference's VTK output of a 10^3 symmetric
    melt: phi_A = 0.5 everywhere. The symmetric A5B5 melt at u = 0 reproduces
    it byte for byte through the MPS.cpp layout."""
    P = orc.MPS(10, 5, 5, 0.5, 100.0)
    M = P.M
    P.residual(np.zeros(2 * M))
    phiA = P.phi()[:M]
    text = vtk_text(phiA, (10, 10, 10))
    with open(os.path.join(GOLDEN, "test.vtk")) as f:
        assert text == f.read()
    dims, vals = read_vtk(os.path.join(GOLDEN, "test.vtk"))
    assert dims == (10, 10, 10) and np.all(vals == 0.5)


def test_diis_coefficients_solve_a_c_equals_one(orc):
    rng = np.random.default_rng(8)
    B = rng.normal(size=(6, 6))
    A = B @ B.T + np.eye(6)
    A[0, 0] = 0.0  # the first DIIS step leaves A[0][0] unset (MPS.cpp:234-235)
    np.testing.assert_allclose(orc.diis_coeffs(A), np.linalg.solve(A, np.ones(6)), rtol=1e-10)


def test_contact_average_3d_sc_is_neighbour_mean(orc):
    lat = LatticeSpec(3, (4, 5, 3), lowerbound=("periodic", "mirror1", "mirror2"),
                      upperbound=("periodic", "mirror2", "mirror1"))
    L = orc.Lattice(lat)
    P = np.random.default_rng(9).random(60)
    out = L.sides(_embed(lat, P))[lat.interior_mask()]
    Pp = _padded(P.reshape(4, 5, 3), lat)
    ref = sum(_shift(Pp, d) for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))) / 6
    np.testing.assert_allclose(out, ref.reshape(-1), rtol=1e-14)


@pytest.mark.parametrize("bounds", [("periodic",) * 3, ("mirror1", "mirror2", "periodic")])
def test_order2_restricted_mass_balance(orc, bounds):
    """Second order: sum_r phi(r) = theta needs the backward sweep
    (PropagateB) to be the adjoint of the forward one (PropagateF) including
    the mirror component swaps (Lat3D.cpp:1276-1290)."""
    lat = LatticeSpec(3, (5, 4, 6), order=2, lowerbound=bounds, upperbound=bounds[::-1])
    sysd = SystemSpec(("A", "B", "S"), np.array([[0, 0.8, 0.3], [0.8, 0, 0.5], [0.3, 0.5, 0]]),
                      [MoleculeSpec("p", "(A)4(B)7", "restricted", theta=30.0),
                       MoleculeSpec("s", "(S)1", "solvent")], stiffness=0.45)
    S = orc.System(lat, sysd)
    x = np.random.default_rng(17).normal(0, 0.3, S.nvar)
    S.compute_phi(x)
    m = lat.interior_mask()
    w = np.ones(lat.M)
    MX, MY, MZ = lat.dims
    w3 = w.reshape(MX, MY, MZ)
    n = lat.n_layers
    for ax in range(3):
        lo, up = lat.lowerbound[ax], lat.upperbound[ax]
        if "periodic" in (lo, up):
            continue
        for bound, layer in ((lo, 1), (up, n[ax])):
            if bound == "mirror2":  # the mirror plane runs through this layer: half of it is inside
                sl = [slice(None)] * 3
                sl[ax] = layer
                w3[tuple(sl)] *= 0.5
    tot = np.sum((w * (S.phi(0) + S.phi(1)))[m])
    assert abs(tot - 30.0) < 1e-11 * 30


def test_order2_uniform_field_closed_form(orc):
    lat = LatticeSpec(3, (4, 5, 3), order=2, lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    sysd = SystemSpec(("A", "S"), np.array([[0, 0.6], [0.6, 0]]),
                      [MoleculeSpec("p", "(A)9", "restricted", theta=12.0), MoleculeSpec("s", "(S)1", "solvent")],
                      stiffness=0.404611)
    S = orc.System(lat, sysd)
    S.compute_phi(np.full(S.nvar, 0.25))
    assert abs(S.lngn(0) - (math.log(60) - 9 * 0.25)) < 1e-12
    np.testing.assert_allclose(S.phi(0)[lat.interior_mask()], 12.0 / 60, rtol=1e-12)


def test_torch_melt_reference_matches_oracle(orc):
    """The differentiable torch restatement used as the reference gradient
    reproduces the oracle's melt residual."""
    import torch

    import torch_ref

    n, na, nb = 5, 3, 4
    lat, sysd, x = configs.melt_3d(n=n, n_a=na, n_b=nb, chi=0.4, sigma=0.2, seed=9)
    fo = orc.System(lat, sysd).residual(x)
    u = torch_ref.to_interior(torch.from_numpy(x), n)
    f = torch_ref.melt_residual(u, [0] * na + [1] * nb, [[0, 0.4], [0.4, 0]], float(n ** 3), (0.0, 0.0))
    np.testing.assert_allclose(torch_ref.to_padded(f, n).numpy(), fo, rtol=0, atol=1e-12)


@pytest.mark.parametrize("grad,bx", [(1, ("mirror1", "mirror1")), (1, ("periodic", "periodic")),
                                     (2, ("mirror1", "mirror1")), (2, ("periodic", "periodic"))])
def test_order2_lumped_forms_match_six_components(orc, grad, bx):
    """Lat1D2ndO's three and Lat2D2ndO's five lumped components, restated
    from their own formulas, reproduce the six-component 3-D sweep on the same
    lattice with the missing axes one periodic layer thick (homogeneous)."""
    n = (7, 6)[:grad]
    low, up = (bx[0], "periodic")[:grad], (bx[1], "periodic")[:grad]
    lat = LatticeSpec(grad, n, order=2, lowerbound=low, upperbound=up, lambda_=1 / 6)
    lat3 = LatticeSpec(3, tuple(n) + (1,) * (3 - grad), order=2, lowerbound=low + ("periodic",) * (3 - grad),
                       upperbound=up + ("periodic",) * (3 - grad))
    sysd = SystemSpec(("A", "B", "S"), np.array([[0, 0.8, 0.3], [0.8, 0, 0.5], [0.3, 0.5, 0]]),
                      [MoleculeSpec("p", "(A)4(B)5", "restricted", theta=9.0),
                       MoleculeSpec("q", "(B)3", "free", phibulk=0.1), MoleculeSpec("s", "(S)1", "solvent")],
                      stiffness=0.45)
    A, B = orc.System(lat, sysd), orc.System(lat3, sysd)
    assert A.nvar == B.nvar
    x = np.random.default_rng(31).normal(0, 0.3, A.nvar)
    fa, fb = A.residual(x), B.residual(x)
    np.testing.assert_allclose(fa, fb, rtol=0, atol=1e-12)
    for m in range(2):
        assert abs(A.lngn(m) - B.lngn(m)) < 1e-12


@pytest.mark.parametrize("grad", [1, 2])
def test_order2_lumped_mass_balance_and_closed_form(orc, grad):
    n = (9, 8)[:grad]
    lat = LatticeSpec(grad, n, order=2, lowerbound=("mirror1",) * grad, upperbound=("mirror1",) * grad,
                      lambda_=1 / 6)
    sysd = SystemSpec(("A", "S"), np.array([[0, 0.6], [0.6, 0]]),
                      [MoleculeSpec("p", "(A)11", "restricted", theta=7.0), MoleculeSpec("s", "(S)1", "solvent")],
                      stiffness=0.3)
    S = orc.System(lat, sysd)
    m = lat.interior_mask()
    S.compute_phi(np.random.default_rng(5).normal(0, 0.3, S.nvar))
    assert abs(np.sum(S.phi(0)[m]) - 7.0) < 1e-11
    # uniform field: every component is g^N, ln G_N = ln(V g^N)
    S.compute_phi(np.full(S.nvar, 0.2))
    V = int(np.prod(n))
    assert abs(S.lngn(0) - (math.log(V) - 11 * 0.2)) < 1e-12
    np.testing.assert_allclose(S.phi(0)[m], 7.0 / V, rtol=1e-12)


# ---- curved second-order lattices ------------------------------------------

def _ring_volumes(geometry, n, offset=0.0):
    x = np.arange(1, n + 1, dtype=float) + offset
    if geometry == "cylinder":
        return np.pi * (x ** 2 - (x - 1) ** 2)
    return 4 / 3 * np.pi * (x ** 3 - (x - 1) ** 3)


def _restricted_mass(orc, lat, geometry, theta=300.0):
    sysd = SystemSpec(("A", "S"), np.array([[0, 0.4], [0.4, 0]]),
                      [MoleculeSpec("p", "(A)12", "restricted", theta=theta), MoleculeSpec("s", "(S)1", "solvent")],
                      stiffness=0.5)
    O = orc.System(lat, sysd)
    O.residual(np.random.default_rng(1).normal(0, 0.1, O.nvar))
    MX, MY, MZ = lat.dims
    w = np.zeros((MX, MY, MZ))
    Lx = _ring_volumes(geometry, lat.n_layers[0], lat.offset_first_layer)
    if lat.gradients == 1:
        w[1:-1, 0, 0] = Lx
    else:
        w[1:-1, 1:-1, 0] = Lx[:, None]
    return (w.reshape(-1) * O.phi(0)).sum() / theta - 1


@pytest.mark.parametrize("geometry", ["cylinder", "sphere"])
@pytest.mark.parametrize("lower, offset", [("mirror1", 0.0), ("surface", 4.0)])
def test_curved_order2_restricted_mass_balance(orc, geometry, lower, offset):
    """Lat1DCyl2ndO / Lat1DSphere2ndO: the forward and backward sweeps are
    adjoint in the layer-volume inner product, so a restricted chain's
    segments sum to theta exactly (an index or table slip breaks it). A
    mirror2 bound on a curved lattice does not conserve even in first order
    (its image layer has another volume), so the bounds here are mirror1 and
    surface."""
    lat = LatticeSpec(1, (30,), geometry=geometry, lowerbound=(lower,), upperbound=("mirror1",), order=2,
                      offset_first_layer=offset)
    assert abs(_restricted_mass(orc, lat, geometry)) < 1e-12


def test_cylinder_2d_order2_mass_balance_and_backward_quirk(orc):
    """Lat2DCyl2ndO.cpp:223-227: PropagateB computes the outward split and then
    reads the plain outward neighbour. With that split applied (a test switch
    in the oracle) the 2-D cylinder conserves mass exactly, so everything else
    in the restatement is pinned; as the reference has it, mass drifts."""
    lat = LatticeSpec(2, (12, 10), geometry="cylinder", lowerbound=("mirror1", "mirror1"),
                      upperbound=("mirror1", "periodic"), order=2)
    try:
        orc.set_cyl2d_split_outward(True)
        assert abs(_restricted_mass(orc, lat, "cylinder")) < 1e-12
    finally:
        orc.set_cyl2d_split_outward(False)
    assert abs(_restricted_mass(orc, lat, "cylinder")) > 1e-4


@pytest.mark.parametrize("backward", [False, True])
def test_cylinder_2d_order2_reduces_to_1d(orc, backward):
    """An axially uniform field on the 2-D cylinder: one step of the five
    component form equals the three component Lat1DCyl2ndO step (the +-y and
    angular components equal the 1-D in-layer component)."""
    n, ny = 9, 4
    S = 0.3
    l1 = orc.Lattice(LatticeSpec(1, (n,), geometry="cylinder", lowerbound=("mirror1",), upperbound=("mirror1",),
                                 order=2))
    l2 = orc.Lattice(LatticeSpec(2, (n, ny), geometry="cylinder", lowerbound=("mirror1", "periodic"),
                                 upperbound=("mirror1", "periodic"), order=2))
    rng = np.random.default_rng(5)
    M1, M2 = l1.M, l2.M
    prof = rng.uniform(0.5, 1.5, (3, n))  # radial, in-layer, inward components per ring
    g1 = np.zeros(M1)
    g1[1:-1] = rng.uniform(0.8, 1.2, n)
    G1 = np.zeros(3 * M1)
    for k in range(3):
        G1[k * M1 + 1:k * M1 + 1 + n] = prof[k]
    MY = ny + 2
    g2 = np.zeros((n + 2, MY))
    g2[1:-1, 1:-1] = g1[1:-1, None]
    G2 = np.zeros((5, n + 2, MY))
    for k5, k3 in enumerate([0, 1, 1, 1, 2]):
        G2[k5, 1:-1, 1:-1] = prof[k3][:, None]
    try:
        orc.set_cyl2d_split_outward(True)  # the 1-D form splits both radial steps
        out2 = l2.propagate2(G2.reshape(-1), g2.reshape(-1), S, backward).reshape(5, n + 2, MY)
    finally:
        orc.set_cyl2d_split_outward(False)
    out1 = l1.propagate2(G1, g1, S, backward).reshape(3, M1)
    for k5, k3 in enumerate([0, 1, 1, 1, 2]):
        np.testing.assert_allclose(out2[k5, 1:-1, 1:-1], np.repeat(out1[k3, 1:-1, None], ny, axis=1),
                                   rtol=1e-13, atol=0)


# ---- force ensemble, normalisations, second-order connect (per-rank primitives)

def _walk_moments(P, axis_coord):
    m = P.sum()
    mean = (P * axis_coord).sum() / m
    var = (P * (axis_coord - mean) ** 2).sum() / m
    return m, mean, var


@pytest.mark.parametrize("f", [0.0, 0.3, -0.7])
def test_force_1d_drift_closed_form(orc, f):
    """A walk under a force on a flat 1-D lattice, g = 1, from one layer, far
    from both walls: per step the mass is kept, the mean moves by f1 - f_1 and
    the variance grows by f1 + f_1 - (f1 - f_1)^2, with f1 = l1 e^f / n,
    f_1 = l1 e^-f / n, n = l1 (e^f + e^-f - 2) + 1 (Lat1DFlat1stO.cpp:110-130)."""
    n, l1, steps, z0 = 200, 0.25, 40, 100
    lat = LatticeSpec(1, (n,), lowerbound=("mirror1",), upperbound=("mirror1",), lambda_=l1)
    L = orc.Lattice(lat)
    G = np.zeros(lat.M)
    G[z0] = 1.0
    g = np.ones(lat.M)
    for _ in range(steps):
        G = L.propagate_force(G, g, f)
    norm = l1 * (math.exp(f) + math.exp(-f) - 2) + 1
    f1, f_1 = l1 * math.exp(f) / norm, l1 * math.exp(-f) / norm
    z = np.arange(lat.M, dtype=float)
    inner = slice(1, n + 1)
    m, mean, var = _walk_moments(G[inner], z[inner])
    assert abs(m - 1) < 1e-12
    assert abs(mean - (z0 + steps * (f1 - f_1))) < 1e-10
    assert abs(var - steps * (f1 + f_1 - (f1 - f_1) ** 2)) < 1e-9
    if f == 0.0:  # no force: the plain propagator
        G0 = np.zeros(lat.M)
        G0[z0] = 1.0
        np.testing.assert_allclose(L.propagate_force(G0, g, 0.0), L.propagate(G0, g), rtol=0, atol=1e-15)


@pytest.mark.parametrize("kind, n_dn, n_up, n_in, norm", [
    ("standard", 1, 1, 4, lambda ef: 4 + ef + 1 / ef),
    ("fcc", 4, 4, 4, lambda ef: 4 + 4 * ef + 4 / ef),
    ("hex", 3, 3, 6, lambda ef: 4 + 4 * ef + 4 / ef)])
def test_force_3d_drift_closed_form(orc, kind, n_dn, n_up, n_in, norm):
    """A walk under a force along z on a periodic 3-D lattice, g = 1, from one
    site: each step multiplies the mass by (n_in + n_dn e^f + n_up e^-f) fnorm
    and moves the mean z by (n_dn e^f - n_up e^-f) / (n_in + n_dn e^f + n_up
    e^-f), n_dn / n_up the neighbours one layer below / above, fnorm the
    reference's normaliser. Simple cubic and FCC keep the mass; HEX, whose
    normaliser the reference takes from FCC, does not."""
    f = 0.4
    ef = math.exp(f)
    n, steps = (12, 12, 40), 8
    lat = LatticeSpec(3, n, lattice_type=kind, lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    L = orc.Lattice(lat)
    G = np.zeros(lat.M)
    MX, MY, MZ = lat.dims
    r0 = (6 * MY + 6) * MZ + 20
    G[r0] = 1.0
    g = np.ones(lat.M)
    for _ in range(steps):
        G = L.propagate_force(G, g, f)
    w = n_in + n_dn * ef + n_up / ef
    zc = np.tile(np.arange(MZ, dtype=float), MX * MY)
    inner = lat.interior_mask().reshape(-1)
    m, mean, _ = _walk_moments(G[inner], zc[inner])
    assert abs(m - (w / norm(ef)) ** steps) < 1e-12
    assert abs(mean - (20 + steps * (n_dn * ef - n_up / ef) / w)) < 1e-10


def test_norm_phi_restricted_gives_theta(orc):
    """NormPhiRestr of phi = Gi puts exactly C into sum_z w(z) phi(z) (the
    ComputeLnGN weights); second order with phi = sum_k Gi^k likewise, and the
    per-axis parts scale with it. NormPhiFree is a plain scale."""
    rng = np.random.default_rng(3)
    lat = LatticeSpec(1, (30,), geometry="cylindrical", lowerbound=("mirror1",), upperbound=("mirror2",),
                      lambda_=1 / 6)
    L = orc.Lattice(lat)
    Gi = rng.random(lat.M) + 0.1
    phi, _ = L.norm_phi_restr(Gi, None, Gi, 2.5)
    assert abs(math.exp(L.lngn(phi)) - 2.5) < 1e-12
    lat2 = LatticeSpec(3, (5, 6, 4), order=2, lowerbound=("mirror1", "periodic", "mirror2"),
                       upperbound=("mirror2", "periodic", "mirror1"))
    L2 = orc.Lattice(lat2)
    Gi6 = rng.random(6 * lat2.M) + 0.1
    tot = Gi6.reshape(6, -1).sum(axis=0)
    PHI = rng.random(3 * lat2.M)
    phi2, PHI2 = L2.norm_phi_restr(tot, PHI, Gi6, 7.0)
    # six equal components phi2 / 6: the six-component ComputeLnGN sums w phi2
    assert abs(6 * math.exp(L2.lngn2(np.tile(phi2, 6) / 6)) - 7.0) < 1e-11
    np.testing.assert_allclose(PHI2 / PHI, phi2[0] / tot[0], rtol=1e-14)
    p3, P3 = L.norm_phi_free(Gi, None, 0.5)
    np.testing.assert_array_equal(p3, Gi * 0.5)


def test_connect2_axis_parts_add_up(orc):
    """ConnectG with PHI: the three per-axis parts add up to the total added."""
    rng = np.random.default_rng(4)
    lat = LatticeSpec(3, (4, 5, 3), order=2, lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    L = orc.Lattice(lat)
    A, B = rng.random(6 * lat.M), rng.random(6 * lat.M)
    out0 = rng.random(lat.M)
    out, OUT = L.connect2(A, B, out0, np.zeros(3 * lat.M))
    np.testing.assert_allclose(OUT.reshape(3, -1).sum(axis=0), out - out0, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(out - out0, (A * B).reshape(6, -1).sum(axis=0), rtol=1e-14)
