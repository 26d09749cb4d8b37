"""Chemical potentials and the grand potential of the oracle
(SF_MolList1stO.cpp:389-452, :637-647, :1001-1100), pinned by what the
thermodynamics fixes: the Flory-Huggins exchange chemical potentials of a
homogeneous polymer solution, a vanishing grand potential there, and the Gibbs
adsorption equation dOmega = -sum_m Gamma_m dmu_m across converged wall
profiles."""
import math

import numpy as np
import pytest

from field_amd import configs
from field_amd.spec import LatticeSpec, MoleculeSpec, SystemSpec


def _solution(N, phi, chi, restricted=False):
    """A_N in a monomeric solvent on a periodic 6^3 lattice: u = 0 is the
    homogeneous solution."""
    n = 6
    lat = LatticeSpec(3, (n, n, n), lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    pol = (MoleculeSpec("p", f"(A){N}", "restricted", theta=phi * n ** 3) if restricted
           else MoleculeSpec("p", f"(A){N}", "free", phibulk=phi))
    sysd = SystemSpec(("A", "S"), np.array([[0.0, chi], [chi, 0.0]]), [pol, MoleculeSpec("s", "(S)1", "solvent")])
    return lat, sysd


@pytest.mark.parametrize("N, phi, chi, restricted", [(10, 0.2, 0.7, False), (1, 0.3, 2.5, False),
                                                     (40, 0.05, 0.4, True)])
def test_flory_huggins_chemical_potentials(orc, N, phi, chi, restricted):
    """mu_p = ln phi + (1 - N)(1 - phi) + N chi (1 - phi)^2 and
    mu_s = ln(1 - phi) + (1 - 1/N) phi + chi phi^2; the grand potential of the
    homogeneous solution is 0."""
    lat, sysd = _solution(N, phi, chi, restricted)
    O = orc.System(lat, sysd)
    f = O.residual(np.zeros(sysd.nvar(lat)))
    assert np.abs(f).max() < 1e-12
    mu_p = math.log(phi) + (1 - N) * (1 - phi) + N * chi * (1 - phi) ** 2
    mu_s = math.log(1 - phi) + (1 - 1 / N) * phi + chi * phi ** 2
    assert abs(O.chemical_potential(0) - mu_p) < 1e-12 * max(1, abs(mu_p))
    assert abs(O.chemical_potential(1) - mu_s) < 1e-12
    assert abs(O.grand_potential()) < 1e-10


def _wall_state(orc, phib):
    lat, sysd, x0 = configs.adsorption_1d(n_layers=40, N=10, phibulk=phib, chi_as=0.5, chi_s=-1.0)
    O = orc.System(lat, sysd)
    x, its, norm, _ = O.solve_diis(x0, tol=1e-12, max_iter=20000)
    assert norm < 1e-12
    O.residual(x)
    inner = lat.interior_mask().reshape(-1)
    # excess numbers of chains per molecule over the liquid layers
    gam = [np.sum(O.phi(sysd.states.index("A"))[inner] - O.phibulk(0)) / 10,
           np.sum(O.phi(sysd.states.index("S"))[inner] - O.phibulk(1))]
    return O.grand_potential(), [O.chemical_potential(0), O.chemical_potential(1)], gam


def test_gibbs_adsorption_equation(orc):
    """Adsorption of A_10 from a solution at a wall (chi_s = -1): the grand
    potential of converged profiles at phi_b (1 +- d) changes by
    -sum_m Gamma_m dmu_m, Gamma_m the excess number of molecules m."""
    pb, d = 0.01, 1e-3
    om_m, mu_m, _ = _wall_state(orc, pb * (1 - d))
    om_p, mu_p, _ = _wall_state(orc, pb * (1 + d))
    _, _, gam = _wall_state(orc, pb)
    lhs = om_p - om_m
    rhs = -sum(g * (b - a) for g, a, b in zip(gam, mu_m, mu_p))
    print(f"dOmega = {lhs:.6e}, -sum Gamma dmu = {rhs:.6e}")
    assert abs(lhs) > 1e-7
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


def _site_volumes(orc, lat):
    """Volume of every interior site: the weight ComputeLnGN gives it, read
    as exp(ln G_N) of a unit vector on that site (layer volumes L(z) on curved
    lattices, 1/2 on a mirror2 boundary layer, 1 elsewhere)."""
    L = orc.Lattice(lat)
    w = np.zeros(lat.M)
    for i in np.flatnonzero(lat.interior_mask().reshape(-1)):
        e = np.zeros(lat.M)
        e[i] = 1.0
        w[i] = math.exp(L.lngn(e))
    return w


def _exclusive_sf_cases():
    import cases
    out = []
    for name in cases.NAMES:
        lat, sysd = cases.case(name)
        if sysd.form != "sfbox":
            continue  # the melt form has no grand potential
        owner = {}
        shared = False
        for m, seq in enumerate(sysd.sequences()):
            for t in set(seq):
                shared |= t in owner
                owner[t] = m
        if not shared:  # each state belongs to one molecule, so its phi is that molecule's
            out.append(name)
    return out


@pytest.mark.parametrize("name", _exclusive_sf_cases())
def test_helmholtz_minus_grand_potential_is_sum_mu_n(orc, name):
    """At a converged profile F - Omega = sum_m mu_m n_m, n_m the number of
    molecules m (sum over sites of volume x phi_m, divided by N_m). The
    identity holds only where the residual vanishes, and it ties the free
    energy to the grand and chemical potentials pinned above on walls, curved
    lattices, pores, restricted chains and second order alike: a wrong
    boundary subtraction or layer weight in F breaks it."""
    import cases
    lat, sysd = cases.case(name)
    O = orc.System(lat, sysd)
    x, its, err, rc = O.solve_diis(cases.start(lat, sysd, seed=3, sigma=0.01), tol=1e-11, max_iter=30000)
    assert rc == 0 and err < 1e-11
    O.residual(x)
    w = _site_volumes(orc, lat)
    F, omega = O.free_energy(), O.grand_potential()
    mu_n = 0.0
    for m, seq in enumerate(sysd.sequences()):
        phi_m = sum(O.phi(t) for t in set(seq))
        mu_n += O.chemical_potential(m) * np.sum(w * phi_m) / len(seq)
    print(f"{name}: F = {F:.10f}, Omega = {omega:.10f}, F - Omega - sum mu n = {F - omega - mu_n:.3e}")
    assert abs(F - omega - mu_n) <= 1e-8 * max(1.0, abs(F))
    # not trivially so: the profile is inhomogeneous and F differs from Omega
    assert abs(F - omega) > 1e-3
