"""The benchmark workloads as problem descriptions, with seeded initial fields.

Random numbers: splitmix64 (Steele, Lea, Flood 2014) gives 64-bit words, the top
53 bits make a uniform double in [0,1), and Box-Muller turns pairs into normal
deviates. The same generator feeds every consumer, so the GPU and the CPU
oracle start from identical fields.
"""
from __future__ import annotations

import math

import numpy as np

from .spec import LatticeSpec, MoleculeSpec, SystemSpec

_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive splitmix64 outputs for `seed` (uint64)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed: int, n: int, sigma: float = 1.0) -> np.ndarray:
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0,1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * math.pi * u2)
    z[1::2] = r * np.sin(2 * math.pi * u2)
    return sigma * z[:n]


# ---------------------------------------------------------------------------
def melt_3d(n=64, n_a=12, n_b=28, chi=0.4, sigma=0.1, seed=3, lattice_type="standard"):
    """A(n_a)B(n_b) diblock melt on a periodic n^3 simple-cubic box in the
    MPS.cpp form. n = V/N chains in floating point; the chi-term bulk
    fractions are MPS.cpp's integer quotients N2/N, N1/N (0 for a diblock)."""
    lat = LatticeSpec(3, (n, n, n), lattice_type=lattice_type,
                      lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    N = n_a + n_b
    V = n ** 3
    sysd = SystemSpec(
        states=("A", "B"),
        chi=np.array([[0.0, chi], [chi, 0.0]]),
        molecules=[MoleculeSpec("AB", f"(A){n_a}(B){n_b}", "restricted", theta=float(V))],
        form="melt",
        melt_phibulk=(float(n_a // N), float(n_b // N)),
    )
    x = np.zeros((2, lat.M))
    mask = lat.interior_mask()
    rnd = normal(seed, 2 * V, sigma).reshape(2, V)
    x[:, mask] = rnd
    return lat, sysd, x.reshape(-1)


def adsorption_1d(n_layers=100, N=100, phibulk=0.01, chi_as=0.5, chi_s=-1.0, lam=1.0 / 6.0):
    """Homopolymer A_N in monomeric solvent S next to a frozen wall W in the
    lowerbound surface layer; upperbound mirror1 (config of the 1-D runs)."""
    lat = LatticeSpec(1, (n_layers,), lowerbound=("surface",), upperbound=("mirror1",),
                      lambda_=lam)
    wall = np.zeros(lat.M)
    wall[0] = 1.0
    chi = np.array([[0.0, chi_as, chi_s], [chi_as, 0.0, 0.0], [chi_s, 0.0, 0.0]])
    sysd = SystemSpec(
        states=("A", "S", "W"),
        chi=chi,
        molecules=[MoleculeSpec("pol", f"(A){N}", "free", phibulk=phibulk),
                   MoleculeSpec("sol", "(S)1", "solvent")],
        frozen={"W": wall},
    )
    return lat, sysd, np.zeros(sysd.nvar(lat))


def cylinder_2d(nr=64, nz=128, sigma=0.05, seed=2):
    """A20B20 diblock (phi_b 0.05) in solvent inside a cylindrical pore with an
    attractive wall W at the outer radius; mirror at the axis and both ends."""
    lat = LatticeSpec(2, (nr, nz), geometry="cylindrical",
                      lowerbound=("mirror1", "mirror1"), upperbound=("surface", "mirror1"),
                      lambda_=1.0 / 6.0)
    MX, MY, _ = lat.dims
    wall = np.zeros((MX, MY))
    wall[MX - 1, :] = 1.0
    states = ("A", "B", "S", "W")
    chi = np.zeros((4, 4))
    for a, b, v in (("A", "B", 1.0), ("A", "S", 0.0), ("B", "S", 1.2), ("A", "W", -0.5)):
        i, j = states.index(a), states.index(b)
        chi[i, j] = chi[j, i] = v
    sysd = SystemSpec(
        states=states, chi=chi,
        molecules=[MoleculeSpec("AB", "(A)20(B)20", "free", phibulk=0.05),
                   MoleculeSpec("sol", "(S)1", "solvent")],
        frozen={"W": wall.reshape(-1)},
    )
    nv = sysd.nvar(lat)
    return lat, sysd, normal(seed, nv, sigma)


def batch_1d(count=512, seed0=4):
    """Parameters of uneven 1-D adsorption problems: layers U{50..400}, N
    U{20..1000}, chi_AS U[0,1], chi_s U[-2,0], phi_b 10^U[-4,-1]; problem i
    draws from seed seed0 + i. Each dict feeds sfx.Batch or, through
    adsorption_problem, a single-system description of the same problem."""
    out = []
    for i in range(count):
        u = uniform(seed0 + i, 5)
        out.append({"n_layers": 50 + int(u[0] * 351), "N": 20 + int(u[1] * 981),
                    "chi_as": float(u[2]), "chi_s": float(-2 * u[3]),
                    "phibulk": float(10 ** (-4 + 3 * u[4])), "lambda_": 1.0 / 6.0})
    return out


def _fh_bisect(fn, lo, hi, iters=200):
    """Root of fn on [lo, hi] with fn(lo) < 0 < fn(hi), by bisection."""
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if fn(mid) < 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def fh_binodal(N, chi):
    """Coexisting volume fractions (dilute, concentrated) of A_N in a monomeric
    solvent under Flory-Huggins, f(phi) = phi/N ln phi + (1-phi) ln(1-phi)
    + chi phi (1-phi) per site, or None when chi <= chi_c = (1 + N^-1/2)^2 / 2.

    The two phases share f'(phi) (exchange potential) and f - phi f'
    (osmotic pressure). The dilute root is found in ln phi, since it falls far
    below 1e-300 for long chains in a poor solvent."""
    if chi <= 0.5 * (1.0 + N ** -0.5) ** 2:
        return None

    def d1(la):  # f'(phi) at phi = exp(la)
        p = math.exp(la)
        return (la + 1.0) / N - math.log1p(-p) - 1.0 + chi * (1.0 - 2.0 * p)

    def press(la):  # f - phi f' at phi = exp(la)
        p = math.exp(la)
        return math.log1p(-p) + p * (1.0 - 1.0 / N) + chi * p * p

    pc = 1.0 / (1.0 + math.sqrt(N))
    spin = lambda p: 1.0 / (N * p) + 1.0 / (1.0 - p) - 2.0 * chi  # f'' = 0
    s1 = _fh_bisect(lambda p: -spin(p), 1e-300, pc)
    s2 = _fh_bisect(spin, pc, 1.0 - 1e-16)
    ls1, ls2 = math.log(s1), math.log(s2)

    # f' rises on (0, s1) and on (s2, 1) and falls in between, so every common
    # slope t in (f'(s2), f'(s1)) has one root a(t) < s1 and one b(t) > s2. The
    # tangent intercepts differ by h(t) = press(a) - press(b) with
    # dh/dt = b - a > 0: h has one root, the binodal.
    def roots(t):
        la = _fh_bisect(lambda v: d1(v) - t, -1e6, ls1)
        lb = _fh_bisect(lambda v: d1(v) - t, ls2, math.log1p(-1e-16))
        return la, lb

    def h(t):
        la, lb = roots(t)
        return press(la) - press(lb)

    la, lb = roots(_fh_bisect(h, d1(ls2), d1(ls1)))
    return math.exp(la), math.exp(lb)


def bulk_is_stable(p):
    """True when the bulk of a batch_1d-style problem is a stable one-phase
    Flory-Huggins solution: chi_as <= chi_c, or phi_b below the dilute binodal.
    Otherwise the bulk is metastable or unstable, the wall nucleates a
    condensed layer that grows with the box, and a converged profile need not
    exist."""
    b = fh_binodal(p["N"], p["chi_as"])
    return b is None or p["phibulk"] <= b[0]


def adsorption_problem(p):
    """(lattice, system, zero start) of one batch_1d-style problem dict."""
    return adsorption_1d(p["n_layers"], p["N"], phibulk=p["phibulk"], chi_as=p["chi_as"],
                         chi_s=p["chi_s"], lam=p.get("lambda_", 1.0 / 6.0))


def semiflexible_3d(n=128, N=200, phi=0.3, chi=0.6, stiffness_energy=1.0, sigma=0.1, seed=5):
    """Second-order homopolymer A_N (restricted, theta = phi V) in a monomeric
    solvent on a periodic n^3 simple-cubic lattice. The straight-bond weight is
    S = 1/(1 + 4 exp(-eps)) for a bending energy eps per right angle (kT), so
    each of the four perpendicular directions has (1 - S)/4 and reversal 0."""
    lat = LatticeSpec(3, (n, n, n), order=2, lowerbound=("periodic",) * 3, upperbound=("periodic",) * 3)
    S = 1.0 / (1.0 + 4.0 * math.exp(-stiffness_energy))
    sysd = SystemSpec(
        states=("A", "S"),
        chi=np.array([[0.0, chi], [chi, 0.0]]),
        molecules=[MoleculeSpec("pol", f"(A){N}", "restricted", theta=phi * n ** 3),
                   MoleculeSpec("sol", "(S)1", "solvent")],
        stiffness=S,
    )
    return lat, sysd, normal(seed, sysd.nvar(lat), sigma)
