"""Where the 64^3 A12B28 melt sits relative to its lattice spinodal.

A numpy lattice-RPA Jacobian of the melt residual at the homogeneous state is
pinned against the oracle's central-difference Jacobian on a small lattice,
then its smallest exchange eigenvalue is followed in chi on the 64^3 lattice:
positive at chi = 0.4, negative just above, and 0.05 at chi = 0.38 (the first
step of the chi sweep the melt solve runs)."""
import numpy as np
import pytest

from field_amd import configs


def rpa_jacobian(n, n_a, n_b, chi):
    """2x2 block per wave vector of d f / d u at u = 0 for the MPS-form
    AB melt on a periodic n^3 simple-cubic lattice.

    a(q) = (cos qx + cos qy + cos qz) / 3 is the eigenvalue of the six-neighbour
    average; the Debye matrix S_tt'(q) = (1/N) sum over segment pairs of types
    (t, t') of a^|s-s'|; dphi = -S du, du'_t = du_t - chi (a dphi_t' - phi0_t'
    dphi_tot) for t' != t, df_t = -dphi_tot + du'_t - mean_t du'_t."""
    N = n_a + n_b
    seq = [0] * n_a + [1] * n_b
    k = 2 * np.pi * np.fft.fftfreq(n)
    a = (np.cos(k)[:, None, None] + np.cos(k)[None, :, None] + np.cos(k)[None, None, :]) / 3
    cnt = np.zeros((2, 2, N))
    for s in range(N):
        for t in range(N):
            cnt[seq[s], seq[t], abs(s - t)] += 1
    S = np.zeros((2, 2) + a.shape)
    for d in range(N):
        S += (cnt[:, :, d] / N)[:, :, None, None, None] * a[None, None] ** d
    phi0 = np.array([n_a / N, n_b / N])
    J = np.zeros((2, 2) + a.shape)
    for c in range(2):
        dphi = -S[:, c]
        tot = dphi[0] + dphi[1]
        dup = np.zeros((2,) + a.shape)
        for t in range(2):
            o = 1 - t
            dup[t] = (1.0 if t == c else 0.0) - chi * (a * dphi[o] - phi0[o] * tot)
        mean = dup.mean(axis=0)
        for t in range(2):
            J[t, c] = -tot + dup[t] - mean
    J[:, :, 0, 0, 0] = np.eye(2) - 0.5  # q = 0: the restricted normalisation keeps dphi = 0
    return J


def test_rpa_jacobian_matches_oracle_jacobian(orc):
    """Column by column on a 5^3 lattice with a short chain: the numpy J0 applied
    to a unit vector equals the oracle's central difference of the residual."""
    n, n_a, n_b, chi = 5, 2, 4, 0.7
    lat, sysd, _ = configs.melt_3d(n=n, n_a=n_a, n_b=n_b, chi=chi)
    O = orc.System(lat, sysd)
    J = rpa_jacobian(n, n_a, n_b, chi)
    M = lat.M
    inner = np.flatnonzero(lat.interior_mask())
    rng = np.random.default_rng(2)
    for _ in range(4):
        v = np.zeros((2, n, n, n))
        v[rng.integers(2), rng.integers(n), rng.integers(n), rng.integers(n)] = 1.0
        V = np.fft.fftn(v, axes=(1, 2, 3))
        jv = np.fft.ifftn(np.einsum("ij...,j...->i...", J, V), axes=(1, 2, 3)).real
        x = np.zeros(2 * M)
        x.reshape(2, M)[:, inner] = v.reshape(2, -1)
        h = 1e-6
        fd = (O.residual(h * x) - O.residual(-h * x)) / (2 * h)
        fd = fd.reshape(2, M)[:, inner].reshape(2, n, n, n)
        # the oracle's f carries the uniform gauge of the melt (a shift of every u_t)
        # as a zero mode: compare after removing each component's mean
        fd -= fd.mean(axis=(1, 2, 3), keepdims=True)
        jv -= jv.mean(axis=(1, 2, 3), keepdims=True)
        assert np.abs(fd - jv).max() < 1e-7 * max(1.0, np.abs(fd).max())


def _min_exchange_eig(n, chi):
    J = rpa_jacobian(n, 12, 28, chi)
    k = 2 * np.pi * np.fft.fftfreq(n)
    q2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    ev = np.linalg.eigvals(np.moveaxis(J, (0, 1), (-2, -1)))
    sel = (q2 > 0.09) & (q2 < 4.0)  # away from q = 0 and the checkerboard
    return ev[sel].real.min()


@pytest.mark.parametrize("chi, lo, hi", [(0.38, 0.045, 0.055), (0.4, 1e-4, 1e-3), (0.401, -1.0, -1e-4)])
def test_melt_spinodal_on_64_cubed(chi, lo, hi):
    assert lo < _min_exchange_eig(64, chi) < hi
