"""Differentiable restatement of the melt-form residual on a periodic
simple-cubic lattice, in plain torch float64 (test infrastructure: the
reference gradient for the hand-written tangent sweep).

It follows MPS.cpp:158-198 as generalised to an arbitrary two-state chain:
G_s = g_t(s) * A[G_{s-1}] with A the six-neighbour average (Lat3D1stO.cpp:53-260,
periodic halos = torch.roll), phi_t = theta/(N G_N) sum_{t(s)=t} G_s G_inv_s / g_t
(restricted normalisation, SF_Copol1stO.cpp:1241-1280), contact average
<phi> = A[phi] (Lat3D.cpp:1729-1784), and
f_t = -1 + 1/phi_tot + u'_t - mean_t u'_t,
u'_t = u_t - sum_j chi_tj <phi_j>/phi_tot + sum_j chi_tj phib_j (SF_Solve.cpp:527-591).
"""
import torch


def _avg6(G):
    return (torch.roll(G, 1, 0) + torch.roll(G, -1, 0) + torch.roll(G, 1, 1) + torch.roll(G, -1, 1)
            + torch.roll(G, 1, 2) + torch.roll(G, -1, 2)) / 6.0


def melt_residual(u, seq, chi, theta, phib):
    """u: (2, n, n, n) fields; seq: state index per rank; returns f (2, n, n, n)."""
    g = torch.exp(-u)
    N = len(seq)
    G = [g[seq[0]]]
    for s in range(1, N):
        G.append(g[seq[s]] * _avg6(G[-1]))
    Gi = [None] * N
    Gi[N - 1] = g[seq[N - 1]]
    for s in range(N - 2, -1, -1):
        Gi[s] = g[seq[s]] * _avg6(Gi[s + 1])
    Q = G[-1].sum()
    c = theta / (N * Q)
    phi = [torch.zeros_like(u[0]), torch.zeros_like(u[0])]
    for s in range(N):
        phi[seq[s]] = phi[seq[s]] + G[s] * Gi[s]
    phi = [c * phi[t] / g[t] for t in range(2)]
    rho = [_avg6(p) for p in phi]
    pt = phi[0] + phi[1]
    up = []
    for t in range(2):
        v = u[t].clone()
        for j in range(2):
            v = v - chi[t][j] * rho[j] / pt + chi[t][j] * phib[j]
        up.append(v)
    av = 0.5 * (up[0] + up[1])
    return torch.stack([-1.0 + 1.0 / pt + up[t] - av for t in range(2)])


def to_interior(x, n):
    """(2*M,) halo-padded melt-form vector -> (2, n, n, n)."""
    return x.reshape(2, n + 2, n + 2, n + 2)[:, 1:-1, 1:-1, 1:-1]


def to_padded(f, n):
    out = torch.zeros(2, n + 2, n + 2, n + 2, dtype=f.dtype, device=f.device)
    out[:, 1:-1, 1:-1, 1:-1] = f
    return out.reshape(-1)
