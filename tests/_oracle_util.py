"""Independent numpy re-derivations used to pin the oracle on small inputs."""
import math

import numpy as np

DENSITY = 0.02198
LAMBDA = 6.0596


def put_in_bc(s, L):
    return s - L * np.floor(s * (1.0 / L) + 0.5)


def table_V(tab, dr, r):
    k = int(r / dr)
    if k <= 0 or k >= len(tab):
        return 0.0
    return tab[k]


def norm3(s):
    return math.sqrt(((0.0 + s[0] * s[0]) + s[1] * s[1]) + s[2] * s[2])


def brute_V_slice(beads_m, L, tab, dr, n=None):
    n = beads_m.shape[0] if n is None else n
    tot = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            s = put_in_bc(beads_m[j] - beads_m[i], L)
            tot += 1.0 * table_V(tab, dr, norm3(s))
    return tot


def brute_vnn(beads_m, i, L, rc, tab, dr):
    tot = 0.0
    for j in range(beads_m.shape[0]):
        if j == i:
            continue
        s = put_in_bc(beads_m[j] - beads_m[i], L)
        if ((0.0 + s[0] * s[0]) + s[1] * s[1]) + s[2] * s[2] < rc * rc:
            tot += table_V(tab, dr, norm3(s))
    return tot


def phasor_rho(beads, q):
    """rho_q(m) = sum_i exp(i q.r_i(m)), beads [M][N][3], q [nq][3] -> [nq][M]."""
    ph = np.einsum("qd,mnd->qmn", q, beads)
    return np.exp(1j * ph).sum(axis=2)


def lattice_rho(beads, n, L):
    """FP64 density modes rho_n(m) = sum_i exp(2 pi i n . r_i(m) / L) for integer
    lattice vectors n [nq][3]; beads [M][N][3] -> [nq][M] complex.

    Per slice the axis phasors exp(2 pi i n_d x_d / L) are tabulated for every
    n_d in range, the x-y products form the rows of one matrix and one product
    with the z table gives every (n_x, n_y, n_z) at once.  Only n_x >= 0 is
    formed; rho_{-n} = conj(rho_n) (real positions) supplies the rest."""
    n = np.asarray(n, dtype=np.int64)
    M, N, _ = beads.shape
    h = np.abs(n).max(axis=0)
    s = np.asarray(beads, dtype=np.float64) / L
    kx = np.arange(0, h[0] + 1)
    ky = np.arange(-h[1], h[1] + 1)
    kz = np.arange(-h[2], h[2] + 1)
    neg = n[:, 0] < 0
    nn = np.where(neg[:, None], -n, n)
    flat = (nn[:, 0] * ky.size + (nn[:, 1] + h[1])) * kz.size + (nn[:, 2] + h[2])
    out = np.empty((n.shape[0], M), np.complex128)
    for m in range(M):
        ex = np.exp(2j * np.pi * np.outer(kx, s[m, :, 0]))
        ey = np.exp(2j * np.pi * np.outer(ky, s[m, :, 1]))
        ez = np.exp(2j * np.pi * np.outer(kz, s[m, :, 2]))
        rows = (ex[:, None, :] * ey[None, :, :]).reshape(-1, N)
        out[:, m] = (rows @ ez.T).reshape(-1)[flat]
    out[neg] = np.conj(out[neg])
    return out


def isf_from_rho(rho, n_particles):
    """F(q, tau) = sum_m Re rho(m) conj(rho(m + tau)) / N for every lag (cyclic in
    m), from rho [nq][M] through the correlation theorem: the sum over m equals
    (1/M) sum_k |R(k)|^2 exp(-2 pi i k tau / M), R the DFT of rho over m."""
    R = np.fft.fft(rho, axis=1)
    return (np.fft.fft(np.abs(R) ** 2, axis=1).real / rho.shape[1]) / n_particles
