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
