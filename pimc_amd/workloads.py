"""Parameters of the measured workloads (He-4 at 0.02198 A^-3, Aziz 1979).

Sizes follow the reference's Prism(density, N) box (src/container.cpp:84-109);
tau = 1/(M T) as in src/setup.cpp:999-1008 but without forcing M even.
"""
import math

DENSITY = 0.02198
LAMBDA = 6.0596
SEED0 = 139853      # the reference's base seed (src/pdrive.cpp:41,63)


def box_side(n, density=DENSITY):
    return (1.0 * n / density) ** (1.0 / 3.0)


WORKLOADS = {
    "small_n16_m125": dict(N=16, M=125, T=2.0, tau=0.004, rc_frac=1.0, q=("max_int", "4 4 4")),
    "walker_n256_m167": dict(N=256, M=167, T=1.5, tau=0.004, rc_frac=1.0, q=("max_float", "3 0 0")),
    "large_n2048_m250": dict(N=2048, M=250, T=1.0, tau=0.004, rc_frac=0.5, q=("max_float", "4 0 0")),
}


def lattice_q(side, qmag_max=None, nmax=None):
    """Lattice wavevectors in the reference's max_float/max_int order (src/estimator.cpp:487-541)."""
    import numpy as np
    if nmax is None:
        nm = 1 + int(qmag_max * side / 2.0 / math.pi)
        lim = qmag_max
    else:
        nm = nmax
        q1 = nm * 2.0 * math.pi / side
        lim = math.sqrt(((0.0 + q1 * q1) + q1 * q1) + q1 * q1)
    r = np.arange(-nm, nm + 1)
    n = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    q = n * 2.0 * math.pi / side
    mag = np.sqrt(((0.0 + q[:, 0] * q[:, 0]) + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2])
    return q[mag <= lim]


def batched_walkers(n=512, seed=SEED0):
    """The batched workload's walkers: N uniform over {16, 32, 64, 128, 256}, M
    over {64, 128, 256, 512}, T uniform in [1, 3] K, tau = 1/(M T), rc = L.
    Returns dicts (side, M, N, tau, lam, rc, seed, kind); walker k's worldlines
    are the Gaussian bridge (kind 1) of MTRand(seed + k), rounded to FP32."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        N = int(rng.choice([16, 32, 64, 128, 256]))
        M = int(rng.choice([64, 128, 256, 512]))
        T = float(rng.uniform(1.0, 3.0))
        L = box_side(N)
        out.append(dict(side=L, M=M, N=N, tau=1.0 / (M * T), lam=LAMBDA, rc=L, seed=seed + k, kind=1))
    return out


def served_chains(n=1024, seed=SEED0 + 1):
    """The serving workload's chains: N uniform in [32, 512], M in [64, 512], T
    uniform in [1, 3] K, tau = 1/(M T), rc = L; chain k's worldlines are the
    Gaussian bridge (kind 1) of MTRand(SEED0 + 10000 + k)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        N = int(rng.integers(32, 513))
        M = int(rng.integers(64, 513))
        T = float(rng.uniform(1.0, 3.0))
        L = box_side(N)
        out.append(dict(side=L, M=M, N=N, tau=1.0 / (M * T), lam=LAMBDA, rc=L, seed=SEED0 + 10000 + k, kind=1))
    return out
