"""Seeded synthetic worldlines for the measured workloads.

The random stream is the reference's MTRand (include/MersenneTwister.h): numpy's
legacy MT19937 seeding is init_genrand, and the transforms are written out here
(rand = u/4294967295, randExc = u/4294967296, randNorm = Marsaglia polar keeping
the first variate, :325-363).  Slice 0 is Prism::randPosition
(src/container.cpp:155-160).  kind 0: slices m > 0 are randNorm(previous,
sqrt(2 lambda tau)) per axis followed by putInBC.  kind 1: a closed Gaussian
bridge built with the staging construction (src/move.cpp:270-291) from slice 0
back to slice 0.
"""
import math

import numpy as np


class _Stream:
    def __init__(self, seed):
        self._bg = np.random.RandomState(seed)._bit_generator
        self._buf = np.empty(0, dtype=np.uint64)
        self._pos = 0

    def raw(self, n):
        if self._pos + n > self._buf.size:
            more = max(n, 1 << 16)
            self._buf = np.concatenate([self._buf[self._pos:], self._bg.random_raw(more)])
            self._pos = 0
        out = self._buf[self._pos:self._pos + n]
        self._pos += n
        return out

    def rand_exc(self, n):
        return self.raw(n).astype(np.float64) * (1.0 / 4294967296.0)

    def polar(self, n):
        """n variates x*sqrt(-2 log r / r), one per accepted (x, y) pair, in order."""
        out = np.empty(n)
        got = 0
        while got < n:
            need = n - got
            k = int(need * 1.35) + 16
            u = self.raw(2 * k).astype(np.float64) * (1.0 / 4294967295.0)
            x = 2.0 * u[0::2] - 1.0
            y = 2.0 * u[1::2] - 1.0
            r = x * x + y * y
            ok = ~((r >= 1.0) | (r == 0.0))
            idx = np.nonzero(ok)[0]
            if idx.size > need:
                # hand back the raw words after the last pair used
                last = idx[need - 1]
                self._pos -= 2 * (k - last - 1)
                idx = idx[:need]
            xs, rs = x[idx], r[idx]
            # libm log (math.log), not numpy's vectorised log, so the last bit
            # matches the reference's std::log
            logs = np.fromiter((math.log(v) for v in rs), dtype=np.float64, count=rs.size)
            out[got:got + idx.size] = xs * np.sqrt(-2.0 * logs / rs)
            got += idx.size
        return out


def put_in_bc(r, L):
    return r - L * np.floor(r * (1.0 / L) + 0.5)


def generate(seed, side, M, N, lam, tau, kind=0):
    """beads [M][N][3] FP64."""
    s = _Stream(seed)
    b = np.empty((M, N, 3))
    b[0] = side * (-0.5 + s.rand_exc(3 * N).reshape(N, 3))
    if kind == 0:
        sd = math.sqrt(2.0 * lam * tau)
        g = s.polar(3 * N * (M - 1)).reshape(M - 1, N, 3)
        for m in range(1, M):
            b[m] = put_in_bc(b[m - 1] + g[m - 1] * sd, side)
    else:
        s2lt = math.sqrt(2.0) * math.sqrt(lam * tau)
        g = s.polar(3 * N * (M - 1)).reshape(N, M - 1, 3)
        end = b[0].copy()
        for k in range(M - 1):
            f1 = 1.0 * (M - k - 1)
            f2 = 1.0 / (1.0 * (M - k))
            sq = s2lt * math.sqrt(f1 * f2)
            r = put_in_bc(end - b[k], side)
            r = r * f2
            r = r + b[k]
            r = r + g[:, k, :] * sq
            b[k + 1] = put_in_bc(r, side)
    return b
