"""Writes tests/golden/state_m4_n5.txt (a state file in the reference's text
format, from the independent writer state_text.py) and state_m4_n5.npz (its
arrays).  Three closed worldlines joined by a permutation, one open worm of two
beads, scattered columns; counters for the totals, three moves and two
estimators; MTRand save words of seed 139853 after 1000 draws.

    python tests/golden/make_state_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as orc  # noqa: E402
from tests._state_util import scattered_paths  # noqa: E402
from tests.golden.state_text import state_text  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    M, N = 4, 5
    beads, nxt, prv, worm = scattered_paths(rng, M, N, K=3, worm_len=2, side=9.0)
    beads[0, 0] = (-0.0, 1e-300, 4.5)          # signed zero, a denormal-range value, the box edge
    counters = np.array([[1234, 5678], [11, 20], [0, 0], [999999, 1000000], [7, 7], [4321, 4321]], np.int64)
    mt = orc.MTRand(139853)
    for _ in range(1000):
        mt.randInt()
    words = mt.save()
    text = state_text(N, counters.tolist(), beads.tolist(), nxt.tolist(), prv.tolist(), worm.tolist(),
                      [int(w) for w in words])
    with open(os.path.join(HERE, "state_m4_n5.txt"), "w") as f:
        f.write(text)
    np.savez(os.path.join(HERE, "state_m4_n5.npz"), beads=beads, next=nxt, prev=prv, worm=worm,
             counters=counters, random=np.asarray(words, dtype=np.uint64), num_worldlines=N)


if __name__ == "__main__":
    main()
