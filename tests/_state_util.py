"""Random worldline topologies for the state-file tests: K closed worldlines
joined by a random permutation across the M-1 -> 0 link, plus optionally one
open worm of `worm_len` beads, scattered over random columns on every slice
(so the arrays are not left packed).  Links are {slice, ptcl}, XXX = (-1, -1)."""
import numpy as np


def scattered_paths(rng, M, N, K, worm_len=0, side=10.0):
    assert K + (1 if worm_len else 0) <= N and worm_len <= M
    beads = rng.uniform(-0.5 * side, 0.5 * side, size=(M, N, 3))
    nxt = -np.ones((M, N, 2), np.int32)
    prv = -np.ones((M, N, 2), np.int32)
    worm = np.zeros((M, N), np.uint32)
    col = [rng.permutation(N) for _ in range(M)]       # worldline k sits in column col[m][k]
    perm = rng.permutation(K)
    for m in range(M):
        for k in range(K):
            c = col[m][k]
            worm[m, c] = 1
            m2 = (m + 1) % M
            k2 = k if m + 1 < M else perm[k]
            nxt[m, c] = (m2, col[m2][k2])
            prv[m2, col[m2][k2]] = (m, c)
    if worm_len:
        # the worm occupies logical index K on slices s0 .. s0 + worm_len - 1 (no wrap)
        s0 = int(rng.integers(0, M - worm_len + 1))
        for t in range(worm_len):
            m = s0 + t
            c = col[m][K]
            worm[m, c] = 1
            if t + 1 < worm_len:
                c2 = col[m + 1][K]
                nxt[m, c] = (m + 1, c2)
                prv[m + 1, c2] = (m, c)
    return beads, nxt, prv, worm


def walk_positions(beads, nxt, worm):
    """The set of worldlines as tuples of positions, each read along next links
    from its first bead (a bead with no prev, or the smallest (slice, column) of
    a closed cycle) -- independent of the columns the beads sit in."""
    M, N = worm.shape
    prv_of = {}
    for m in range(M):
        for i in range(N):
            if worm[m, i] and tuple(nxt[m, i]) != (-1, -1):
                prv_of[tuple(nxt[m, i])] = (m, i)
    seen, lines = set(), []
    starts = [(m, i) for m in range(M) for i in range(N) if worm[m, i] and (m, i) not in prv_of]
    for m in range(M):
        for i in range(N):
            if worm[m, i]:
                starts.append((m, i))
    for s in starts:
        if s in seen:
            continue
        line, b = [], s
        while b != (-1, -1) and b not in seen:
            seen.add(b)
            line.append(tuple(beads[b[0], b[1]]))
            b = tuple(nxt[b[0], b[1]])
        lines.append(tuple(line))
    return lines
