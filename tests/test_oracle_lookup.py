"""Pins the oracle's LookupTable (src/lookuptable.cpp:26-516,
include/lookuptable.h:139-253):
- grid dimensions numNNGrid = max(1, floor(side/rc + EPS)) and the grid index
  closed form int(|pos + side/2 - EPS| / (size + EPS));
- after updateGrid each cell's labels list its beads in ascending particle order;
- the 27-neighbour table keeps the last occurrence of each distinct cell;
- the interaction list equals the brute-force r^2 < rc^2 set (cell walking loses
  nothing), and for a single cell it is the slice in label order minus self;
- updateBead/delBead reproduce a swap-with-last model written independently."""
import itertools

import numpy as np
import pytest

from tests._oracle_util import DENSITY, LAMBDA, put_in_bc

EPS = 1e-7


def state(orc, N, M, frac, seed=139853):
    L = orc.box_side(N, DENSITY)
    st = orc.State(L, M, N, 0.004, LAMBDA, frac * L)
    b = orc.gen_worldlines(seed, L, M, N, LAMBDA, 0.004, 0)
    st.set_beads(b)
    return st, b, L


def grid_index(p, L, n):
    size = L / (1.0 * n)
    return np.array([int(abs(p[i] + 0.5 * L - EPS) / (size + EPS)) for i in range(3)])


@pytest.mark.parametrize("frac,n", [(1.0, 1), (0.5, 2), (0.34, 2), (0.3, 3), (0.22, 4)])
def test_grid_dims_and_index(orc, frac, n):
    st, b, L = state(orc, 64, 3, frac)
    dims, cap = st.grid_dims()
    assert list(dims) == [max(1, int(np.floor(L / (frac * L) + EPS)))] * 3 == [n] * 3
    assert cap == 64 // n ** 3 + 5 or cap > 64 // n ** 3 + 5
    d = st.grid_dump()
    for m in range(3):
        for i in range(64):
            assert np.array_equal(d["grid"][m, i], grid_index(b[m, i], L, n))


def test_labels_in_particle_order(orc):
    st, b, L = state(orc, 64, 4, 0.22)
    d = st.grid_dump()
    n = d["dims"][0]
    for m in range(4):
        cells = {}
        for i in range(64):
            g = d["grid"][m, i]
            cells.setdefault((g[0] * n + g[1]) * n + g[2], []).append(i)
        for c in range(n ** 3):
            lst = cells.get(c, [])
            assert d["num_labels"][c, m] == len(lst)
            assert list(d["hash"][c, m, :len(lst)]) == lst
            for lab, i in enumerate(lst):
                assert d["bead_label"][m, i] == lab


def test_neighbour_table_dedup(orc):
    for frac, n in ((1.0, 1), (0.5, 2), (0.22, 4)):
        st, _, _ = state(orc, 16, 2, frac)
        nn = st.grid_nn()
        shifts = [(a, b_, c) for a in (-1, 0, 1) for b_ in (-1, 0, 1) for c in (-1, 0, 1)]
        for cell in range(n ** 3):
            g = np.array([cell // (n * n), (cell // n) % n, cell % n])
            full = [tuple((g + np.array(s)) % n) for s in shifts]
            for m in range(27):
                if tuple(nn[cell, m]) == (-1, -1, -1):
                    assert full[m] in full[m + 1:]
                else:
                    assert tuple(nn[cell, m]) == full[m] and full[m] not in full[m + 1:]


@pytest.mark.parametrize("frac", [1.0, 0.5, 0.22])
def test_interaction_list_set(orc, frac):
    st, b, L = state(orc, 64, 3, frac)
    rc2 = (frac * L) ** 2
    for m in range(3):
        for i in (0, 13, 63):
            lst, sep = st.interaction_list(m, i)
            got = sorted(j for (_, j) in lst)
            ref = []
            for j in range(64):
                if j != i:
                    s = put_in_bc(b[m, j] - b[m, i], L)
                    if ((0.0 + s[0] ** 2) + s[1] ** 2) + s[2] ** 2 < rc2:
                        ref.append(j)
            assert got == ref
            if frac == 1.0:
                assert [j for (_, j) in lst] == [j for j in range(64) if j != i]


def test_update_bead_swap_with_last(orc):
    N, M = 48, 3
    st, b, L = state(orc, N, M, 0.22)
    d = st.grid_dump()
    n = d["dims"][0]
    size = L / n
    # independent model: per (cell, slice) python lists
    lists = {}
    for m in range(M):
        for c in range(n ** 3):
            lists[(c, m)] = list(d["hash"][c, m, :d["num_labels"][c, m]])
    where = {(m, i): tuple(d["grid"][m, i]) for m in range(M) for i in range(N)}
    rng = np.random.default_rng(11)
    for step in range(400):
        m, i = int(rng.integers(M)), int(rng.integers(N))
        p = put_in_bc(b[m, i] + rng.normal(scale=2.0, size=3), L)
        g = tuple(grid_index(p, L, n))
        if g != where[(m, i)]:
            old = (where[(m, i)][0] * n + where[(m, i)][1]) * n + where[(m, i)][2]
            lst = lists[(old, m)]
            k = lst.index(i)
            last = lst.pop()
            if last != i:
                lst[k] = last
            lists[((g[0] * n + g[1]) * n + g[2], m)].append(i)
            where[(m, i)] = g
        st.update_bead(m, i, p)
        b[m, i] = p
    d = st.grid_dump()
    for (c, m), lst in lists.items():
        assert list(d["hash"][c, m, :d["num_labels"][c, m]]) == lst
        for lab, i in enumerate(lst):
            assert d["bead_label"][m, i] == lab
