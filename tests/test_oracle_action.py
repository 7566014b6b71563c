"""Pins the oracle's LocalAction (src/action.cpp:456-1123), minimum image
(include/container.h:50-59) and Path/Worm factors (src/worm.cpp:108-118):
- V(slice) and the total potential action against a numpy brute force over i<j
  and, for N <= 8, against the mpmath closed form at the accumulated bin radius;
- minimum image against the nearest of the 27 periodic images;
- Vnn(bead) against a brute-force r^2 < rc^2 filter;
- the reference's own checkMove invariant (src/move.cpp:190-243): with rc >= maxSep
  the per-bead action change equals the change of the full potential action;
- translation and particle-permutation invariance;
- worm factors 1 / 0.5 / 0 for NONE/SPECIAL/HEADTAIL pairs;
- EnergyEstimator: V = A/(M tau) + N^2 tailV/Vol and the kinetic estimator written out."""
import itertools
import math

import mpmath as mp
import numpy as np
import pytest

from tests._oracle_util import DENSITY, LAMBDA, brute_V_slice, brute_vnn, norm3, put_in_bc, table_V

from tests.test_oracle_aziz import v_mp


def make_state(orc, N, M, tau, rc_frac=1.0, seed=139853, kind=0):
    L = orc.box_side(N, DENSITY)
    st = orc.State(L, M, N, tau, LAMBDA, rc_frac * L)
    b = orc.gen_worldlines(seed, L, M, N, LAMBDA, tau, kind)
    st.set_beads(b)
    return st, b, L


def test_min_image_is_nearest_image(orc):
    rng = np.random.default_rng(1)
    L = 9.0
    for _ in range(300):
        a, c = rng.uniform(-L / 2, L / 2, 3), rng.uniform(-L / 2, L / 2, 3)
        s = put_in_bc(a - c, L)
        best = min(np.linalg.norm(a - c + L * np.array(n)) for n in itertools.product((-1, 0, 1), repeat=3))
        assert abs(np.linalg.norm(s) - best) < 1e-12
    # the oracle's separations (beadSep of the interaction list) are minimum image
    st, b, L = make_state(orc, 8, 4, 0.004)
    lst, sep = st.interaction_list(2, 3)
    for (sl, j), s in zip(lst, sep):
        assert np.array_equal(s, put_in_bc(b[sl, j] - b[2, 3], L))


def test_V_slice_brute_force(orc):
    st, b, L = make_state(orc, 16, 5, 0.004)
    tab, dr = st.aziz.table(0), st.aziz.dr
    for m in range(5):
        assert st.V_slice(m)[1] == pytest.approx(brute_V_slice(b[m], L, tab, dr), rel=1e-13, abs=1e-12)
        assert st.V_slice(m)[0] == 0.0


def test_total_action_mpmath(orc):
    N, M, tau = 6, 4, 0.004
    st, b, L = make_state(orc, N, M, tau)
    dr = 1.0e-6 * 2.9673
    n = int(math.sqrt(3 * (L / 2) ** 2) / dr)
    rk = np.concatenate([[0.0], np.cumsum(np.full(n - 1, dr))])
    tot = mp.mpf(0)
    for m in range(M):
        for i in range(N):
            for j in range(i + 1, N):
                k = int(norm3(put_in_bc(b[m, j] - b[m, i], L)) / dr)
                if 0 < k < n:
                    tot += v_mp(float(rk[k]))
    ref = tau * tot
    assert st.total_action() == pytest.approx(float(ref), rel=1e-12)
    # shift scales tau, VFactor weights even/odd slices
    st.set_action(shift=2, vfactor=(2.0 / 3.0, 4.0 / 3.0))
    ev = sum(st.V_slice(m)[1] for m in range(0, M, 2))
    od = sum(st.V_slice(m)[1] for m in range(1, M, 2))
    assert st.total_action() == pytest.approx(2 * tau * (2 / 3 * ev + 4 / 3 * od), rel=1e-13)


def test_vnn_cutoff_brute_force(orc):
    for frac in (1.0, 0.5, 0.22):
        st, b, L = make_state(orc, 32, 3, 0.004, rc_frac=frac)
        tab, dr = st.aziz.table(0), st.aziz.dr
        for m in range(3):
            for i in (0, 7, 31):
                assert st.Vnn(m, i) == pytest.approx(brute_vnn(b[m], i, L, frac * L, tab, dr), rel=1e-12,
                                                     abs=1e-12)


def test_vnn_slice_counts_pairs_once(orc):
    st, b, L = make_state(orc, 16, 3, 0.004, rc_frac=1.0)
    for m in range(3):
        assert st.Vnn_slice(m) == pytest.approx(st.V_slice(m)[1], rel=1e-12)


def test_checkmove_invariant(orc):
    N, M, tau = 16, 8, 0.004
    st, b, L = make_state(orc, N, M, tau)
    rng = np.random.default_rng(3)
    beads = [(m, int(rng.integers(N))) for m in (1, 2, 3, 5)]
    a0 = st.total_action()
    u0 = st.action_list(beads)
    for (m, i) in beads:
        st.update_bead(m, i, put_in_bc(b[m, i] + rng.normal(scale=0.3, size=3), L))
    a1 = st.total_action()
    u1 = st.action_list(beads)
    assert (u1 - u0) == pytest.approx(a1 - a0, rel=1e-9, abs=1e-9 * abs(a0))
    # range form walks next() links: slices 2..5 of particle 4
    r = st.action_range((2, 4), (5, 4))
    assert r == pytest.approx(st.action_list([(2, 4), (3, 4), (4, 4), (5, 4)]), rel=1e-15)
    assert st.action_bead(3, 4) == pytest.approx(tau * st.Vnn(3, 4), rel=1e-15)


def test_translation_and_permutation_invariance(orc):
    N, M, tau = 16, 6, 0.004
    st, b, L = make_state(orc, N, M, tau)
    a0 = st.total_action()
    perm = np.random.default_rng(2).permutation(N)
    st.set_beads(b[:, perm, :])
    assert st.total_action() == pytest.approx(a0, rel=1e-12)
    st.set_beads(put_in_bc(b + np.array([1.234, -2.5, 0.77]), L))
    assert st.total_action() == pytest.approx(a0, rel=1e-7)


def test_worm_factors(orc):
    N, M, tau = 8, 4, 0.004
    st, b, L = make_state(orc, N, M, tau)
    base = [st.V_slice(m)[1] for m in range(M)]
    on = np.ones((M, N), dtype=np.uint8)
    nb = np.full(M, N, dtype=np.int32)
    nxt = np.zeros((M, N, 2), dtype=np.int32)
    prv = np.zeros((M, N, 2), dtype=np.int32)
    for m in range(M):
        for i in range(N):
            nxt[m, i] = ((m + 1) % M, i)
            prv[m, i] = ((m - 1) % M, i)
    tab, dr = st.aziz.table(0), st.aziz.dr

    def v(m, i, j):
        return table_V(tab, dr, norm3(put_in_bc(b[m, j] - b[m, i], L)))

    st.set_topology(on, nb, nxt, prv, head=(1, 2), tail=(1, 5), special1=(2, 0))
    # slice 1: head(2)-tail(5) pair weight 0, pairs with one of them 0.5
    expect = base[1] - v(1, 2, 5)
    for j in range(N):
        for h in (2, 5):
            if j not in (2, 5):
                expect -= 0.5 * v(1, h, j)
    assert st.V_slice(1)[1] == pytest.approx(expect, rel=1e-12)
    expect2 = base[2] - sum(0.5 * v(2, 0, j) for j in range(1, N))
    assert st.V_slice(2)[1] == pytest.approx(expect2, rel=1e-12)
    assert st.V_slice(0)[1] == pytest.approx(base[0], rel=1e-15)


def test_energy_estimator(orc):
    N, M, tau = 16, 10, 0.004
    st, b, L = make_state(orc, N, M, tau, kind=1)
    e = st.energy()
    A = st.total_action()
    tail = N * N / L ** 3 * st.aziz.tailV
    assert e["V"] == pytest.approx(A / (M * tau) + tail, rel=1e-12)
    assert e["V_int"] == pytest.approx(A / (M * tau) + tail, rel=1e-12)
    vel = put_in_bc(np.roll(b, -1, axis=0) - b, L)
    K = 1.5 * N / tau - (vel ** 2).sum() / (4 * LAMBDA * tau * tau * M)
    assert e["K"] == pytest.approx(K, rel=1e-11)
    assert e["E/N"] == pytest.approx((e["K"] + e["V"]) / N, rel=1e-14)
    # E_mu = E - mu N (src/estimator.cpp:1011); mu = 0 leaves E
    e2 = st.energy(mu=-7.2)
    assert e2["E_mu"] == pytest.approx(e2["E"] + 7.2 * N, rel=1e-14)
    assert e["E_mu"] == e["E"]


def _numpy_sep_hist(bm, L, n=None):
    """updateSepHist over the i<j pairs of one slice (src/action.cpp:216-224):
    50 bins of dSep = sqrt(3) L / 100 (:192); a bin index of 50 or more is dropped."""
    n = bm.shape[0] if n is None else n
    i, j = np.triu_indices(n, 1)
    s = put_in_bc(bm[j] - bm[i], L)
    r = np.sqrt(((0.0 + s[:, 0] * s[:, 0]) + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2])
    k = np.floor(r / (0.5 * math.sqrt(3.0) * L / 50.0)).astype(np.int64)
    return np.bincount(k[k < 50], minlength=50)


def test_sep_hist_numpy(orc):
    """The oracle's separation histogram (the side effect of V(slice)) against a
    numpy histogram of the same slice's minimum-image pair distances, for several
    sizes and both worldline kinds; after several V(slice) calls it holds the last
    slice only (the histogram is cleared per call, :902-947)."""
    for N, M, seed, kind in ((16, 5, 139853, 0), (64, 7, 139854, 1), (33, 3, 139855, 0), (200, 2, 139856, 1)):
        st, b, L = make_state(orc, N, M, 0.004, seed=seed, kind=kind)
        for m in range(M):
            _, h = st.V_slice(m, hist=True)
            assert np.array_equal(h, _numpy_sep_hist(b[m], L)), (N, m)
        _, h2 = st.V_slice(M - 1, hist=True)
        _, h0 = st.V_slice(0, hist=True)
        assert np.array_equal(h0, _numpy_sep_hist(b[0], L))
        if M > 1:
            assert not np.array_equal(h0, h2)


def test_sep_hist_edges(orc):
    """Bin edges of the histogram: a coincident pair counts in bin 0, a pair half
    a box apart on every axis (|r| = sqrt(3) L / 2, bin 50) is dropped, and pairs
    just under half a box apart on one axis count in bin int(L / 2 / dSep) = 28."""
    N, M = 4, 1
    L = orc.box_side(N, DENSITY)
    dsep = 0.5 * math.sqrt(3.0) * L / 50.0
    b = np.zeros((M, N, 3))
    b[0, 0] = [-0.25 * L] * 3
    b[0, 1] = [0.25 * L] * 3            # |r| = sqrt(3) L / 2 from bead 0
    b[0, 2] = [-0.25 * L] * 3           # coincident with bead 0, also the corner from bead 1
    b[0, 3] = [-0.25 * L + 0.5 * L * (1 - 1e-9)] + [-0.25 * L] * 2   # |r| just under L/2 from 0 and 2
    st = orc.State(L, M, N, 0.004, LAMBDA, L)
    st.set_beads(b)
    _, h = st.V_slice(0, hist=True)
    expect = _numpy_sep_hist(b[0], L)
    assert np.array_equal(h, expect)
    assert h[0] == 1                    # the coincident pair
    assert h.sum() == 6 - 2             # both corner pairs dropped
    assert h[int(0.5 * L * (1 - 1e-9) / dsep)] == 2 == h[28]


def test_total_action_threads(orc):
    """The threaded total action (the all-cores CPU baseline) is the serial one
    bit for bit: the same per-slice sums added in slice order."""
    st, b, L = make_state(orc, 24, 9, 0.004, kind=1)
    a = st.total_action()
    for t in (1, 2, 4, 16):
        assert st.total_action(threads=t) == a
    st.set_action(2, (2.0 / 3.0, 4.0 / 3.0))
    assert st.total_action(threads=3) == st.total_action()
