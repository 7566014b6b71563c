"""The state-file component (pimc_state_*, host code of the library) against the
reference's text format and load semantics (src/pimc.cpp:926-986, :1006-1027,
:1108-1258; include/common.h:244-397; src/path.cpp:145-189).

* Format: the library's writer is byte-identical to an independent writer built
  from the reference's stream operators (tests/golden/state_text.py) on the
  committed fixture; parsing the fixture returns its arrays (positions at the
  16 significant digits the format keeps), the counters and the MTRand words;
  write -> read -> write is the identity.
* leftPack: the library equals the oracle's restatement bit for bit; the
  oracle is pinned by invariants (active beads packed left with the same count
  per slice, every worldline read along next links keeps its positions, links
  stay mutual).
* Load set-up: same slice count -> the file's arrays with positions put in the
  box and numBeadsAtSlice; other slice count -> the classical path of
  loadClassicalState, written here in numpy.
No GPU: these are host functions of the library."""
import os

import numpy as np
import pytest

from tests._state_util import scattered_paths, walk_positions
from tests.golden.state_text import state_text

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "state_m4_n5")


@pytest.fixture(scope="module")
def pa():
    import pimc_amd
    return pimc_amd


def r16(x):
    return np.vectorize(lambda v: float("%.16g" % v))(x)


def test_fixture_parse(pa):
    ref = np.load(FIX + ".npz")
    st = pa.StateFile.read(FIX + ".txt")
    assert st.dims() == (4, 5, 5)
    a = st.arrays()
    assert np.array_equal(a["beads"], r16(ref["beads"]))
    assert np.signbit(a["beads"][0, 0, 0])                 # "-0" stays negative zero
    assert np.array_equal(a["next"], ref["next"]) and np.array_equal(a["prev"], ref["prev"])
    assert np.array_equal(a["worm"], ref["worm"])
    assert np.array_equal(st.counters(), ref["counters"])
    assert np.array_equal(st.random(), ref["random"])


def test_fixture_format_bytes(pa):
    ref = np.load(FIX + ".npz")
    with open(FIX + ".txt") as f:
        text = f.read()
    st = pa.StateFile.create(4, 5)
    st.set(beads=ref["beads"], next=ref["next"], prev=ref["prev"], worm=ref["worm"])
    st.set_counters(ref["counters"])
    st.set_random(ref["random"])
    assert st.text() == text
    # and the parsed file reprints to the same bytes
    assert pa.StateFile.parse(text).text() == text


def test_round_trip_and_independent_writer(pa, tmp_path):
    rng = np.random.default_rng(5)
    M, N = 7, 9
    beads, nxt, prv, worm = scattered_paths(rng, M, N, K=6, worm_len=3, side=20.0)
    beads = beads * 10.0 ** rng.integers(-8, 8, size=beads.shape)   # exponents in the text
    st = pa.StateFile.create(M, N)
    st.set(beads=beads, next=nxt, prev=prv, worm=worm)
    st.set_counters([[1, 2], [3, 4]])
    text = st.text()
    assert text == state_text(N, [(1, 2), (3, 4)], beads.tolist(), nxt.tolist(), prv.tolist(), worm.tolist(), None)
    path = tmp_path / "state"
    st.write(str(path))
    back = pa.StateFile.read(str(path))
    a = back.arrays()
    assert np.array_equal(a["beads"], r16(beads))
    assert back.random() is None and back.text() == text


def test_parse_errors(pa):
    with open(FIX + ".txt") as f:
        text = f.read()
    for bad in (text.replace("(0,3) x (0,4)", "(0,3) x (0,5)", 2),      # link extents differ from beads
                text[: len(text) // 2],                                 # truncated
                text.replace("(1,", "(9,", 1)):                         # link outside the arrays
        with pytest.raises(pa.PimcError):
            pa.StateFile.parse(bad)


@pytest.mark.parametrize("seed", range(6))
def test_left_pack_matches_oracle(pa, orc, seed):
    rng = np.random.default_rng(seed)
    M, N = int(rng.integers(2, 9)), int(rng.integers(3, 12))
    K = int(rng.integers(1, N))
    wl = int(rng.integers(0, M + 1)) if K < N else 0
    beads, nxt, prv, worm = scattered_paths(rng, M, N, K, worm_len=wl)
    st = pa.StateFile.create(M, N)
    st.set(beads=beads, next=nxt, prev=prv, worm=worm)
    st.left_pack()
    got = st.arrays()
    b, n, p, w = orc.left_pack(beads, nxt, prv, worm)
    assert np.array_equal(got["beads"], b) and np.array_equal(got["worm"], w)
    assert np.array_equal(got["next"], n) and np.array_equal(got["prev"], p)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_left_pack_invariants(orc, seed):
    rng = np.random.default_rng(100 + seed)
    M, N = int(rng.integers(2, 9)), int(rng.integers(3, 12))
    K = int(rng.integers(1, N))
    wl = int(rng.integers(1, M + 1)) if K < N else 0
    beads, nxt, prv, worm = scattered_paths(rng, M, N, K, worm_len=wl)
    b, n, p, w = orc.left_pack(beads, nxt, prv, worm)
    counts = worm.sum(axis=1)
    assert np.array_equal(w.sum(axis=1), counts)
    for m in range(M):
        assert np.all(w[m, :counts[m]] == 1) and np.all(w[m, counts[m]:] == 0)
        assert np.all(n[m, counts[m]:] == -1) and np.all(p[m, counts[m]:] == -1)
    for m in range(M):
        for i in range(counts[m]):
            if tuple(n[m, i]) != (-1, -1):
                assert tuple(p[n[m, i, 0], n[m, i, 1]]) == (m, i)
            if tuple(p[m, i]) != (-1, -1):
                assert tuple(n[p[m, i, 0], p[m, i, 1]]) == (m, i)
    assert sorted(walk_positions(beads, nxt, worm)) == sorted(walk_positions(b, n, w))


def test_to_path_same_slices(pa):
    rng = np.random.default_rng(11)
    M, N, side = 6, 8, 12.0
    beads, nxt, prv, worm = scattered_paths(rng, M, N, K=5, worm_len=2, side=1.5 * 12.0)   # some outside the box
    st = pa.StateFile.create(M, N)
    st.set(beads=beads, next=nxt, prev=prv, worm=worm)
    st.left_pack()
    a = st.arrays()
    p = st.to_path(M, side)
    expect = a["beads"] - side * np.floor(a["beads"] / side + 0.5)
    assert np.allclose(p["beads"], expect, rtol=0, atol=1e-12)
    assert np.all(np.abs(p["beads"]) <= 0.5 * side)
    assert np.array_equal(p["bead_on"], (a["worm"] != 0).astype(np.uint8))
    assert np.array_equal(p["num_beads"], a["worm"].sum(axis=1))
    assert np.array_equal(p["next"], a["next"]) and np.array_equal(p["prev"], a["prev"])


def test_to_path_classical(pa):
    """A file with another slice count: loadClassicalState puts each worldline
    active on the file's slice 0 at its slice-0 position on every slice."""
    ref = np.load(FIX + ".npz")
    st = pa.StateFile.read(FIX + ".txt")
    M, side = 10, 9.0
    p = st.to_path(M, side)
    act = np.nonzero(ref["worm"][0])[0]
    k = len(act)
    x0 = r16(ref["beads"][0, act])
    x0 = x0 - side * np.floor(x0 / side + 0.5)
    for m in range(M):
        assert np.array_equal(p["beads"][m, :k], x0)
        assert np.all(p["bead_on"][m, :k] == 1) and np.all(p["bead_on"][m, k:] == 0)
        for i in range(k):
            assert tuple(p["next"][m, i]) == ((m + 1) % M, i) and tuple(p["prev"][m, i]) == ((m - 1) % M, i)
        assert np.all(p["next"][m, k:] == -1) and np.all(p["prev"][m, k:] == -1)
    assert np.all(p["num_beads"] == k)
