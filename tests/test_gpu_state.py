"""A saved configuration scored on the device: a state file (scattered columns,
a permutation across the imaginary-time boundary, optionally an open worm) is
written, read back, left packed and loaded (pimc_state_*, Context.load_state);
the total action, the energy estimator and the virial estimator on the device
equal the oracle's on the same arrays and links (src/pimc.cpp:1108-1258,
src/path.cpp:145-189).  Tolerances as the other parity tests: 1e-11 relative in
the table mode."""
import numpy as np
import pytest

from tests._state_util import scattered_paths

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA = 0.02198, 6.0596


@pytest.fixture(scope="module")
def gpu():
    import pimc_amd
    return pimc_amd


def loaded(orc, gpu, M, N, K, worm_len, seed, tau=0.02):
    L = orc.box_side(N, DENSITY)
    rng = np.random.default_rng(seed)
    beads, nxt, prv, worm = scattered_paths(rng, M, N, K, worm_len=worm_len, side=L)
    st = gpu.StateFile.create(M, N)
    st.set(beads=beads, next=nxt, prev=prv, worm=worm)
    st.set_counters([[10, 20]])
    st.left_pack()
    st = gpu.StateFile.parse(st.text())              # through the text format
    ctx = gpu.Context(L, M, N, tau, LAMBDA, L, v_mode=gpu.V_TABLE)
    p = ctx.load_state(st)
    ref = orc.State(L, M, N, tau, LAMBDA, L)
    ref.set_beads(p["beads"])
    ref.set_topology(p["bead_on"], p["num_beads"], p["next"], p["prev"])
    return ctx, ref, p


@pytest.mark.parametrize("seed", [1, 2])
def test_closed_state_scores(orc, gpu, seed):
    ctx, ref, p = loaded(orc, gpu, 12, 16, 16, 0, seed)
    assert np.all(p["num_beads"] == 16)
    a, b = ctx.potentialAction(), ref.total_action()
    assert abs(a - b) <= 1e-11 * abs(b)
    eg, er = ctx.energy(next_links=p["next"]), ref.energy()
    for k in er:
        assert abs(eg[k] - er[k]) <= 1e-11 * max(abs(er[k]), 1.0), (k, eg[k], er[k])
    vg, vr = ctx.virial_energy(3, next_links=p["next"]), ref.virial_energy(3)
    for k in vr:
        assert abs(vg[k] - vr[k]) <= 1e-10 * max(abs(vr[k]), 1.0), (k, vg[k], vr[k])


def test_worm_state_scores(orc, gpu):
    """Eleven closed worldlines and a worm of five beads among 16 columns: the
    slices carry 11 or 12 active beads after leftPack."""
    ctx, ref, p = loaded(orc, gpu, 10, 16, 11, 5, 7)
    assert set(p["num_beads"].tolist()) == {11, 12}
    a, b = ctx.potentialAction(), ref.total_action()
    assert abs(a - b) <= 1e-11 * abs(b)
    eg, er = ctx.energy(next_links=p["next"]), ref.energy()
    for k in er:
        assert abs(eg[k] - er[k]) <= 1e-11 * max(abs(er[k]), 1.0), (k, eg[k], er[k])
    # the virial estimator needs closed worldlines
    with pytest.raises(gpu.PimcError):
        ctx.virial_energy(3, next_links=p["next"])


def test_energy_links_change_between_calls(orc, gpu):
    """The energy estimator keeps the last call's links on the device: the same
    links again give the same bits, other links (two worldlines exchanged
    across the last link) the values of a context that saw only those, and the
    first links after them the first values again."""
    ctx, ref, p = loaded(orc, gpu, 12, 16, 16, 0, 1)
    nxt = np.array(p["next"], dtype=np.int32, copy=True)
    M = nxt.shape[0]
    swapped = nxt.copy()
    swapped[M - 1, 0], swapped[M - 1, 1] = nxt[M - 1, 1].copy(), nxt[M - 1, 0].copy()
    e1 = ctx.energy(next_links=nxt)
    e1b = ctx.energy(next_links=nxt)
    e2 = ctx.energy(next_links=swapped)
    e1c = ctx.energy(next_links=nxt)
    other, _, _ = loaded(orc, gpu, 12, 16, 16, 0, 1)
    e2_ref = other.energy(next_links=swapped)
    assert e1 == e1b == e1c
    assert e2 == e2_ref
    assert e1 != e2   # the exchange changes the kinetic link sum
