"""The oracle's GSF / Li-Broughton correction (src/action.cpp:456-806), pinned by
identities that do not restate its loops:

* gradVSquared of two beads is 2 |gradV(r12)|^2 (gradV is odd in r);
* gradVSquared(slice) equals a numpy sum of the per-bead forces;
* the reference's checkMove invariant: with every pair inside the cutoff, the
  change of one bead's action including gradVnnSquared equals the change of the
  total action including gradVSquared (src/move.cpp:190-243) -- this pins the
  three-body term against the two-body one, sign and pair structure included;
* the thermodynamic energy equals (1/M) dA/dtau - lambda/(tau M) dA/dlambda +
  tail, with the derivatives taken numerically from the total action (A is a
  cubic in tau and linear in lambda at fixed positions)."""
import numpy as np
import pytest

from tests._oracle_util import DENSITY, LAMBDA

GSF = dict(vfactor=(2.0 / 3.0, 4.0 / 3.0), gradvfactor=(0.0, 2.0 / 9.0), period=2)
LB = dict(vfactor=(1.0, 1.0), gradvfactor=(1.0 / 12.0, 1.0 / 12.0), period=2)


def make(orc, N, M, tau, rc_frac=1.0, seed=7, action=GSF, kind=0):
    L = orc.box_side(N, DENSITY)
    st = orc.State(L, M, N, tau, LAMBDA, rc_frac * L)
    st.set_beads(orc.gen_worldlines(seed, L, M, N, LAMBDA, tau, kind))
    st.set_action(1, action["vfactor"])
    st.set_correction(action["gradvfactor"], action["period"])
    return st, L


def test_two_beads(orc):
    st, L = make(orc, 2, 4, 0.01)
    az = st.aziz
    b = st.beads()
    for m in range(4):
        sep = b[m, 0] - b[m, 1]
        sep -= L * np.floor(sep / L + 0.5)
        g = np.array(az.gradV(sep))
        assert np.isclose(st.gradV2_slice(m), 2.0 * g @ g, rtol=1e-13, atol=0)


def test_slice_sum_numpy(orc):
    N, M = 9, 6
    st, L = make(orc, N, M, 0.01, seed=11)
    az = st.aziz
    b = st.beads()
    for m in range(M):
        F = np.zeros((N, 3))
        for i in range(N):
            for j in range(N):
                if i != j:
                    sep = b[m, i] - b[m, j]
                    sep -= L * np.floor(sep / L + 0.5)
                    F[i] += np.array(az.gradV(sep))
        assert np.isclose(st.gradV2_slice(m), np.sum(F * F), rtol=1e-12, atol=0)


@pytest.mark.parametrize("action", [GSF, LB])
def test_checkmove_invariant(orc, action):
    N, M, tau = 12, 8, 0.02
    st, L = make(orc, N, M, tau, rc_frac=1.0, action=action, seed=5)
    rng = np.random.default_rng(2)
    for trial in range(12):
        m, p = int(rng.integers(M)), int(rng.integers(N))
        b = st.beads()
        new = b[m, p] + rng.normal(scale=0.4, size=3)
        new -= L * np.floor(new / L + 0.5)
        a_tot0, a_b0 = st.total_action(), st.action_bead(m, p)
        g_tot0, g_b0 = st.gradV2_slice(m), st.gradVnn2(m, p)
        st.update_bead(m, p, new)
        a_tot1, a_b1 = st.total_action(), st.action_bead(m, p)
        g_tot1, g_b1 = st.gradV2_slice(m), st.gradVnn2(m, p)
        # the three-body bead term carries the whole change of the slice's force sum
        assert abs((g_tot1 - g_tot0) - (g_b1 - g_b0)) <= 1e-10 * (abs(g_tot0) + abs(g_tot1))
        assert abs((a_tot1 - a_tot0) - (a_b1 - a_b0)) <= 1e-10 * (abs(a_tot0) + abs(a_tot1))
    # and the correction is really there: it moves the action
    st2, _ = make(orc, N, M, tau, rc_frac=1.0, action=dict(action, gradvfactor=(0.0, 0.0)), seed=5)
    assert st.total_action() != st2.total_action()


def test_energy_is_the_tau_lambda_derivative(orc):
    N, M, tau = 10, 12, 0.02
    seed = 3
    st, L = make(orc, N, M, tau, seed=seed)
    beads = st.beads()

    def action(t, lam):
        s = orc.State(L, M, N, t, lam, L)
        s.set_beads(beads)
        s.set_action(1, GSF["vfactor"])
        s.set_correction(GSF["gradvfactor"], GSF["period"])
        return s.total_action()

    h_t, h_l = 1e-4 * tau, 1e-4 * LAMBDA
    dA_dtau = (action(tau + h_t, LAMBDA) - action(tau - h_t, LAMBDA)) / (2 * h_t)
    dA_dlam = (action(tau, LAMBDA + h_l) - action(tau, LAMBDA - h_l)) / (2 * h_l)
    # the tail from the primitive estimator at the same positions: V_prim = A_prim/(tau M) + tail
    prim = orc.State(L, M, N, tau, LAMBDA, L)
    prim.set_beads(beads)
    tail = prim.energy()["V"] - prim.total_action() / (tau * M)
    expected = dA_dtau / M - LAMBDA / (tau * M) * dA_dlam + tail
    got = st.energy()["V"]
    assert abs(got - expected) <= 1e-8 * abs(expected)
    # the operator estimator averages the even slices only (period 2)
    vop = np.mean([st.V_slice(m)[1] for m in range(0, M, 2)]) + tail
    assert np.isclose(st.energy()["V_int"], vop, rtol=1e-12)


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_bare_list_and_correction(orc, shift):
    """barePotentialAction(beads, values) (src/action.cpp:647-689) is
    VFactor*shift*tau times a numpy pair sum of the table V over the slice;
    potentialActionCorrection (:696-716) scales as (shift*tau)^3 and, at shift 1,
    is the bead action minus its bare part (the bead action's correction being
    pinned by the checkMove invariant above)."""
    N, M, tau = 10, 8, 0.02
    st, L = make(orc, N, M, tau, seed=13)
    az = st.aziz
    b = st.beads()
    rng = np.random.default_rng(shift)
    sp = np.array([[int(rng.integers(M)), int(rng.integers(N))] for _ in range(10)], dtype=np.int32)
    st.set_action(shift, GSF["vfactor"])
    bare = np.array(st.bare_action_list(sp))
    cor = np.array([st.action_correction(int(m), int(p)) for m, p in sp])
    for k, (m, p) in enumerate(sp):
        seps = b[m, p] - np.delete(b[m], p, axis=0)
        seps -= L * np.floor(seps / L + 0.5)
        vnn = sum(az.V(s) for s in seps)
        assert np.isclose(bare[k], GSF["vfactor"][m % 2] * shift * tau * vnn, rtol=1e-12, atol=0)
    st.set_action(1, GSF["vfactor"])
    cor1 = np.array([st.action_correction(int(m), int(p)) for m, p in sp])
    assert np.allclose(cor, shift ** 3 * cor1, rtol=1e-12, atol=0)
    assert np.all(cor1[sp[:, 0] % 2 == 0] == 0.0) and np.any(cor1 != 0.0)
    bare1 = np.array(st.bare_action_list(sp))
    for k, (m, p) in enumerate(sp):
        assert np.isclose(bare1[k] + cor1[k], st.action_bead(int(m), int(p)), rtol=1e-13, atol=0)
