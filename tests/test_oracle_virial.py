"""The oracle's virial and centroid-virial pair terms (src/action.cpp:1446-1805)
and VirialEnergyEstimator::accumulate (src/estimator.cpp:1086-1251), pinned by
what the mathematics fixes rather than by restating the reference's loops:

* rDOTgradUterm1 of a cluster that does not wrap equals the pair virial
  sum_{i<j} gradV(r_ij).r_ij (the bead form and the pair form agree when
  r_i - r_j is the minimum image);
* rDOTgradUterm2 of two beads is 2 gvf tau^3 lambda d2V(k) dVdr(k) |r| (the
  transverse part of the T-matrix drops out of T F for a pair);
* for beads whose pairs all lie beyond the potential minimum (dV/dr > 0, so
  |gradV| = dV/dr), the T-matrix is the Jacobian dF_i/dr_i: rDOTgradUterm2 and
  deltadotgradUterm2 against finite differences of the numpy forces;
* deltadotgradUterm1 against the window centroid written as a mean of the
  unwrapped worldline, bead counted twice;
* secondderivPotentialActionTau summed over slices equals the numerical second
  tau-derivative of the total action;
* the estimator: E_th equals the EnergyEstimator's E; with a window of one the
  centroid-virial E equals E_th; V_op is the even-slice mean plus the tail.
The cVCov2 key (src/estimator.cpp:1244) is pinned by reading only: the map
inserts the unregistered key at index 0, so K_op carries totEcv*beta*dEdB and
CvCov2 stays 0."""
import numpy as np
import pytest

from tests._oracle_util import DENSITY, LAMBDA, put_in_bc

GSF = dict(vfactor=(2.0 / 3.0, 4.0 / 3.0), gradvfactor=(0.0, 2.0 / 9.0), period=2)
LB = dict(vfactor=(1.0, 1.0), gradvfactor=(1.0 / 12.0, 1.0 / 12.0), period=2)


def state(orc, L, beads, tau, action=GSF):
    M, N = beads.shape[0], beads.shape[1]
    st = orc.State(L, M, N, tau, LAMBDA, L)
    st.set_beads(beads)
    st.set_action(1, action["vfactor"])
    st.set_correction(action["gradvfactor"], action["period"])
    return st


def forces(az, x, L):
    n = x.shape[0]
    F = np.zeros((n, 3))
    for i in range(n):
        for j in range(n):
            if i != j:
                F[i] += np.array(az.gradV(put_in_bc(x[i] - x[j], L)))
    return F


def test_pair_virial_identity(orc):
    """Beads inside a 5 A cube of a 30 A box: no minimum image wraps."""
    L, M, N, tau = 30.0, 4, 7, 0.02
    rng = np.random.default_rng(1)
    beads = rng.uniform(-2.5, 2.5, size=(M, N, 3)) + 1.0
    st = state(orc, L, beads, tau, LB)
    az = st.aziz
    out = st.virial_terms(3)
    for m in range(M):
        pair = 0.0
        for i in range(N):
            for j in range(i + 1, N):
                s = beads[m, i] - beads[m, j]
                pair += np.dot(az.gradV(s), s)
        assert np.isclose(out[m, 0], LB["vfactor"][m % 2] * tau * pair, rtol=1e-11, atol=0)


def test_two_beads_term2(orc):
    L, M, tau = 20.0, 4, 0.02
    rng = np.random.default_rng(2)
    beads = np.zeros((M, 2, 3))
    for m in range(M):
        beads[m, 0] = rng.uniform(-1, 1, 3)
        d = rng.normal(size=3)
        beads[m, 1] = beads[m, 0] + d / np.linalg.norm(d) * rng.uniform(2.2, 5.0)
    st = state(orc, L, beads, tau, LB)
    az = st.aziz
    out = st.virial_terms(2)
    for m in range(M):
        s = beads[m, 0] - beads[m, 1]
        r = np.linalg.norm(s)
        dvdr = np.dot(az.gradV(s), s) / r
        expect = 2.0 * LB["gradvfactor"][m % 2] * tau ** 3 * LAMBDA * az.grad2V(s) * dvdr * r
        assert np.isclose(out[m, 1], expect, rtol=1e-10, atol=0)


def lattice_beads(M, rng):
    """27 beads on a 3.6 A simple cubic lattice in a 10.8 A box, jittered by
    0.05 A: every minimum-image pair lies beyond 3.4 A, past the Aziz minimum
    (2.967 A) where dV/dr > 0, and no separation component comes near L/2, so
    the forces are smooth in every bead position."""
    g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], dtype=float) * 3.6 - 3.6
    return np.stack([g + rng.uniform(-0.05, 0.05, size=g.shape) for _ in range(M)]), 10.8


def force_on(az, x, i, L):
    F = np.zeros(3)
    for j in range(x.shape[0]):
        if j != i:
            F += np.array(az.gradV(put_in_bc(x[i] - x[j], L)))
    return F


@pytest.mark.parametrize("action", [GSF, LB])
def test_term2_is_the_force_jacobian(orc, action):
    M, tau, W = 4, 0.03, 2
    rng = np.random.default_rng(3)
    beads, L = lattice_beads(M, rng)
    st = state(orc, L, beads, tau, action)
    az = st.aziz
    out = st.virial_terms(W)
    h = 2e-3
    for m in range(M):
        gvf = action["gradvfactor"][m % 2]
        x = beads[m]
        F = forces(az, x, L)
        # window centroid: mean of the unwrapped worldline, this bead twice
        TF = np.zeros_like(F)
        for i in range(x.shape[0]):
            J = np.zeros((3, 3))
            for b in range(3):
                xp, xm = x.copy(), x.copy()
                xp[i, b] += h
                xm[i, b] -= h
                J[:, b] = (force_on(az, xp, i, L) - force_on(az, xm, i, L)) / (2 * h)
            TF[i] = J @ F[i]
        delta = window_delta(beads, L, m, W)
        pref = 2.0 * gvf * tau ** 3 * LAMBDA
        if gvf == 0.0:
            assert out[m, 1] == 0.0 and out[m, 3] == 0.0
            continue
        e_r, e_d = pref * np.sum(TF * x), pref * np.sum(TF * delta)
        assert abs(out[m, 1] - e_r) <= 2e-3 * pref * np.sum(np.abs(TF * x)), (m, out[m, 1], e_r)
        assert abs(out[m, 3] - e_d) <= 2e-3 * pref * np.sum(np.abs(TF * delta)), (m, out[m, 3], e_d)


def window_delta(beads, L, m, W):
    M, N = beads.shape[0], beads.shape[1]
    delta = np.zeros((N, 3))
    for i in range(N):
        pts = [beads[m, i].copy(), beads[m, i].copy()]
        fwd, bwd = beads[m, i].copy(), beads[m, i].copy()
        for g in range(1, W):
            fwd = fwd + put_in_bc(beads[(m + g) % M, i] - beads[(m + g - 1) % M, i], L)
            bwd = bwd + put_in_bc(beads[(m - g) % M, i] - beads[(m - g + 1) % M, i], L)
            pts += [fwd, bwd]
        delta[i] = put_in_bc(beads[m, i] - np.mean(pts, axis=0), L)
    return delta


@pytest.mark.parametrize("W", [1, 3, 5])
def test_delta_term1_window_centroid(orc, W):
    N, M, tau = 6, 12, 0.02
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(9, L, M, N, LAMBDA, tau, 1)
    st = state(orc, L, beads, tau)
    az = st.aziz
    out = st.virial_terms(W)
    for m in range(M):
        F = forces(az, beads[m], L)
        d = window_delta(beads, L, m, W)
        expect = GSF["vfactor"][m % 2] * tau * np.sum(F * d)
        assert np.isclose(out[m, 2], expect, rtol=1e-9, atol=1e-9 * np.sum(np.abs(F * d)))
        if W == 1:
            assert out[m, 2] == 0.0 and out[m, 3] == 0.0


def test_second_tau_derivative(orc):
    N, M, tau = 8, 10, 0.02
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(4, L, M, N, LAMBDA, tau, 1)
    out = state(orc, L, beads, tau).virial_terms(3)

    def action(t):
        return state(orc, L, beads, t).total_action()

    h = 1e-3 * tau
    d2 = (action(tau + h) - 2 * action(tau) + action(tau - h)) / h ** 2
    assert np.isclose(out[:, 5].sum(), d2, rtol=1e-5)
    st = state(orc, L, beads, tau)
    for m in range(M):
        g2 = st.gradV2_slice(m)
        gvf = GSF["gradvfactor"][m % 2]
        assert np.isclose(out[m, 4], gvf * tau ** 3 * LAMBDA * g2, rtol=1e-13, atol=0)


@pytest.mark.parametrize("action", [GSF, LB, dict(vfactor=(1.0, 1.0), gradvfactor=(0.0, 0.0), period=1)])
def test_estimator_identities(orc, action):
    N, M, tau = 8, 12, 0.02
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(6, L, M, N, LAMBDA, tau, 1)
    st = state(orc, L, beads, tau, action)
    e = st.energy()
    v5, v1 = st.virial_energy(5), st.virial_energy(1)
    assert np.isclose(v5["E_th"], e["E"], rtol=1e-13)
    assert np.isclose(v1["E"], v1["E_th"], rtol=1e-12)
    tail = st.energy()["V_int"] - np.mean([st.V_slice(m)[1] for m in range(0, M, action["period"])])
    vop = np.mean([st.V_slice(m).sum() for m in range(0, M, 2)]) + tail
    assert np.isclose(v5["V_op"], vop, rtol=1e-12)
    assert np.isclose(v5["K_cv"] + v5["V_cv"], v5["E"], rtol=1e-12)
    # the unregistered cVCov2 key lands on index 0 (K_op)
    assert v5["CvCov2"] == 0.0
    beta = M * tau
    assert np.isclose(v5["K_op"], v5["E"] - v5["V_op"] + v5["E"] * beta * v5["dEdB"], rtol=1e-12)
    assert np.isclose(v5["E/N"], v5["E"] / N, rtol=1e-14)
