"""Pins the oracle's wavevectors (src/estimator.cpp:439-570), S(q)
(:3705-3737) and F(q,tau) (:3923-3961):
- q lists against a numpy enumeration in the same lexicographic order;
- S(q) against the phasor form sum_m |rho_q(m)|^2 / N (numpy), S(0) = N per
  slice, and S(G) = N for a perfect simple-cubic lattice;
- F(q,tau) against the phasor correlation sum_m Re rho(m) rho*(m+tau) / N,
  F(q,0) equal to the S(q) contribution, and F(tau) = F(M - tau)."""
import itertools
import math

import numpy as np
import pytest

from tests._oracle_util import DENSITY, LAMBDA, phasor_rho


def enum_q(L, nmax, qmag_max):
    out = []
    for n in itertools.product(*[range(-nmax[i], nmax[i] + 1) for i in range(3)]):
        q = np.array(n) * 2.0 * math.pi / L
        if math.sqrt(((0.0 + q[0] * q[0]) + q[1] * q[1]) + q[2] * q[2]) <= qmag_max:
            out.append(q)
    return np.array(out)


def test_qvectors(orc):
    L = orc.box_side(16, DENSITY)
    q = orc.qvectors(L, "max_int", "4 4 4")
    assert q.shape == (729, 3)
    assert np.array_equal(q, enum_q(L, [4, 4, 4], math.sqrt(3 * (4 * 2 * math.pi / L) ** 2)))
    L2 = orc.box_side(256, DENSITY)
    q2 = orc.qvectors(L2, "max_float", "3 0 0")
    nm = 1 + int(3.0 * L2 / 2.0 / math.pi)
    assert np.array_equal(q2, enum_q(L2, [nm] * 3, 3.0))
    assert q2.shape[0] == 5377
    q3 = orc.qvectors(L, "int", "1 0 0 0 2 -1")
    assert np.allclose(q3, np.array([[1, 0, 0], [0, 2, -1]]) * 2 * math.pi / L, rtol=0, atol=0)
    q4 = orc.qvectors(L, "float", "0.5 0.25 0.1")
    assert q4[0, 2] == float(np.float32(0.1))


def _state(orc, N, M, seed=139853):
    L = orc.box_side(N, DENSITY)
    st = orc.State(L, M, N, 0.004, LAMBDA, L)
    b = orc.gen_worldlines(seed, L, M, N, LAMBDA, 0.004, 0)
    st.set_beads(b)
    return st, b, L


def test_ssf_phasor_identity(orc):
    st, b, L = _state(orc, 12, 5)
    q = orc.qvectors(L, "max_int", "2 2 2")
    got = st.ssf(q)
    rho = phasor_rho(b, q)
    ref = (np.abs(rho) ** 2).sum(axis=1) / 12
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-10)
    i0 = np.where(np.all(q == 0, axis=1))[0][0]
    assert got[i0] == 5 * 12


def test_ssf_perfect_lattice(orc):
    n = 3
    N = n ** 3
    L = orc.box_side(N, DENSITY)
    a = L / n
    pts = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)]) * a - L / 2 + a / 4
    b = np.broadcast_to(pts, (2, N, 3)).copy()
    st = orc.State(L, 2, N, 0.004, LAMBDA, L)
    st.set_beads(b)
    G = np.array([[2 * math.pi / a, 0, 0], [0, 2 * math.pi / a, 2 * math.pi / a]])
    assert np.allclose(st.ssf(G), 2 * N, rtol=1e-12)


def test_isf_identities(orc):
    M, N = 6, 5
    st, b, L = _state(orc, N, M)
    q = np.array([[2 * math.pi / L, 0, 0], [0, 4 * math.pi / L, -2 * math.pi / L], [0.3, 0.1, 0.2]])
    got = st.isf(q)
    rho = phasor_rho(b, q)
    ref = np.zeros((3, M))
    for t in range(M):
        ref[:, t] = (rho * np.conj(np.roll(rho, -t, axis=1))).real.sum(axis=1) / N
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-10)
    assert np.allclose(got[:2, 0], st.ssf(q[:2]), rtol=1e-12)
    for t in range(1, M):
        assert got[0, t] == pytest.approx(got[0, M - t], rel=1e-12)


def test_lattice_rho_and_isf_helpers(orc):
    """The FP64 phasor helpers the full-size GPU tests compare against: the
    factored lattice density modes equal the direct sum exp(i q.r) (negative n_x
    included), and the FFT correlation equals the defining lag sum, the oracle's
    S(q) and F(q,tau) included."""
    from tests._oracle_util import isf_from_rho, lattice_rho
    M, N = 7, 9
    st, b, L = _state(orc, N, M)
    nvec = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [2, -3, 1], [-2, 3, -1], [0, 4, -4], [-4, 0, 2], [3, 3, 3]])
    q = nvec * 2.0 * math.pi / L
    rho = lattice_rho(b, nvec, L)
    assert np.allclose(rho, phasor_rho(b, q), rtol=1e-12, atol=1e-11)
    f = isf_from_rho(rho, N)
    direct = np.zeros((len(q), M))
    for t in range(M):
        direct[:, t] = (rho * np.conj(np.roll(rho, -t, axis=1))).real.sum(axis=1) / N
    assert np.allclose(f, direct, rtol=1e-12, atol=1e-10)
    assert np.allclose(f, st.isf(q), rtol=1e-10, atol=1e-9)
    assert np.allclose(f[:, 0], st.ssf(q), rtol=1e-10, atol=1e-9)
