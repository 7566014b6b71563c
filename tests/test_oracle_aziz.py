"""Pins the oracle's Aziz HFDHE2 potential (src/potential.cpp:1741-1909) and the
TabulatedPotential direct lookup (include/potential.h:163-260):
- valueV / valuedVdr / valued2Vdr2 against the closed form evaluated in mpmath
  at 40 digits, and against mpmath numerical derivatives of that closed form;
- spot values of the 1979 curve (V(rm) = -10.79975 K, sigma = 2.638504 A,
  minimum at 2.967409 A) found by mpmath root finding;
- the table: entry n equals valueV at the n-fold *accumulated* radius
  (np.cumsum replays r += dr), and V(sep) = lookupV[int(|sep|/dr)] with 0 for
  k <= 0 and k >= tableLength;
- the tail correction against 2*pi * integral_rc^inf r^2 V_disp(r) dr."""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 40
P1979 = dict(epsilon=10.8, rm=2.9673, D=1.241314, alpha=13.353384, beta=0.0,
             C6=1.3732412, C8=0.4253785, C10=0.1781, A=0.5448504e6)


def v_mp(r, p=P1979):
    x = mp.mpf(r) / p["rm"]
    rep = p["A"] * mp.e ** (-p["alpha"] * x + p["beta"] * x * x)
    if x < mp.mpf("1e-7"):
        return mp.mpf(0)
    if x < mp.mpf("0.01"):
        return p["epsilon"] * rep
    F = mp.e ** (-(p["D"] / x - 1) ** 2) if x < p["D"] else mp.mpf(1)
    att = -(p["C6"] * x ** -6 + p["C8"] * x ** -8 + p["C10"] * x ** -10) * F
    return p["epsilon"] * (rep + att)


@pytest.fixture(scope="module")
def az(orc):
    return orc.Aziz(1979, max_sep=8.0, rc=7.0, derivs=True)


def test_params(az):
    p = az.params()
    for k, v in P1979.items():
        assert p[k] == v


def test_valueV_closed_form(az):
    rs = np.concatenate([np.linspace(0.005, 0.03, 13), np.linspace(0.03, 12.0, 400)])
    for r in rs:
        ref = v_mp(r)
        got = az.valueV(float(r))
        assert abs(got - float(ref)) <= 1e-13 * max(1.0, abs(float(ref))), r


def test_derivatives_against_numerical(az):
    for r in np.linspace(1.0, 9.0, 41):
        d1 = mp.diff(lambda t: v_mp(t), mp.mpf(float(r)))
        d2 = mp.diff(lambda t: v_mp(t), mp.mpf(float(r)), 2)
        assert abs(az.valuedVdr(float(r)) - float(d1)) <= 1e-10 * max(1.0, abs(float(d1)))
        assert abs(az.valued2Vdr2(float(r)) - float(d2)) <= 1e-9 * max(1.0, abs(float(d2)))


def test_spot_values(az):
    assert abs(az.valueV(2.9673) - (-10.79975)) < 5e-6
    sigma = mp.findroot(lambda t: v_mp(t), 2.6)
    rmin = mp.findroot(lambda t: mp.diff(v_mp, t), 2.96)
    assert abs(float(sigma) - 2.638504) < 1e-6
    assert abs(float(rmin) - 2.967409) < 1e-6
    assert abs(az.valueV(float(sigma))) < 1e-10


def test_table_is_accumulated_grid(az):
    n = az.table_length
    dr = az.dr
    assert dr == 1.0e-6 * 2.9673
    assert n == int(8.0 / dr)
    r = np.concatenate([[0.0], np.cumsum(np.full(n - 1, dr))])
    tab = az.table(0)
    rng = np.random.default_rng(5)
    for k in np.concatenate([[0, 1, 2, 3370, 3371, 10000, n - 1], rng.integers(0, n, 300)]):
        assert tab[k] == az.valueV(float(r[k])), k
    # the accumulated radius drifts from k*dr; the table follows the drift
    assert r[n - 1] != (n - 1) * dr


def test_direct_lookup(az):
    n = az.table_length
    dr = az.dr
    tab, tab1, tab2 = az.table(0), az.table(1), az.table(2)
    rng = np.random.default_rng(7)
    for _ in range(500):
        s = rng.normal(size=3) * rng.uniform(0.0, 4.0)
        r = math.sqrt(((0.0 + s[0] * s[0]) + s[1] * s[1]) + s[2] * s[2])
        k = int(r / dr)
        expect = 0.0 if (k <= 0 or k >= n) else tab[k]
        assert az.V(s) == expect
        if 0 < k < n:
            g = az.gradV(s)
            assert np.array_equal(g, (tab1[k] / r) * s)
            assert az.grad2V(s) == tab2[k]
    assert az.V([1e-7, 0, 0]) == 0.0
    assert az.V([9.0, 0, 0]) == 0.0


def test_tail_correction(orc):
    for rc in (4.0, 11.334, 22.668):
        a = orc.Aziz(1979, max_sep=2.0, rc=rc)
        p = P1979
        disp = lambda r: -p["epsilon"] * (p["C6"] * (p["rm"] / r) ** 6 + p["C8"] * (p["rm"] / r) ** 8
                                         + p["C10"] * (p["rm"] / r) ** 10)
        ref = 2 * mp.pi * mp.quad(lambda r: r * r * disp(r), [rc, mp.inf])
        assert abs(a.tailV - float(ref)) <= 1e-12 * abs(float(ref))


def test_branch_bins(az):
    n = az.table_length
    dr = az.dr
    r = np.concatenate([[0.0], np.cumsum(np.full(n - 1, dr))])
    x = r / 2.9673
    b = az.branch_bins()
    for i, thr in enumerate((1e-7, 0.01, 1.241314)):
        k = b[i]
        assert not (x[k] < thr) and (k == 0 or x[k - 1] < thr)
