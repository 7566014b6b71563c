"""Pins the oracle's MTRand (include/MersenneTwister.h:76-419) against numpy's
independent MT19937 (legacy RandomState seeding is init_genrand) and the C++
standard's known answer for std::mt19937 ([rand.predef]: 10000th output of the
default seed 5489 is 4123659995)."""
import math

import numpy as np


def test_known_answers(orc):
    mt = orc.MTRand(5489)
    v = [mt.randInt() for _ in range(10000)]
    assert v[-1] == 4123659995
    assert orc.MTRand(139853).randInt() == 3298052538


def test_stream_matches_numpy_mt19937(orc):
    for seed in (1, 139853, 139853 + 7, 0xFFFFFFFF):
        ref = np.random.RandomState(seed)._bit_generator.random_raw(2000)
        mt = orc.MTRand(seed)
        got = np.array([mt.randInt() for _ in range(2000)], dtype=np.uint64)
        assert np.array_equal(got, ref), seed


def test_transforms_closed_form(orc):
    seed = 139853 + 3
    raw = np.random.RandomState(seed)._bit_generator.random_raw(64)
    mt = orc.MTRand(seed)
    for k in range(32):
        assert mt.rand() == float(raw[k]) * (1.0 / 4294967295.0)
    for k in range(32, 64):
        assert mt.randExc() == float(raw[k]) * (1.0 / 4294967296.0)


def _masked(raw_iter, n):
    used = n
    for s in (1, 2, 4, 8, 16):
        used |= used >> s
    while True:
        i = next(raw_iter) & used
        if i <= n:
            return i


def test_randint_n_and_randnorm(orc):
    seed = 424242
    raw = iter(int(x) for x in np.random.RandomState(seed)._bit_generator.random_raw(20000))
    mt = orc.MTRand(seed)
    for n in (0, 1, 2, 7, 100, 255, 1000, 2**31 + 5):
        for _ in range(20):
            assert mt.randInt(n) == _masked(raw, n)
    # Marsaglia polar, written out independently
    for _ in range(200):
        while True:
            x = 2.0 * (next(raw) * (1.0 / 4294967295.0)) - 1.0
            y = 2.0 * (next(raw) * (1.0 / 4294967295.0)) - 1.0
            r = x * x + y * y
            if not (r >= 1.0 or r == 0.0):
                break
        expect = 1.5 + x * math.sqrt(-2.0 * math.log(r) / r) * 0.25
        assert mt.randNorm(1.5, 0.25) == expect


def test_save_load_continues_bitwise(orc):
    a = orc.MTRand(139853)
    for _ in range(1000):
        a.randInt()
    st = a.save()
    b = orc.MTRand(1)
    b.load(st)
    assert [a.randInt() for _ in range(700)] == [b.randInt() for _ in range(700)]
