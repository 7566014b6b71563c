"""The product-side worldline generator (pimc_amd/worldlines.py) reproduces the
oracle's seeded generator bit for bit (same MTRand stream and transforms)."""
import numpy as np
import pytest

from pimc_amd import worldlines


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("N,M", [(16, 125), (33, 17), (256, 12)])
def test_generator_matches_oracle(orc, kind, N, M):
    L = orc.box_side(N, 0.02198)
    for seed in (139853, 139854):
        a = worldlines.generate(seed, L, M, N, 6.0596, 0.004, kind)
        b = orc.gen_worldlines(seed, L, M, N, 6.0596, 0.004, kind)
        assert np.array_equal(a, b)


def test_lattice_q_matches_reference_order(orc):
    from pimc_amd.workloads import box_side, lattice_q
    for N, kind, text, kw in ((16, "max_int", "4 4 4", dict(nmax=4)), (256, "max_float", "3 0 0", dict(qmag_max=3.0))):
        L = box_side(N)
        assert np.array_equal(lattice_q(L, **kw), orc.qvectors(L, kind, text))


@pytest.mark.parametrize("kind", [0, 1])
def test_native_generator_matches_oracle(orc, kind):
    from pimc_amd import api
    for (N, M) in ((16, 125), (256, 20), (7, 3)):
        L = orc.box_side(N, 0.02198)
        a = api.gen_worldlines(139853 + N, L, M, N, 6.0596, 0.004, kind)
        b = orc.gen_worldlines(139853 + N, L, M, N, 6.0596, 0.004, kind)
        assert np.array_equal(a, b)
