"""pimc_ssf_noalloc against pimc_ssf through the C-ABI: the same S(q), bit for
bit, on two calls in a row (the second with a smaller q list), for a q list
mixing lattice and general wavevectors and longer than one scratch run."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853


def test_ssf_noalloc_matches_ssf(orc):
    import pimc_amd
    from pimc_amd.workloads import lattice_q
    N, M = 256, 167
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(SEED0, L, M, N, LAMBDA, 0.004, 0)
    ctx = pimc_amd.Context(L, M, N, 0.004, LAMBDA, L)
    ctx.upload_path(beads)
    lat = lattice_q(L, qmag_max=4.5)                       # 17 845 lattice q
    assert lat.shape[0] == 17845
    rng = np.random.default_rng(4)
    gen = rng.uniform(-2.0, 2.0, size=(24, 3))             # general q (minimum-image pair sums)
    # runs: 2000 lattice, 16 general, 15 845 lattice (longer than the 12 264 q one
    # 32 MiB run holds at M = 167: split in two), 8 general
    q1 = np.concatenate([lat[:2000], gen[:16], lat[2000:], gen[16:]])
    print(f"call 1: {q1.shape[0]} q", flush=True)
    a1 = ctx.ssf_noalloc(q1)
    b1 = ctx.ssf(q1)
    assert np.array_equal(a1, b1)
    i0 = int(np.where(np.all(q1 == 0, axis=1))[0][0])
    assert a1[i0] == pytest.approx(N * M, rel=1e-12)
    # second call, smaller, reusing the scratch left by the first
    q2 = np.concatenate([gen[:3], lat[::397], [[2 * math.pi / L, 0, 0]]])
    print(f"call 2: {q2.shape[0]} q", flush=True)
    a2 = ctx.ssf_noalloc(q2)
    b2 = ctx.ssf(q2)
    assert np.array_equal(a2, b2)
    # against the oracle on a few of them
    st = orc.State(L, M, N, 0.004, LAMBDA, L)
    st.set_beads(beads)
    ref = st.ssf(q2[:6])
    assert np.all(np.abs(a2[:6] - ref) <= 1e-5 * np.maximum(ref, M))
    assert ctx.ssf_noalloc(np.zeros((0, 3))).shape == (0,)
    ctx.close()
