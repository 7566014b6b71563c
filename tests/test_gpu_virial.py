"""GPU parity of the virial pair terms (src/action.cpp:1446-1786, :776-787) and
of VirialEnergyEstimator::accumulate (src/estimator.cpp:1086-1251) against the
oracle on identical seeded worldlines, with periodic links and with a
permutation across the imaginary-time boundary.

Tolerances, relative to the largest magnitude of the compared column: 1e-9 in
the FP64 closed-form mode (dV/dr and d2V/dr2 at x = k dr / rm against the tables'
accumulated radii), 1e-6 in the FP32-transcendental mode (exp through v_exp_f32
with a split exponent), 1e-11 in the table mode (per-pair values from the same
tables, summation order differs).  The estimator's potential columns carry V(slice),
whose FP32-transcendental mode is good to 1e-6 (TOL_E)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA = 0.02198, 6.0596
TOL = {0: 1e-6, 1: 1e-9, 2: 1e-11}
TOL_E = {0: 1e-6, 1: 1e-9, 2: 1e-11}
GSF = dict(vfactor=(2.0 / 3.0, 4.0 / 3.0), gradvfactor=(0.0, 2.0 / 9.0), period=2)
LB = dict(vfactor=(1.0, 1.0), gradvfactor=(1.0 / 12.0, 1.0 / 12.0), period=2)
PRIM = dict(vfactor=(1.0, 1.0), gradvfactor=(0.0, 0.0), period=1)


@pytest.fixture(scope="module")
def gpu():
    import pimc_amd
    return pimc_amd


def make(orc, gpu, N, M, tau, mode, action, seed=77, swap=False):
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(seed, L, M, N, LAMBDA, tau, 1)
    ref = orc.State(L, M, N, tau, LAMBDA, L)
    ref.set_beads(beads)
    ref.set_action(1, action["vfactor"])
    ref.set_correction(action["gradvfactor"], action["period"])
    ctx = gpu.Context(L, M, N, tau, LAMBDA, L, v_mode=mode)
    ctx.upload_path(beads)
    ctx.set_action_factors(action["vfactor"], action["gradvfactor"], action["period"])
    nxt = None
    if swap:
        # particles 0 and 1 exchange across the M-1 -> 0 link
        nxt = np.zeros((M, N, 2), dtype=np.int32)
        prv = np.zeros((M, N, 2), dtype=np.int32)
        for m in range(M):
            for i in range(N):
                nxt[m, i] = ((m + 1) % M, i)
                prv[m, i] = ((m - 1) % M, i)
        nxt[M - 1, 0], nxt[M - 1, 1] = (0, 1), (0, 0)
        prv[0, 1], prv[0, 0] = (M - 1, 0), (M - 1, 1)
        ref.set_topology(np.ones((M, N), dtype=np.uint8), np.full(M, N, dtype=np.int32), nxt, prv)
    return ref, ctx, nxt


def close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return np.all(np.abs(a - b) <= tol * scale)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("name,action", [("gsf", GSF), ("li_broughton", LB), ("primitive", PRIM)])
def test_virial_terms(orc, gpu, mode, name, action):
    ref, ctx, _ = make(orc, gpu, 24, 20, 0.02, mode, action)
    for W in (1, 5):
        g, r = ctx.virial_terms(W), ref.virial_terms(W)
        assert close(g, r, TOL[mode]), (name, mode, W, np.abs(g - r).max(axis=0), np.abs(r).max(axis=0))
        if action["gradvfactor"][0] == 0.0:
            assert np.all(g[0::2, 1] == 0.0) and np.all(g[0::2, 4] == 0.0)


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("swap", [False, True])
def test_virial_energy(orc, gpu, mode, swap):
    ref, ctx, nxt = make(orc, gpu, 32, 24, 0.015, mode, GSF, swap=swap)
    for W in (1, 3, 5):
        g = ctx.virial_energy(W, mu=0.7, next_links=nxt)
        r = ref.virial_energy(W, mu=0.7)
        for k in r:
            assert abs(g[k] - r[k]) <= TOL_E[mode] * max(abs(r[k]), 1.0), (W, k, g[k], r[k])
        assert g["CvCov2"] == 0.0


@pytest.mark.parametrize("mode", [0, 2])
def test_virial_terms_larger(orc, gpu, mode):
    """Two workgroups per slice (N = 300 > 256) and LDS staging."""
    ref, ctx, _ = make(orc, gpu, 300, 6, 0.01, mode, LB, seed=5)
    g, r = ctx.virial_terms(3), ref.virial_terms(3)
    assert close(g, r, TOL[mode])


def test_virial_rejects_open_links(orc, gpu):
    ref, ctx, nxt = make(orc, gpu, 8, 6, 0.02, 2, GSF, swap=True)
    bad = nxt.copy()
    bad[2, 3] = (3, 4)   # two beads link to (3, 4)
    with pytest.raises(gpu.PimcError):
        ctx.virial_terms(2, next_links=bad)
    with pytest.raises(gpu.PimcError):
        ctx.virial_terms(0)


def test_virial_links_change_between_calls(orc, gpu):
    """The previous call's links stay on the device: a call with other links,
    or with the first links again after a refused call, uses the links it is
    given (bit-identical to a context that saw only those links)."""
    M, N = 24, 32
    _, ctx, swap = make(orc, gpu, N, M, 0.015, 2, GSF, swap=True)
    per = np.zeros((M, N, 2), dtype=np.int32)
    for m in range(M):
        per[m, :, 0] = (m + 1) % M
        per[m, :, 1] = np.arange(N)
    a = ctx.virial_terms(3, next_links=swap)
    a_again = ctx.virial_terms(3, next_links=swap)
    b = ctx.virial_terms(3, next_links=per)
    bad = swap.copy()
    bad[2, 3] = (3, 4)
    with pytest.raises(gpu.PimcError):
        ctx.virial_terms(3, next_links=bad)
    a_after = ctx.virial_terms(3, next_links=swap)
    _, only_swap, _ = make(orc, gpu, N, M, 0.015, 2, GSF)
    _, only_per, _ = make(orc, gpu, N, M, 0.015, 2, GSF)
    a_ref = only_swap.virial_terms(3, next_links=swap)
    b_ref = only_per.virial_terms(3, next_links=per)
    assert not np.array_equal(a_ref, b_ref)   # the exchange changes the link-window terms
    for x in (a, a_again, a_after):
        assert np.array_equal(x, a_ref)
    assert np.array_equal(b, b_ref)
