"""Whole Markov runs through the device: pimc_chain's Staging / Bisection /
CenterOfMass moves (Delta V on the GPU) against the oracle's replay of
src/move.cpp on the same worldlines.

* Same seed: every Metropolis decision and every proposed position must match
  the oracle's run, so the final worldlines are bitwise equal and the reported
  dU agree to 1e-9 relative (FP64 closed form and table modes: the device dU
  differs from the host's only by summation order, ~1e-15 relative, so a
  decision flip would need u within that of exp(-dU)).
* Independent seeds, FP32 pair terms on the device: the mean energies per
  particle (thermodynamic E and V, centroid-virial E) of a long oracle chain and
  of 64 device-resident chains agree within 3 combined standard errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853


@pytest.fixture(scope="module")
def gpu():
    import pimc_amd
    return pimc_amd


def make_pair(orc, gpu, N, M, tau, rc_frac, kind, mode, seed_beads=SEED0):
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(seed_beads, L, M, N, LAMBDA, tau, kind)
    ref = orc.State(L, M, N, tau, LAMBDA, rc_frac * L)
    ref.set_beads(beads)
    ctx = gpu.Context(L, M, N, tau, LAMBDA, rc_frac * L, v_mode=mode)
    ctx.upload_path(beads)
    return ref, ctx, L


@pytest.mark.parametrize("mode,rc_frac,N,M", [(2, 1.0, 16, 40), (1, 0.5, 32, 64), (2, 0.5, 64, 48)])
def test_markov_run_matches_oracle(orc, gpu, mode, rc_frac, N, M):
    ref, ctx, L = make_pair(orc, gpu, N, M, 0.01, rc_frac, kind=1, mode=mode)
    seed, Mbar, levels, com_delta = 4242, 8, 3, 0.5
    moves = orc.Moves(ref, Mbar, levels, True, 1, com_delta)
    rng = orc.MTRand(seed)
    chain = gpu.Chain(ctx, seed, Mbar, levels, True, 1, com_delta)
    kinds = ("staging", "bisection", "com")
    n_acc = 0
    for k in range(600):
        kind = kinds[k % 3]
        acc_r, info_r = getattr(moves, kind)(rng)
        acc_g, info_g = chain.move(kind)
        assert bool(acc_r) == acc_g, f"move {k} ({kind}): oracle {acc_r} device {acc_g}"
        assert info_g[2] == info_r[2]
        assert abs(info_g[1] - info_r[1]) <= 1e-9 * (1.0 + abs(info_r[1])), (k, kind, info_r, info_g)
        n_acc += acc_g
    print(f"mode {mode} N={N} M={M} rc={rc_frac}L: {n_acc}/600 accepted, decisions identical")
    assert 0 < n_acc < 600
    assert np.array_equal(chain.positions(), ref.beads())
    assert np.array_equal(ctx.download_path(), ref.beads())
    a_g, a_r = ctx.potentialAction(), ref.total_action()
    assert abs(a_g - a_r) <= 1e-9 * abs(a_r)


def blocked(x, nb=20):
    x = np.asarray(x)[: len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
    return x.mean(), x.std(ddof=1) / np.sqrt(nb)


def test_markov_energy_independent_seeds(orc, gpu):
    """Independent runs agree within statistical error bars.  N=16, M=20,
    tau=0.02 (T = 2.5 K), Staging / Bisection / CenterOfMass in rotation.
    Oracle: one chain of 250 000 sweeps after 2 000 of burn-in, sampled every 4
    sweeps.  Device: 64 chains with their own worldlines and seeds run on the
    device (one launch per 4 sweeps of every chain), FP32 pair terms, each
    sampled after every launch.  Compared: the thermodynamic E/N and V/N of
    EnergyEstimator and the centroid-virial E/N of VirialEnergyEstimator
    (window 5), each within 3 combined standard errors (20 blocks per side);
    the oracle's standard error of E/N is about 0.1 K."""
    N, M, tau = 16, 20, 0.02
    Mbar, levels, com_delta = 6, 2, 0.5
    kinds = ("staging", "bisection", "com")
    ref, _, L = make_pair(orc, gpu, N, M, tau, 1.0, kind=1, mode=0)
    moves = orc.Moves(ref, Mbar, levels, True, 1, com_delta)
    rng = orc.MTRand(777)
    r_e, r_v, r_vir = [], [], []
    for k in range(252000):
        for kind in kinds:
            getattr(moves, kind)(rng)
        if k >= 2000 and k % 4 == 0:
            e = ref.energy()
            r_e.append(e["E/N"])
            r_v.append(e["V/N"])
            if k % 40 == 0:
                r_vir.append(ref.virial_energy(5)["E/N"])
    n_ch = 64
    ctxs, chains = [], []
    for c in range(n_ch):
        ctx = gpu.Context(L, M, N, tau, LAMBDA, L, v_mode=0)
        ctx.upload_path(orc.gen_worldlines(SEED0 + 500 + c, L, M, N, LAMBDA, tau, 1))
        ctxs.append(ctx)
        chains.append(gpu.Chain(ctx, 20000 + c, Mbar, levels, True, 1, com_delta))
    gpu.run_chains_device(chains, [k % 3 for k in range(6000)])          # burn-in
    g_e, g_v, g_vir = [], [], []
    rounds = 1000
    for r in range(rounds):
        gpu.run_chains_device(chains, [k % 3 for k in range(12)])
        es = [ctx.energy() for ctx in ctxs]
        g_e.append(np.mean([e["E/N"] for e in es]))
        g_v.append(np.mean([e["V/N"] for e in es]))
        if r % 10 == 0:
            g_vir.append(np.mean([ctx.virial_energy(5)["E/N"] for ctx in ctxs]))
    for name, a, b in (("E/N", r_e, g_e), ("V/N", r_v, g_v), ("E_vir/N", r_vir, g_vir)):
        (mr, sr), (mg, sg) = blocked(a), blocked(b)
        print(f"{name}: oracle {mr:.4f} +- {sr:.4f} ({len(a)} samples)  device {mg:.4f} +- {sg:.4f} "
              f"({len(b)} rounds x {n_ch} chains)", flush=True)
        assert sr > 0 and sg > 0
        assert abs(mr - mg) <= 3.0 * np.hypot(sr, sg), name
    assert blocked(r_e)[1] <= 0.15


@pytest.mark.parametrize("mode,rc_frac,N,M", [(2, 1.0, 16, 40), (2, 0.25, 128, 32), (1, 0.5, 64, 48)])
def test_device_resident_run_matches_oracle(orc, gpu, mode, rc_frac, N, M):
    """pimc_chain_run_device: proposals, draws and decisions on the device in
    one launch, continuing the host stream (host moves before and after).  The
    decisions and stage lengths equal the oracle's; dU and the final positions
    agree to 1e-9 (the device's log / exp may differ from the host's in the
    last ulp); decisions inside the 1e-10 guard band are counted."""
    ref, ctx, L = make_pair(orc, gpu, N, M, 0.01, rc_frac, kind=1, mode=mode)
    seed, Mbar, levels, com_delta = 99, 8, 3, 0.5
    moves = orc.Moves(ref, Mbar, levels, True, 1, com_delta)
    rng = orc.MTRand(seed)
    chain = gpu.Chain(ctx, seed, Mbar, levels, True, 1, com_delta)
    kinds = ("staging", "bisection", "com")
    seq = [kinds[k % 3] for k in range(900)]
    for kind in seq[:30]:                           # host-driven first
        acc_r, _ = getattr(moves, kind)(rng)
        acc_g, _ = chain.move(kind)
        assert bool(acc_r) == acc_g
    acc_d, info_d, near = chain.run_device(seq[30:870])
    for k, kind in enumerate(seq[30:870]):
        acc_r, info_r = getattr(moves, kind)(rng)
        assert bool(acc_r) == bool(acc_d[k]), f"device move {k} ({kind}): oracle {acc_r} device {acc_d[k]}"
        assert info_d[k, 2] == info_r[2] and info_d[k, 0] == 1.0
        assert abs(info_d[k, 1] - info_r[1]) <= 1e-9 * (1.0 + abs(info_r[1])), (k, kind, info_r, info_d[k])
    for kind in seq[870:]:                          # and the stream continues on the host
        acc_r, _ = getattr(moves, kind)(rng)
        acc_g, _ = chain.move(kind)
        assert bool(acc_r) == acc_g
    print(f"device run mode {mode} N={N} M={M}: {int(acc_d.sum())}/840 accepted, {near} near the threshold")
    assert 0 < acc_d.sum() < 840
    assert np.allclose(ctx.download_path(), ref.beads(), rtol=0, atol=1e-9)
    a_g, a_r = ctx.potentialAction(), ref.total_action()
    assert abs(a_g - a_r) <= 1e-9 * abs(a_r)


def test_many_chains_one_launch(orc, gpu):
    """pimc_chain_run_device_many: six independent chains (their own worldlines
    and seeds, uneven N and M) in one launch equal the same chains run one at a
    time (bitwise: one workgroup each, same arithmetic), and the first equals the
    oracle's replay decision for decision."""
    cfg = [(16, 40), (24, 32), (32, 48), (16, 64), (40, 24), (20, 36)]
    kinds = [k % 3 for k in range(240)]
    runs = []
    for rep in range(2):
        chains, ctxs = [], []
        for c, (N, M) in enumerate(cfg):
            ref, ctx, L = make_pair(orc, gpu, N, M, 0.01, 1.0, kind=1, mode=2, seed_beads=SEED0 + c)
            ctxs.append((ref, ctx))
            chains.append(gpu.Chain(ctx, 1000 + c, 6, 2, True, 1, 0.4))
        if rep == 0:
            acc, info, near = gpu.run_chains_device(chains, kinds)
        else:
            acc = np.zeros((len(cfg), len(kinds)), bool)
            info = np.zeros((len(cfg), len(kinds), 3))
            for c, ch in enumerate(chains):
                a, i, _ = ch.run_device(kinds)
                acc[c], info[c] = a, i
        runs.append((acc, info, [ctx.download_path() for _, ctx in ctxs], ctxs))
    (a0, i0, p0, ctxs), (a1, i1, p1, _) = runs
    assert np.array_equal(a0, a1) and np.array_equal(i0, i1)
    for x, y in zip(p0, p1):
        assert np.array_equal(x, y)
    ref = ctxs[0][0]
    moves = orc.Moves(ref, 6, 2, True, 1, 0.4)
    rng = orc.MTRand(1000)
    names = ("staging", "bisection", "com")
    for k, kd in enumerate(kinds):
        acc_r, info_r = getattr(moves, names[kd])(rng)
        assert bool(acc_r) == bool(a0[0, k]), k
        assert abs(i0[0, k, 1] - info_r[1]) <= 1e-9 * (1.0 + abs(info_r[1]))
    assert np.allclose(p0[0], ref.beads(), rtol=0, atol=1e-9)
    print(f"6 chains x 240 moves in one launch; acceptance {a0.mean():.3f}")
