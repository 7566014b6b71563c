"""GPU parity of the GSF / Li-Broughton gradient correction against the oracle
(src/action.cpp:456-806, src/setup.cpp:1232-1256) on identical seeded worldlines.

Tolerances: total action, bead actions, dU and energy rel <= 1e-6 in the FP32
mode (FP32 transcendentals in V and dV/dr), 1e-9 in the FP64 closed form, 1e-11
in the table mode (per-pair values from the host's tables, summation order
differs); Markov runs: identical decisions and bitwise-equal worldlines."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853
TOL = {0: 1e-6, 1: 1e-9, 2: 1e-11}
GSF = dict(vfactor=(2.0 / 3.0, 4.0 / 3.0), gradvfactor=(0.0, 2.0 / 9.0), period=2)
LB = dict(vfactor=(1.0, 1.0), gradvfactor=(1.0 / 12.0, 1.0 / 12.0), period=2)


@pytest.fixture(scope="module")
def gpu():
    import pimc_amd
    return pimc_amd


def pair(orc, gpu, N, M, tau, rc_frac, mode, action, kind=1, seed=SEED0):
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(seed, L, M, N, LAMBDA, tau, kind)
    ref = orc.State(L, M, N, tau, LAMBDA, rc_frac * L)
    ref.set_beads(beads)
    ref.set_action(1, action["vfactor"])
    ref.set_correction(action["gradvfactor"], action["period"])
    ctx = gpu.Context(L, M, N, tau, LAMBDA, rc_frac * L, v_mode=mode)
    ctx.upload_path(beads)
    ctx.set_action_factors(action["vfactor"], action["gradvfactor"], action["period"])
    return ref, ctx, beads, L


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("name,action", [("gsf", GSF), ("li_broughton", LB)])
def test_total_action_and_energy(orc, gpu, mode, name, action):
    ref, ctx, _, _ = pair(orc, gpu, 32, 40, 0.01, 1.0, mode, action)
    a, b = ctx.potentialAction(), ref.total_action()
    print(f"{name} mode {mode}: total action rel err {abs(a - b) / abs(b):.2e}")
    assert abs(a - b) <= TOL[mode] * abs(b)
    eg, er = ctx.energy(), ref.energy()
    for k in ("K", "V", "V_int", "E", "E/N"):
        assert abs(eg[k] - er[k]) <= TOL[mode] * max(abs(er[k]), 1.0), (k, eg[k], er[k])


@pytest.mark.parametrize("N,M", [(33, 20), (200, 30), (256, 167)])
def test_gradv_squared_sizes(orc, gpu, N, M):
    """gradVSquared through each unordered pair once (mixed mode, slices of at
    most 256 beads), odd and even sizes, against the oracle's GSF action."""
    ref, ctx, _, _ = pair(orc, gpu, N, M, 0.004, 1.0, 0, GSF)
    a, b = ctx.potentialAction(), ref.total_action()
    assert abs(a - b) <= TOL[0] * abs(b), (N, M, a, b)


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("N,M,rc_frac", [(24, 32, 1.0), (256, 24, 0.25)])
def test_bead_actions_and_delta(orc, gpu, mode, N, M, rc_frac):
    """potentialAction(vector) at shift 1 and a segment move's dU, with and
    without a cutoff (rc = L/4: 4 cells per side, lists restricted to r < rc)."""
    ref, ctx, beads, L = pair(orc, gpu, N, M, 0.01, rc_frac, mode, GSF)
    rng = np.random.default_rng(4)
    scale = abs(ref.total_action())
    for trial in range(4):
        seg = int(rng.integers(3, 9))
        start, p = int(rng.integers(M)), int(rng.integers(N))
        sp = np.array([[(start + k) % M, p] for k in range(seg)], dtype=np.int32)
        a_ref = ref.action_list(sp)
        a_gpu = ctx.action_list(sp)
        assert abs(a_gpu - a_ref) <= TOL[mode] * scale, (trial, a_gpu, a_ref)
        new = beads[sp[:, 0], sp[:, 1]] + rng.normal(scale=0.3, size=(seg, 3))
        new -= L * np.floor(new / L + 0.5)
        out = ctx.delta_action(sp, new)
        for k in range(seg):
            ref.update_bead(int(sp[k, 0]), int(sp[k, 1]), new[k])
        a_new = ref.action_list(sp)
        for k in range(seg):
            ref.update_bead(int(sp[k, 0]), int(sp[k, 1]), beads[sp[k, 0], sp[k, 1]])
        assert abs(out[0] - a_ref) <= TOL[mode] * scale
        assert abs(out[1] - a_new) <= TOL[mode] * scale
        assert abs(out[2] - (a_new - a_ref)) <= TOL[mode] * scale


def test_shift_above_one_has_no_correction(orc, gpu):
    ref, ctx, _, _ = pair(orc, gpu, 24, 32, 0.01, 1.0, 2, GSF)
    sp = np.array([[3, 5], [5, 5], [7, 5]], dtype=np.int32)   # odd slices: gradVFactor = 2/9
    ref.set_action(2, GSF["vfactor"])
    r2 = ref.action_list(sp)
    ref.set_action(1, GSF["vfactor"])
    r1 = ref.action_list(sp)
    a1, a2 = ctx.action_list(sp, shift=1), ctx.action_list(sp, shift=2)
    assert abs(a2 - r2) <= 1e-11 * abs(r2)
    assert abs(a1 - r1) <= 1e-11 * abs(r1)
    # the bare part scales with shift tau, so a1 - a2/2 is the shift-1 correction
    # alone (gradVnnSquared carries cross terms: its sign is not fixed)
    cor_ref = r1 - r2 / 2.0
    assert cor_ref != 0.0
    assert abs((a1 - a2 / 2.0) - cor_ref) <= 1e-9 * abs(cor_ref)


def test_gsf_markov_run_matches_oracle(orc, gpu):
    ref, ctx, _, _ = pair(orc, gpu, 16, 40, 0.01, 1.0, 2, GSF)
    seed, Mbar, levels, com_delta = 4242, 8, 3, 0.5
    moves = orc.Moves(ref, Mbar, levels, True, 1, com_delta)
    rng = orc.MTRand(seed)
    chain = gpu.Chain(ctx, seed, Mbar, levels, True, 1, com_delta)
    kinds = ("staging", "bisection", "com")
    n_acc = 0
    for k in range(450):
        kind = kinds[k % 3]
        acc_r, info_r = getattr(moves, kind)(rng)
        acc_g, info_g = chain.move(kind)
        assert bool(acc_r) == acc_g, f"move {k} ({kind}): oracle {acc_r} device {acc_g}"
        assert abs(info_g[1] - info_r[1]) <= 1e-9 * (1.0 + abs(info_r[1]))
        n_acc += acc_g
    print(f"GSF Markov run: {n_acc}/450 accepted, decisions identical")
    assert np.array_equal(chain.positions(), ref.beads())


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("shift", [1, 2])
@pytest.mark.parametrize("name,action", [("gsf", GSF), ("li_broughton", LB)])
def test_bare_action_and_correction_parts(orc, gpu, mode, shift, name, action):
    """barePotentialAction(beads, values) and potentialActionCorrection(bead)
    (src/action.cpp:647-716) per bead, at shift 1 and 2; at shift 1 their sum is
    the bead action the moves use."""
    ref, ctx, _, _ = pair(orc, gpu, 24, 32, 0.01, 1.0, mode, action)
    ref.set_action(shift, action["vfactor"])
    rng = np.random.default_rng(9)
    sp = np.array([[int(rng.integers(32)), int(rng.integers(24))] for _ in range(12)], dtype=np.int32)
    bare, cor = ctx.action_parts(sp, shift=shift)
    bare_ref = np.array(ref.bare_action_list(sp))
    cor_ref = np.array([ref.action_correction(int(m), int(p)) for m, p in sp])
    scale = np.abs(bare_ref).max()
    assert np.all(np.abs(bare - bare_ref) <= TOL[mode] * scale), (bare, bare_ref)
    cscale = max(np.abs(cor_ref).max(), 1e-300)
    assert np.any(cor_ref != 0.0)
    assert np.all(np.abs(cor - cor_ref) <= TOL[mode] * cscale), (cor, cor_ref)
    # even slices carry no GSF correction (gradVFactor[0] = 0)
    if name == "gsf":
        assert np.all(cor[sp[:, 0] % 2 == 0] == 0.0)
    if shift == 1:
        total = ctx.action_list(sp, shift=1)
        assert abs(total - (bare.sum() + cor.sum())) <= 1e-12 * abs(total)
    assert ctx.potentialActionCorrection(sp, shift=shift) == pytest.approx(cor.sum(), rel=1e-14, abs=0)
    assert np.array_equal(ctx.barePotentialAction(sp, shift=shift), bare)


@pytest.mark.parametrize("mode,N,M,rc_frac", [(2, 16, 40, 1.0), (1, 24, 32, 1.0), (2, 64, 24, 0.25)])
def test_gsf_device_resident_run(orc, gpu, mode, N, M, rc_frac):
    """The device-resident chain under the GSF action: every shift-1 bead action
    carries gradVFactor tau^3 lambda gradVnnSquared, evaluated on the device at
    the old and new positions; decisions equal the oracle's replay, dU agrees to
    1e-9 relative in the table mode and 1e-7 in the FP64 closed form (whose
    per-pair values differ from the table's by ~1e-13, amplified where dU is a
    small difference of large corrected bead actions; the FP32-transcendental
    mode's bead actions are checked against the action scale in
    test_bead_actions_and_delta), and the run continues the host stream."""
    ref, ctx, _, _ = pair(orc, gpu, N, M, 0.01, rc_frac, mode, GSF)
    seed, Mbar, levels, com_delta = 77, 8, 3, 0.5
    moves = orc.Moves(ref, Mbar, levels, True, 1, com_delta)
    rng = orc.MTRand(seed)
    chain = gpu.Chain(ctx, seed, Mbar, levels, True, 1, com_delta)
    kinds = [k % 3 for k in range(300)]
    names = ("staging", "bisection", "com")
    acc, info, near = chain.run_device(kinds)
    tol = 1e-9 if mode == 2 else 1e-7
    for k, kd in enumerate(kinds):
        acc_r, info_r = getattr(moves, names[kd])(rng)
        assert bool(acc_r) == bool(acc[k]), f"move {k} ({names[kd]}): oracle {acc_r} device {acc[k]}"
        assert abs(info[k, 1] - info_r[1]) <= tol * (1.0 + abs(info_r[1])), (k, info[k], info_r)
    for k in range(30):                               # the host continues the stream
        acc_r, _ = getattr(moves, names[k % 3])(rng)
        acc_g, _ = chain.move(names[k % 3])
        assert bool(acc_r) == acc_g
    print(f"GSF device run mode {mode} N={N}: {int(acc.sum())}/300 accepted, {near} near the threshold")
    assert np.allclose(ctx.download_path(), ref.beads(), rtol=0, atol=1e-9)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gradv_squared_per_slice(orc, gpu, mode):
    """pimc_gradv_squared: LocalAction::gradVSquared(slice) of every slice
    (src/action.cpp:1188-1223) against the oracle, with and without the
    correction set on the context (the values do not depend on the factors)."""
    ref, ctx, _, _ = pair(orc, gpu, 48, 12, 0.004, 1.0, mode, GSF, kind=0)
    g = ctx.gradVSquared()
    r = np.array([ref.gradV2_slice(m) for m in range(12)])
    assert np.all(np.abs(g - r) <= TOL[mode] * np.abs(r)), np.max(np.abs(g - r) / np.abs(r))
    ctx.set_action_factors((1.0, 1.0), (0.0, 0.0), 1)
    assert np.array_equal(ctx.gradVSquared(), g)


def test_path_version(orc, gpu):
    """pimc_path_version moves with every call that changes what per-slice values
    depend on, and only with those."""
    ref, ctx, beads, L = pair(orc, gpu, 16, 8, 0.004, 1.0, 2, GSF)
    v = ctx.path_version()
    ctx.potentialAction()
    ctx.energy()
    ctx.gradVSquared()
    ctx.delta_action([[2, 3]], beads[2, 3][None] + 0.1)          # evaluated, not committed
    ctx.ssf([[0.0, 0.0, 2 * np.pi / L]])
    ctx.download_path()
    assert ctx.path_version() == v
    steps = [lambda: ctx.update_beads([[1, 1]], beads[1, 1][None] + 0.05),
             lambda: ctx.delta_action([[2, 3]], beads[2, 3][None] + 0.1, commit=True),
             lambda: ctx.upload_path(beads),
             lambda: ctx.set_worm(head=(1, 2)),
             lambda: ctx.set_v_mode(0),
             lambda: ctx.set_action_factors((1.0, 1.0), (0.0, 0.0), 1)]
    for k, step in enumerate(steps):
        step()
        w = ctx.path_version()
        assert w > v, k
        v = w


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("N,M,rc_frac", [(24, 32, 1.0), (300, 16, 0.25)])
def test_delta_one_launch_matches_two(orc, gpu, mode, N, M, rc_frac):
    """With the gradient correction an untimed Delta V runs k_vnn's and
    k_gradvnn's workgroups in one launch, a timed one launches them apart (each
    kernel keeps its own duration): the same bits either way."""
    _, ctx, beads, L = pair(orc, gpu, N, M, 0.01, rc_frac, mode, GSF)
    rng = np.random.default_rng(9)
    for trial in range(3):
        seg = int(rng.integers(3, 9))
        start, p = int(rng.integers(M)), int(rng.integers(N))
        sp = np.array([[(start + k) % M, p] for k in range(seg)], dtype=np.int32)
        new = beads[sp[:, 0], sp[:, 1]] + rng.normal(scale=0.3, size=(seg, 3))
        new -= L * np.floor(new / L + 0.5)
        one = np.asarray(ctx.delta_action(sp, new))
        ctx.set_timing(True, kinds=["delta_action", "gradvnn"])
        two = np.asarray(ctx.delta_action(sp, new))
        ctx.set_timing(False)
        assert np.array_equal(one, two), (trial, one, two)
        assert ctx.kernel_time("gradvnn")[1] == trial + 1
