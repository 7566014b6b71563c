"""GPU parity of libpimc_gpu.so against the CPU oracle on identical seeded
worldlines (the oracle restates the reference's loops in FP64).

Tolerances (written per test):
  total potential action, V per slice, energy:  rel <= 1e-6 (FP32 pair terms),
    1e-9 (FP64 closed form), 1e-12 (FP64 table gather, same bins as the host);
  Delta V of a move: |dU - dU_ref| <= 1e-6 * (|A_old| + |A_new|) in FP32 mode;
  S(q), F(q,tau): rel <= 1e-5 of max(|value|, N*M*1e-3);
  cell lists, interaction lists, separation histograms: bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853
TOL = {0: 1e-6, 1: 1e-9, 2: 1e-12}


@pytest.fixture(scope="module")
def gpu():
    import pimc_amd
    return pimc_amd


def setup_pair(orc, gpu, N, M, tau, rc_frac=1.0, seed=SEED0, kind=0, mode=0):
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(seed, L, M, N, LAMBDA, tau, kind)
    ref = orc.State(L, M, N, tau, LAMBDA, rc_frac * L)
    ref.set_beads(beads)
    ctx = gpu.Context(L, M, N, tau, LAMBDA, rc_frac * L, v_mode=mode)
    ctx.upload_path(beads)
    return ref, ctx, beads, L


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_total_action_n16_m125_seeds(orc, gpu, mode):
    """Small walker N=16, M=125, tau=0.004; 100 seeded configurations."""
    worst = 0.0
    for k in range(100):
        ref, ctx, _, _ = setup_pair(orc, gpu, 16, 125, 0.004, seed=SEED0 + k, mode=mode)
        a, b = ctx.potentialAction(), ref.total_action()
        worst = max(worst, abs(a - b) / abs(b))
    print(f"N16 M125 mode {mode}: worst rel err {worst:.3e}")
    assert worst <= TOL[mode]


def test_launch_shapes_and_autotune(orc, gpu):
    """Every launch shape of the all-pairs kernel (tile size, workgroups per CU,
    LDS staging or cache reads) gives the oracle's action in the table mode, and
    the autotuner keeps one of them."""
    ref, ctx, _, _ = setup_pair(orc, gpu, 100, 24, 0.004, mode=2)
    b = ref.total_action()
    for shape in ((16, 2, 1), (16, 16, 0), (64, 4, 1), (1, 8, 0), (0, 0, -1)):
        ctx.set_launch_shape(*shape)
        a = ctx.potentialAction()
        assert abs(a - b) <= TOL[2] * abs(b), (shape, a, b)
    t = ctx.autotune(2)
    assert t["grid_per_cu"] in (2, 4, 8, 16) and t["stage_in_lds"] in (0, 1) and t["ms"] > 0
    assert abs(ctx.potentialAction() - b) <= TOL[2] * abs(b)
    with pytest.raises(gpu.PimcError):
        ctx.set_launch_shape(16, 100, 1)


def test_timing_kinds(orc, gpu):
    """pimc_set_timing_kinds brackets only the named kernels."""
    ref, ctx, beads, L = setup_pair(orc, gpu, 32, 8, 0.004)
    sp = np.array([[1, 3], [2, 3]], dtype=np.int32)
    new = beads[sp[:, 0], sp[:, 1]] + 0.1
    ctx.set_timing(True, kinds=["pair_action"])
    ctx.potentialAction()
    ctx.delta_action(sp, new)
    assert ctx.kernel_time("pair_action")[1] == 1 and ctx.kernel_time("delta_action")[1] == 0
    ctx.set_timing(True, kinds=["delta_action"])
    ctx.potentialAction()
    ctx.delta_action(sp, new)
    assert ctx.kernel_time("pair_action")[1] == 1 and ctx.kernel_time("delta_action")[1] == 1
    ctx.set_timing(False)


@pytest.mark.parametrize("mode", [0, 2])
def test_total_action_and_slices_n256_m167(orc, gpu, mode):
    """Single walker N=256, M=167, tau=0.004."""
    ref, ctx, _, _ = setup_pair(orc, gpu, 256, 167, 0.004, mode=mode)
    a, vs = ctx.potentialAction(with_slices=True)
    b = ref.total_action()
    assert abs(a - b) <= TOL[mode] * abs(b), (a, b)
    for m in (0, 1, 80, 166):
        v_ref = ref.V_slice(m)[1]
        assert abs(vs[m, 1] - v_ref) <= TOL[mode] * max(abs(v_ref), 1.0)
    # shift and even/odd factors of LocalAction::potentialAction()
    vf = (2.0 / 3.0, 4.0 / 3.0)
    ref.set_action(2, vf)
    a2 = ctx.potentialAction(shift=2, vfactor=vf)
    b2 = ref.total_action()
    assert abs(a2 - b2) <= TOL[mode] * abs(b2)


def test_total_action_bridges_and_odd_sizes(orc, gpu):
    for (N, M, kind) in ((1, 5, 0), (2, 4, 0), (3, 7, 1), (65, 9, 1), (130, 6, 0)):
        ref, ctx, _, _ = setup_pair(orc, gpu, N, M, 0.004, kind=kind)
        a, b = ctx.potentialAction(), ref.total_action()
        assert abs(a - b) <= 1e-6 * max(abs(b), 1e-12), (N, M, a, b)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_total_action_degenerate_pairs(orc, gpu, mode):
    """Pairs outside the table: two coincident beads (bin 0, V = 0) and a pair
    half a box apart on every axis (|r| = maxSep, bin tableLength, V = 0),
    include/potential.h:252-257."""
    N, M = 16, 4
    ref, ctx, beads, L = setup_pair(orc, gpu, N, M, 0.004, mode=mode)
    beads = beads.copy()
    beads[0, 1] = beads[0, 0]
    beads[1, 2] = [-0.25 * L] * 3
    beads[1, 3] = [0.25 * L] * 3
    ref.set_beads(beads)
    ctx.upload_path(beads)
    a, vs = ctx.potentialAction(with_slices=True)
    b = ref.total_action()
    assert abs(a - b) <= TOL[mode] * abs(b), (a, b)
    for m in range(M):
        v_ref = ref.V_slice(m)[1]
        assert abs(vs[m, 1] - v_ref) <= TOL[mode] * max(abs(v_ref), 1.0), (m, vs[m, 1], v_ref)


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("year", [1987, 1995])
def test_total_action_other_aziz_years(orc, gpu, mode, year):
    """The 1987 and 1995 parameter sets (beta != 0 in 1987: the repulsion's
    quadratic exponent term) through the all-pairs kernel."""
    N, M = 64, 12
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(SEED0, L, M, N, LAMBDA, 0.004, 1)
    ref = orc.State(L, M, N, 0.004, LAMBDA, L, year=year)
    ref.set_beads(beads)
    ctx = gpu.Context(L, M, N, 0.004, LAMBDA, L, year=year, v_mode=mode)
    ctx.upload_path(beads)
    a, b = ctx.potentialAction(), ref.total_action()
    assert abs(a - b) <= TOL[mode] * abs(b), (year, mode, a, b)


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("N", [300, 700])
def test_total_action_lds_strides(orc, gpu, mode, N):
    """Slices staged with the 512- and 1024-bead LDS plane strides."""
    ref, ctx, _, _ = setup_pair(orc, gpu, N, 3, 0.004, mode=mode)
    a, b = ctx.potentialAction(), ref.total_action()
    assert abs(a - b) <= TOL[mode] * abs(b), (N, a, b)


@pytest.mark.parametrize("N", [512, 700, 1024, 1537, 2048])
def test_total_action_spatial_order(orc, gpu, N):
    """Slices staged in spatial order with the far-pair path (mixed mode): the
    oracle's action and per-slice V at 1e-6, and the bead-order staging's at
    1e-12 (the same bins; only the far pairs' 1/k^2, the dropped repulsion below
    2^-53 of V and the summation order differ).  Index-order staging is forced
    with stage_in_lds = 1, spatial order with 2; the automatic shape takes
    spatial order above 1024 beads."""
    M = 4
    ref, ctx, _, _ = setup_pair(orc, gpu, N, M, 0.004, rc_frac=0.5, kind=1)
    b = ref.total_action()
    vs_ref = np.array([ref.V_slice(m)[1] for m in range(M)])
    got = {}
    for lds in (1, 2, -1):
        ctx.set_launch_shape(0, 0, lds)
        a, vs = ctx.potentialAction(with_slices=True)
        vs = vs[:, 1].copy()
        assert abs(a - b) <= TOL[0] * abs(b), (lds, a, b)
        assert np.allclose(vs, vs_ref, rtol=TOL[0], atol=0), lds
        got[lds] = (a, vs)
    a1, a2 = got[1][0], got[2][0]
    print(f"N{N}: spatial - index order {abs(a2 - a1) / abs(a1):.2e}, oracle {abs(a2 - b) / abs(b):.2e}")
    assert abs(a2 - a1) <= 1e-12 * abs(a1)
    assert np.allclose(got[2][1], got[1][1], rtol=1e-12, atol=0)
    assert got[-1][0] == (a2 if N > 1024 else a1)
    # the same slices twice: the spatial order depends only on the positions
    ctx.set_launch_shape(0, 0, 2)
    assert ctx.potentialAction() == a2


def test_total_action_spatial_uneven_slices(orc, gpu):
    """The window form over slices of uneven bead counts (numBeadsAtSlice, left
    packed): windows on slices of 256 beads and more, whole-slice staging below
    (255, 100), no pairs at one bead; the oracle at 1e-6, bead-order staging at
    1e-12."""
    N, M = 700, 6
    ref, ctx, beads, L = setup_pair(orc, gpu, N, M, 0.004, rc_frac=0.5, kind=1)
    nb = np.array([700, 255, 100, 1, 256, 517], dtype=np.int32)
    on = np.zeros((M, N), dtype=np.uint8)
    for m in range(M):
        on[m, :nb[m]] = 1
    nxt = np.zeros((M, N, 2), dtype=np.int32)
    prv = np.zeros((M, N, 2), dtype=np.int32)
    for m in range(M):
        for i in range(N):
            nxt[m, i] = ((m + 1) % M, i)
            prv[m, i] = ((m - 1) % M, i)
    ref.set_topology(on, nb, nxt, prv)
    ctx.upload_path(beads, bead_on=on, num_beads=nb)
    b = ref.total_action()
    vs_ref = np.array([ref.V_slice(m)[1] for m in range(M)])
    got = []
    for lds in (1, 2):
        ctx.set_launch_shape(0, 0, lds)
        a, vs = ctx.potentialAction(with_slices=True)
        assert abs(a - b) <= TOL[0] * abs(b), (lds, a, b)
        for m in range(M):
            assert abs(vs[m, 1] - vs_ref[m]) <= TOL[0] * max(abs(vs_ref[m]), 1.0), (lds, m, vs[m, 1], vs_ref[m])
        got.append(vs[:, 1].copy())
    assert np.allclose(got[1], got[0], rtol=1e-12, atol=1e-12)


def test_total_action_spatial_order_worm(orc, gpu):
    """Spatial staging with a worm: slices the worm touches are staged in bead
    order for the pair factors (src/action.cpp:925-940), the others in spatial
    order; the oracle's action at 1e-6 and the bead-order staging's at 1e-12."""
    N, M = 600, 6
    ref, ctx, beads, L = setup_pair(orc, gpu, N, M, 0.004, rc_frac=0.5, kind=1)
    on = np.ones((M, N), dtype=np.uint8)
    nb = np.full(M, N, dtype=np.int32)
    nxt = np.zeros((M, N, 2), dtype=np.int32)
    prv = np.zeros((M, N, 2), dtype=np.int32)
    for m in range(M):
        for i in range(N):
            nxt[m, i] = ((m + 1) % M, i)
            prv[m, i] = ((m - 1) % M, i)
    head, tail, s1 = (1, 20), (3, 500), (4, 7)
    ref.set_topology(on, nb, nxt, prv, head=head, tail=tail, special1=s1)
    ctx.set_worm(head=head, tail=tail, special1=s1)
    b = ref.total_action()
    ctx.set_launch_shape(0, 0, 1)
    a1 = ctx.potentialAction()
    ctx.set_launch_shape(0, 0, 2)
    a2 = ctx.potentialAction()
    assert abs(a2 - b) <= TOL[0] * abs(b), (a2, b)
    assert abs(a2 - a1) <= 1e-12 * abs(a1), (a1, a2)


def _bisection_case(orc, gpu, N, M, seg, seed, mode=0, rc_frac=1.0):
    ref, ctx, beads, L = setup_pair(orc, gpu, N, M, 0.004, rc_frac=rc_frac, seed=seed, mode=mode)
    rng = np.random.default_rng(seed)
    start, p = int(rng.integers(M)), int(rng.integers(N))
    sp = np.array([[(start + k) % M, p] for k in range(1, seg)], dtype=np.int32)
    old = beads[sp[:, 0], sp[:, 1]]
    new = old + rng.normal(scale=np.sqrt(LAMBDA * 0.004) * 2, size=old.shape)
    new -= L * np.floor(new / L + 0.5)
    return ref, ctx, sp, old, new, L


@pytest.mark.parametrize("seg", [8, 16, 32])
@pytest.mark.parametrize("mode", [0, 2])
def test_delta_v_bisection_n256_m167(orc, gpu, seg, mode):
    """Bisection segments of 8/16/32 slices (7/15/31 moved beads) on N=256, M=167."""
    for seed in range(5):
        ref, ctx, sp, old, new, L = _bisection_case(orc, gpu, 256, 167, seg, SEED0 + seed, mode)
        a_old = ref.action_list(sp)
        for (m, i), q in zip(sp, new):
            ref.update_bead(int(m), int(i), q)
        a_new = ref.action_list(sp)
        out = ctx.delta_action(sp, new)
        scale = abs(a_old) + abs(a_new)
        tol = 1e-6 if mode == 0 else 1e-12
        assert abs(out[0] - a_old) <= tol * scale
        assert abs(out[1] - a_new) <= tol * scale
        assert abs(out[2] - (a_new - a_old)) <= tol * scale, (out, a_old, a_new)
        # explicit old positions give the same result
        out2 = ctx.delta_action(sp, new, old_pos=old)
        assert np.array_equal(out, out2)


def test_delta_v_commit_and_cells(orc, gpu):
    ref, ctx, sp, old, new, L = _bisection_case(orc, gpu, 64, 20, 16, SEED0 + 9, rc_frac=0.22)
    out = ctx.delta_action(sp, new, commit=True)
    for (m, i), q in zip(sp, new):
        ref.update_bead(int(m), int(i), q)
    got = ctx.download_path()
    assert np.array_equal(got, ref.beads())
    _compare_cells(ctx.cell_lists(), ref.grid_dump())
    # action of the committed configuration
    b = ref.total_action()
    assert abs(ctx.potentialAction() - b) <= 1e-6 * abs(b)
    assert out[1] == pytest.approx(ref.action_list(sp), rel=1e-6)


def _compare_cells(g, r):
    assert np.array_equal(g["dims"], r["dims"])
    assert np.array_equal(g["grid"], r["grid"])
    assert np.array_equal(g["bead_label"], r["bead_label"])
    assert np.array_equal(g["num_labels"], r["num_labels"])
    cells, M = r["num_labels"].shape
    for c in range(cells):
        for m in range(M):
            n = r["num_labels"][c, m]
            assert np.array_equal(g["hash"][c, m, :n], r["hash"][c, m, :n])


def test_empty_and_invalid_requests(orc, gpu):
    """Empty bead and q lists give zeros and empty results; out-of-range
    locators and two moved beads on one slice raise before any launch; the
    context answers as before afterwards."""
    ref, ctx, beads, L = setup_pair(orc, gpu, 32, 20, 0.004)
    a0 = ctx.potentialAction()
    empty_sp = np.zeros((0, 2), dtype=np.int32)
    assert np.array_equal(ctx.delta_action(empty_sp, np.zeros((0, 3))), np.zeros(3))
    assert ctx.action_list(empty_sp) == 0.0
    assert ctx.ssf(np.zeros((0, 3))).shape == (0,)
    assert ctx.isf(np.zeros((0, 3))).shape == (0, 20)
    with pytest.raises(gpu.PimcError):
        ctx.delta_action([[20, 0]], [[0.0, 0.0, 0.0]])              # slice out of range
    with pytest.raises(gpu.PimcError):
        ctx.delta_action([[0, 32]], [[0.0, 0.0, 0.0]])              # particle out of range
    with pytest.raises(gpu.PimcError):
        ctx.delta_action([[3, 1], [3, 2]], beads[[3, 3], [1, 2]])   # two moved beads on one slice
    with pytest.raises(gpu.PimcError):
        ctx.action_list([[-1, 0]])
    assert ctx.potentialAction() == a0
    sp = np.array([[5, 7], [6, 7]], dtype=np.int32)
    new = beads[sp[:, 0], sp[:, 1]] + 0.1
    new -= L * np.floor(new / L + 0.5)
    out = ctx.delta_action(sp, new)
    old_ref = ref.action_list(sp)
    for (m, i), q in zip(sp, new):
        ref.update_bead(int(m), int(i), q)
    new_ref = ref.action_list(sp)
    scale = abs(old_ref) + abs(new_ref)
    assert abs(out[0] - old_ref) <= 1e-6 * scale
    assert abs(out[2] - (new_ref - old_ref)) <= 1e-6 * scale


@pytest.mark.parametrize("frac", [1.0, 0.5, 0.3, 0.22])
def test_cell_lists_bit_exact(orc, gpu, frac):
    ref, ctx, beads, L = setup_pair(orc, gpu, 96, 6, 0.004, rc_frac=frac)
    _compare_cells(ctx.cell_lists(), ref.grid_dump())
    rng = np.random.default_rng(4)
    sp, pos = [], []
    for _ in range(200):
        m, i = int(rng.integers(6)), int(rng.integers(96))
        q = beads[m, i] + rng.normal(scale=3.0, size=3)
        q -= L * np.floor(q / L + 0.5)
        sp.append((m, i))
        pos.append(q)
        ref.update_bead(m, i, q)
    ctx.update_beads(np.array(sp), np.array(pos))
    _compare_cells(ctx.cell_lists(), ref.grid_dump())
    for (m, i) in ((0, 0), (3, 50), (5, 95)):
        lg, sg = ctx.interaction_list(m, i)
        lr, sr = ref.interaction_list(m, i)
        assert np.array_equal(lg, lr) and np.array_equal(sg, sr)


@pytest.mark.parametrize("N,frac", [(300, 0.3), (700, 0.5)])
def test_cell_lists_many_beads(orc, gpu, N, frac):
    """Labels over several 64-bead groups per slice, in the serial insertion order."""
    ref, ctx, _, _ = setup_pair(orc, gpu, N, 3, 0.004, rc_frac=frac)
    _compare_cells(ctx.cell_lists(), ref.grid_dump())


@pytest.mark.parametrize("frac", [1.0, 0.5, 0.22])
def test_action_list_vnn_cutoff(orc, gpu, frac):
    ref, ctx, beads, L = setup_pair(orc, gpu, 128, 8, 0.004, rc_frac=frac, mode=2)
    sp = np.array([[m, (7 * m + 3) % 128] for m in range(8)], dtype=np.int32)
    a, pb = ctx.action_list(sp, per_bead=True)
    for k, (m, i) in enumerate(sp):
        assert pb[k] == pytest.approx(ref.Vnn(int(m), int(i)), rel=1e-12, abs=1e-12)
    assert a == pytest.approx(ref.action_list(sp), rel=1e-12)
    r = ctx.action_range((2, 5), (6, 5))
    assert r == pytest.approx(ref.action_range((2, 5), (6, 5)), rel=1e-12)


def test_worm_factors(orc, gpu):
    N, M = 12, 6
    ref, ctx, beads, L = setup_pair(orc, gpu, N, M, 0.004, mode=2)
    on = np.ones((M, N), dtype=np.uint8)
    nb = np.full(M, N, dtype=np.int32)
    nxt = np.zeros((M, N, 2), dtype=np.int32)
    prv = np.zeros((M, N, 2), dtype=np.int32)
    for m in range(M):
        for i in range(N):
            nxt[m, i] = ((m + 1) % M, i)
            prv[m, i] = ((m - 1) % M, i)
    head, tail, s1 = (1, 2), (1, 5), (3, 0)
    ref.set_topology(on, nb, nxt, prv, head=head, tail=tail, special1=s1)
    ctx.set_worm(head=head, tail=tail, special1=s1)
    assert ctx.potentialAction() == pytest.approx(ref.total_action(), rel=1e-12)
    sp = np.array([[1, 2], [3, 0], [4, 4]], dtype=np.int32)
    assert ctx.action_list(sp) == pytest.approx(ref.action_list(sp), rel=1e-12)
    # Delta V of the three beads (distinct slices, the short-request kernel) with
    # the worm's pair factors, then committed
    new = beads[sp[:, 0], sp[:, 1]] + np.random.default_rng(4).normal(scale=0.2, size=(3, 3))
    new -= L * np.floor(new / L + 0.5)
    a_old = ref.action_list(sp)
    for (m, i), q in zip(sp, new):
        ref.update_bead(int(m), int(i), q)
    a_new = ref.action_list(sp)
    out = ctx.delta_action(sp, new, commit=True)
    scale = abs(a_old) + abs(a_new)
    assert abs(out[0] - a_old) <= 1e-12 * scale and abs(out[1] - a_new) <= 1e-12 * scale
    assert ctx.potentialAction() == pytest.approx(ref.total_action(), rel=1e-12)


def test_energy_and_sep_hist(orc, gpu):
    ref, ctx, beads, L = setup_pair(orc, gpu, 64, 40, 0.004, kind=1)
    e_g, e_r = ctx.energy(), ref.energy()
    for k in ("K", "V", "V_int", "E", "V/N"):
        assert e_g[k] == pytest.approx(e_r[k], rel=1e-6), k
    vs, h = ctx.potential_slices(sep_hist=True)
    _, h_ref = ref.V_slice(39, hist=True)
    assert np.array_equal(h, h_ref)


def test_ssf_n16_m125(orc, gpu):
    """S(q) on N=16, M=125 for the 729-vector cube |n_i| <= 4."""
    for seed in (SEED0, SEED0 + 1):
        ref, ctx, beads, L = setup_pair(orc, gpu, 16, 125, 0.004, seed=seed)
        q = orc.qvectors(L, "max_int", "4 4 4")
        sg, sr = ctx.ssf(q), ref.ssf(q)
        assert np.all(np.abs(sg - sr) <= 1e-5 * np.maximum(np.abs(sr), 16 * 125 * 1e-3))


def test_ssf_general_q(orc, gpu):
    ref, ctx, beads, L = setup_pair(orc, gpu, 24, 10, 0.004)
    q = np.array([[0.3, 0.1, -0.2], [1.0, 2.0, 0.5], [2 * np.pi / L, 0, 0]])
    sg, sr = ctx.ssf(q), ref.ssf(q)
    assert np.all(np.abs(sg - sr) <= 1e-5 * np.maximum(np.abs(sr), 24 * 10 * 1e-3))


def test_isf_small(orc, gpu):
    ref, ctx, beads, L = setup_pair(orc, gpu, 16, 20, 0.004)
    q = np.concatenate([orc.qvectors(L, "max_int", "1 1 1")[:10], [[0.37, -0.2, 0.9]]])
    fg, fr = ctx.isf(q), ref.isf(q)
    assert np.all(np.abs(fg - fr) <= 1e-5 * np.maximum(np.abs(fr), 16 * 20 * 1e-3))
    assert np.allclose(fg[:, 1:], fg[:, :0:-1], rtol=1e-12, atol=1e-9)
    # into a page-locked result array: the same bits; a wrong shape is refused
    out = gpu.pinned_empty((q.shape[0], 20))
    assert ctx.isf(q, out=out) is out and np.array_equal(out, fg)
    with pytest.raises(gpu.PimcError):
        ctx.isf(q, out=gpu.pinned_empty((q.shape[0], 19)))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_aziz_batch(orc, gpu, mode):
    az_r = orc.Aziz(1979, max_sep=12.0, rc=10.0, derivs=True)
    az_g = gpu.Aziz(1979, 12.0, 10.0, v_mode=mode)
    rng = np.random.default_rng(5)
    sep = rng.normal(size=(4000, 3)) * rng.uniform(0.0, 5.0, size=(4000, 1))
    sep[0] = [1e-8, 0, 0]
    sep[1] = [20.0, 0, 0]
    # the hard core x = r/rm < 0.01 (r < 0.0297 A), down to the EPS bin, and the
    # bins either side of x = 0.01 and x = D
    rm, D = 2.9673, 1.241314
    radii = np.concatenate([np.geomspace(3e-6, 0.0297, 30), 0.01 * rm * (1 + np.array([-1e-6, 1e-6])),
                            D * rm * (1 + np.array([-1e-6, 1e-6]))])
    sep[2:2 + radii.size] = 0.0
    sep[2:2 + radii.size, 0] = radii
    v = az_g.V(sep)
    vr = np.array([az_r.V(s) for s in sep])
    # per-pair scale: |V| plus the well depth, so the zero crossing keeps a finite scale
    scale = np.abs(vr) + 10.8
    err = np.abs(v - vr) / scale
    print(f"aziz batch mode {mode}: max err/(|V|+eps) {err.max():.3e}")
    assert err.max() <= {0: 5e-7, 1: 1e-9, 2: 0.0}[mode]
    g = az_g.gradV(sep[2:])
    gr = np.array([az_r.gradV(s) for s in sep[2:]])
    assert np.array_equal(g, gr)
    assert np.array_equal(az_g.grad2V(sep[2:]), np.array([az_r.grad2V(s) for s in sep[2:]]))
    assert az_g.tailV == az_r.tailV


def test_total_action_n2048_m250(orc, gpu):
    """Large system N=2048, M=250, rc = L/2, Gaussian-bridge worldlines (full size)."""
    ref, ctx, beads, L = setup_pair(orc, gpu, 2048, 250, 0.004, rc_frac=0.5, kind=1)
    b = ref.total_action()
    for mode in (0, 2):
        ctx.set_v_mode(mode)
        a = ctx.potentialAction()
        print(f"N2048 M250 mode {mode}: rel err {abs(a - b) / abs(b):.3e}")
        assert abs(a - b) <= TOL[mode] * abs(b)


def test_isf_n256_m167_properties(orc, gpu):
    """F(q,tau) on the N=256, M=167 q-set (5377 lattice q, |q| <= 3 A^-1, all 167 lags):
    F(q,0) = S(q), F(q,tau) = F(q,M-tau), and the oracle on a few q."""
    from pimc_amd.workloads import lattice_q
    ref, ctx, beads, L = setup_pair(orc, gpu, 256, 167, 0.004)
    q = lattice_q(L, qmag_max=3.0)
    assert q.shape[0] == 5377
    f = ctx.isf(q)
    s = ctx.ssf(q)
    assert np.allclose(f[:, 0], s, rtol=1e-12, atol=1e-9)
    assert np.allclose(f[:, 1:], f[:, :0:-1], rtol=1e-10, atol=1e-7)
    i0 = int(np.where(np.all(q == 0, axis=1))[0][0])
    assert s[i0] == pytest.approx(256 * 167, rel=1e-12)
    pick = [0, 1, 2688, 5376]
    sr = ref.ssf(q[pick])
    assert np.all(np.abs(s[pick] - sr) <= 1e-5 * np.maximum(np.abs(sr), 256 * 167 * 1e-3))


def test_delta_action_prepared(orc, gpu):
    """A request converted once (prepare_delta) gives delta_action's result bit
    for bit, also when committing."""
    ref, ctx, beads, L = setup_pair(orc, gpu, 64, 24, 0.004)
    rng = np.random.default_rng(31)
    sp = np.array([[m, 5] for m in range(3, 11)], dtype=np.int32)
    new = beads[sp[:, 0], sp[:, 1]] + rng.normal(scale=0.2, size=(len(sp), 3))
    a = ctx.delta_action(sp, new)
    prep = ctx.prepare_delta(sp, new)
    b = ctx.delta_action_prepared(prep)[:3].copy()
    assert np.array_equal(a, b)
    c = ctx.delta_action_prepared(prep, commit=True)[:3].copy()
    assert np.array_equal(c, a)
    assert np.array_equal(ctx.download_path()[sp[:, 0], sp[:, 1]], new)
