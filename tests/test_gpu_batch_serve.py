"""GPU tests of the batched and serving forms.

Batched: per-walker potentialAction() of uneven walkers against the oracle
(rel <= 1e-6, mixed mode), a walker's result bit-identical alone and in the mix,
hipGraph replays bit-identical to eager launches with eager work in between.
Serving: Delta V against the oracle, results bit-identical alone and in the mix,
commit / Metropolis semantics, the HBM-budget wait queue, a short synthetic load."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853


def walker(orc, N, M, T, seed, kind=1):
    L = orc.box_side(N, DENSITY)
    tau = 1.0 / (M * T)
    b = orc.gen_worldlines(seed, L, M, N, LAMBDA, tau, kind).astype(np.float32).astype(np.float64)
    b -= L * np.floor(b / L + 0.5)
    return dict(side=L, M=M, N=N, tau=tau, lam=LAMBDA, rc=L), b


def batch_set(orc, n, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        N = int(rng.choice([16, 32, 64, 128]))
        M = int(rng.choice([8, 16, 24]))
        T = float(rng.uniform(1.0, 3.0))
        out.append(walker(orc, N, M, T, SEED0 + k))
    return out


def oracle_action(orc, w, b):
    st = orc.State(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"])
    st.set_beads(b)
    return st.total_action()


def test_batch_against_oracle_and_mix(orc):
    import pimc_amd
    ws = batch_set(orc, 12)
    bt = pimc_amd.Batch([w for w, _ in ws])
    for k, (_, b) in enumerate(ws):
        bt.upload(k, b)
    a = bt.potentialAction()
    for k, (w, b) in enumerate(ws):
        ref = oracle_action(orc, w, b)
        assert abs(a[k] - ref) <= 1e-6 * abs(ref), (k, a[k], ref)
    # the work list: every slice's n(n-1)/2 pairs once, at least one tile per slice
    inf = bt.info()
    assert inf["total_pairs"] == sum(w["M"] * w["N"] * (w["N"] - 1) // 2 for w, _ in ws)
    assert inf["n_items"] >= sum(w["M"] for w, _ in ws)
    assert 1 <= inf["n_groups"] <= inf["n_items"] and inf["last_ms"] >= 0.0
    # alone vs in the mix, same tiles: bit-identical
    for k in (0, 5, 11):
        solo = pimc_amd.Batch([ws[k][0]])
        solo.upload(0, ws[k][1])
        assert solo.potentialAction()[0] == a[k]


def test_batch_graph_replays(orc):
    import pimc_amd
    ws = batch_set(orc, 8, seed=3)
    bt = pimc_amd.Batch([w for w, _ in ws])
    for k, (_, b) in enumerate(ws):
        bt.upload(k, b)
    eager = bt.potentialAction()
    bt.use_graph(True)
    ctx = pimc_amd.Context(ws[0][0]["side"], ws[0][0]["M"], ws[0][0]["N"], ws[0][0]["tau"], LAMBDA, ws[0][0]["rc"])
    ctx.upload_path(ws[0][1])
    for rep in range(4):
        g = bt.potentialAction()
        assert np.array_equal(g, eager), rep
        ctx.potentialAction()          # eager work between replays
        pimc_amd.device_synchronize()
    bt.set_tiles(16, 64)
    g2 = bt.potentialAction()
    for k, (w, b) in enumerate(ws):
        assert abs(g2[k] - eager[k]) <= 1e-12 * abs(eager[k])


def test_batch_autotune_small(orc):
    import pimc_amd
    ws = batch_set(orc, 4, seed=5)
    bt = pimc_amd.Batch([w for w, _ in ws])
    for k, (_, b) in enumerate(ws):
        bt.upload(k, b)
    ref = bt.potentialAction()
    best = bt.autotune(reps=1)
    assert best["candidates"] == 24 and best["ms"] > 0
    after = bt.potentialAction()
    assert np.allclose(after, ref, rtol=1e-12)


def _segment(rng, M, N, L, beads, seg):
    start, p = int(rng.integers(M)), int(rng.integers(N))
    sp = np.array([[(start + k) % M, p] for k in range(1, seg)], dtype=np.int32)
    new = beads[sp[:, 0], sp[:, 1]] + rng.normal(scale=0.3, size=(seg - 1, 3))
    new -= L * np.floor(new / L + 0.5)
    return sp, new


def test_server_delta_v_and_mix(orc):
    import pimc_amd
    chains = [walker(orc, N, M, 1.5, SEED0 + k, kind=0) for k, (N, M) in enumerate([(64, 32), (128, 16), (32, 40)])]
    srv = pimc_amd.Server(1 << 28, streams=2, max_side=40.0)
    ids = [srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b) for w, b in chains]
    rng = np.random.default_rng(1)
    w0, b0 = chains[0]
    sp, new = _segment(rng, w0["M"], w0["N"], w0["side"], b0, 12)
    alone = srv.wait(srv.submit(ids[0], sp, new))
    st = orc.State(w0["side"], w0["M"], w0["N"], w0["tau"], LAMBDA, w0["rc"])
    st.set_beads(b0)
    old = st.action_list(sp)
    for (m, i), q in zip(sp, new):
        st.update_bead(int(m), int(i), q)
    newa = st.action_list(sp)
    scale = abs(old) + abs(newa)
    assert abs(alone["A_old"] - old) <= 1e-6 * scale
    assert abs(alone["dU"] - (newa - old)) <= 1e-6 * scale
    # the same request in a mix of many others
    others = []
    pairs = 2 * (2 * len(sp) * (w0["N"] - 1))   # Vnn at the old and new position, alone and mixed
    for rep in range(30):
        k = 1 + rep % 2
        w, b = chains[k]
        s2, n2 = _segment(rng, w["M"], w["N"], w["side"], b, int(rng.integers(4, 12)))
        others.append(srv.submit(ids[k], s2, n2))
        pairs += 2 * len(s2) * (w["N"] - 1)
    mixed = srv.submit(ids[0], sp, new)
    res = [srv.wait(r) for r in others]
    m = srv.wait(mixed)
    assert np.array_equal(m["raw"][:4], alone["raw"][:4])
    assert all(np.isfinite(r["dU"]) for r in res)
    # the server's counters: every request launched once, its pairs counted once
    st = srv.stats()
    assert st["requests"] == 32 and 1 <= st["batches"] <= 32
    assert st["pair_evals"] == pairs
    assert 0 < st["free_bytes"] < (1 << 28)


def test_server_commit_metropolis_and_budget(orc):
    import pimc_amd
    w, b = walker(orc, 64, 16, 1.5, SEED0, kind=0)
    need = 16 * 64 * (24 + 4 + 1) + 3 * 256
    srv = pimc_amd.Server(int(need * 1.5), streams=2, max_side=40.0)
    c0 = srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    c1 = srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    assert srv.chain_state(c0) == 1 and srv.chain_state(c1) == 0
    rng = np.random.default_rng(2)
    sp, new = _segment(rng, 16, 64, w["side"], b, 6)
    r1 = srv.submit(c1, sp, new, commit_mode=2)          # queued until c1 is admitted
    assert not srv.poll(r1)
    r0 = srv.wait(srv.submit(c0, sp, new, commit_mode=1, u=0.0))
    with np.errstate(over="ignore"):
        assert r0["accepted"] == (0.0 < np.exp(-r0["dU"]))
    srv.remove_chain(c0)
    out = srv.wait(r1)
    assert out["accepted"] and srv.chain_state(c1) == 1
    got = srv.download(c1)
    exp = b.copy()
    exp[sp[:, 0], sp[:, 1]] = new
    assert np.array_equal(got, exp)
    # u = 1 rejects unless dU < 0
    r = srv.wait(srv.submit(c1, sp, b[sp[:, 0], sp[:, 1]], commit_mode=1, u=1.0))
    assert r["accepted"] == (1.0 < np.exp(-r["dU"]))


def test_server_synthetic_load(orc):
    import pimc_amd
    srv = pimc_amd.Server(1 << 30, streams=4, max_side=40.0)
    rng = np.random.default_rng(3)
    for k in range(32):
        N = int(rng.integers(32, 129))
        M = int(rng.integers(64, 129))
        w, b = walker(orc, N, M, float(rng.uniform(1, 3)), SEED0 + k, kind=1)
        srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    st = srv.run_synthetic(32, rate=20000.0, duration_s=0.5)
    assert st["completed"] > 100 and st["p99_us"] >= st["p50_us"] > 0
    st2 = srv.run_synthetic(32, rate=0.0, duration_s=0.3)
    assert st2["completed"] > 100 and st2["req_per_batch"] >= 1.0
    # several client threads (chains split among them), open and closed loop
    st3 = srv.run_synthetic(32, rate=20000.0, duration_s=0.3, threads=4)
    assert st3["completed"] > 100 and st3["p99_us"] >= st3["p50_us"] > 0
    st4 = srv.run_synthetic(32, rate=0.0, duration_s=0.3, threads=4)
    assert st4["completed"] > 100 and 0.0 < st4["accept_frac"] < 1.0
    # with 5 % S(q) requests (257 q each) mixed in
    st5 = srv.run_synthetic(32, rate=20000.0, duration_s=0.5, threads=4, p_ssf=0.05, ssf_nmax=4)
    assert st5["completed"] > 100 and st5["ssf_completed"] > 5
    assert st5["ssf_p99_us"] >= st5["ssf_p50_us"] > 0 and st5["ssf_phasors_per_s"] > 0
    # the chains come back whole after concurrent clients
    for c in (0, 7, 31):
        b = srv.download(c)
        assert np.all(np.isfinite(b))


def test_batch_tiny_walkers(orc):
    """Walkers of 1, 2 and 3 particles (no pair, one pair, odd n) mixed with a
    larger one, and a lone one-particle walker: the batch matches the oracle
    (0 where there is no pair), and so does the single-walker context."""
    import pimc_amd
    ws = [walker(orc, N, M, 1.5, SEED0 + k) for k, (N, M) in enumerate([(1, 8), (2, 12), (3, 10), (16, 8)])]
    bt = pimc_amd.Batch([w for w, _ in ws])
    for k, (_, b) in enumerate(ws):
        bt.upload(k, b)
    a = bt.potentialAction()
    assert a[0] == 0.0
    for k, (w, b) in enumerate(ws[1:], start=1):
        ref = oracle_action(orc, w, b)
        assert abs(a[k] - ref) <= 1e-6 * abs(ref), (k, a[k], ref)
    assert bt.info()["total_pairs"] == 12 * 1 + 10 * 3 + 8 * 120
    lone = pimc_amd.Batch([ws[0][0]])
    lone.upload(0, ws[0][1])
    assert lone.potentialAction()[0] == 0.0
    w, b = ws[0]
    ctx = pimc_amd.Context(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"])
    ctx.upload_path(b)
    assert ctx.potentialAction() == 0.0
    w, b = ws[1]
    ctx = pimc_amd.Context(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"])
    ctx.upload_path(b)
    ref = oracle_action(orc, w, b)
    assert abs(ctx.potentialAction() - ref) <= 1e-6 * abs(ref)


def test_server_remove_while_adding(orc):
    """remove_chain waits (lock released) until its chain's queued request has run;
    chains added meanwhile by another thread grow the server's chain container,
    and the waiting call still releases the right chain's memory and admits the
    next waiting chain."""
    import threading
    import pimc_amd
    w, b = walker(orc, 64, 16, 1.5, SEED0, kind=0)
    need = 16 * 64 * (24 + 4 + 1) + 3 * 256
    srv = pimc_amd.Server(int(need * 1.5), streams=2, max_side=40.0)
    args = (w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    c0 = srv.add_chain(*args)
    c1 = srv.add_chain(*args)
    assert srv.chain_state(c0) == 1 and srv.chain_state(c1) == 0
    rng = np.random.default_rng(9)
    sp, new = _segment(rng, 16, 64, w["side"], b, 6)
    r1 = srv.submit(c1, sp, new, commit_mode=2)     # runs once c1 is admitted
    result = []

    def remove():
        try:
            srv.remove_chain(c1)
            result.append("ok")
        except Exception as ex:       # noqa: BLE001 -- reported by the main thread
            result.append(repr(ex))

    t = threading.Thread(target=remove)
    t.start()
    import time
    time.sleep(0.05)
    assert t.is_alive() and not result
    extra = [srv.add_chain(*args) for _ in range(300)]
    assert all(srv.chain_state(e) == 0 for e in extra)
    srv.remove_chain(c0)                            # frees room: c1 in, its request runs, c1 out
    t.join(timeout=30)
    assert not t.is_alive() and result == ["ok"]
    assert srv.wait(r1)["accepted"]
    assert srv.chain_state(c1) == 2 and srv.chain_state(c0) == 2
    assert srv.chain_state(extra[0]) == 1
    assert all(srv.chain_state(e) == 0 for e in extra[1:])
    got = srv.download(extra[0])
    assert np.array_equal(got, b)
    srv.close()


def test_action_range_inactive_and_bad_end(orc):
    """pimc_action_range over next links: a walk through inactive beads that never
    reaches end's successor, and an end outside the path, are refused with an error
    (no endless walk on the host)."""
    import pimc_amd
    N, M = 8, 6
    L = orc.box_side(N, DENSITY)
    b = orc.gen_worldlines(SEED0, L, M, N, LAMBDA, 0.004, 0)
    on = np.ones((M, N), dtype=np.uint8)
    on[:, N - 1] = 0                                 # the last column inactive (left-packed slices)
    nb = np.full(M, N - 1, dtype=np.int32)
    ctx = pimc_amd.Context(L, M, N, 0.004, LAMBDA, L)
    ctx.upload_path(b, on, nb)
    nxt = np.zeros((M, N, 2), dtype=np.int32)
    for m in range(M):
        for i in range(N):
            nxt[m, i] = ((m + 1) % M, i)
    # start on the inactive worldline, end on another one: the walk cycles forever
    with pytest.raises(pimc_amd.PimcError, match="does not reach"):
        ctx.action_range((0, N - 1), (3, 2), next_links=nxt)
    with pytest.raises(pimc_amd.PimcError, match="out of range"):
        ctx.action_range((0, 1), (M, 1), next_links=nxt)
    with pytest.raises(pimc_amd.PimcError, match="out of range"):
        ctx.action_range((0, 1), (2, -1), next_links=nxt)
    # a proper range still works
    prv = np.array([[[(m - 1) % M, i] for i in range(N)] for m in range(M)], dtype=np.int32)
    st = orc.State(L, M, N, 0.004, LAMBDA, L)
    st.set_beads(b)
    st.set_topology(on, nb, nxt, prv)
    assert ctx.action_range((1, 2), (4, 2), next_links=nxt) == pytest.approx(st.action_range((1, 2), (4, 2)),
                                                                             rel=1e-6)


def test_server_ssf_requests(orc):
    """S(q) requests on resident chains (src/estimator.cpp:3705-3737): equal to
    pimc_ssf on a context holding the same worldlines (bit for bit) and to the
    oracle (1e-5 of max(S, M)); the same alone and inside a mix of Delta V and
    S(q) requests on other chains; queued behind the chain's earlier moves, so a
    committed move is seen; non-lattice q refused."""
    import pimc_amd
    chains = [walker(orc, N, M, 1.5, SEED0 + 40 + k, kind=0) for k, (N, M) in enumerate([(64, 32), (128, 16), (32, 40)])]
    srv = pimc_amd.Server(1 << 28, streams=3, max_side=40.0)
    ids = [srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b) for w, b in chains]
    w0, b0 = chains[0]
    L0 = w0["side"]
    q = orc.qvectors(L0, "max_int", "4 4 4")
    q = q[np.einsum("qd,qd->q", q, q) <= (4 * 2 * np.pi / L0) ** 2 * (1 + 1e-12)]
    assert q.shape[0] == 257
    ctx = pimc_amd.Context(L0, w0["M"], w0["N"], w0["tau"], LAMBDA, w0["rc"])
    ctx.upload_path(b0)
    ref_ctx = ctx.ssf(q)
    alone, lat = srv.wait_ssf(srv.submit_ssf(ids[0], q))
    assert lat > 0
    assert np.array_equal(alone, ref_ctx)
    st = orc.State(L0, w0["M"], w0["N"], w0["tau"], LAMBDA, w0["rc"])
    st.set_beads(b0)
    pick = [0, 1, 128, 200, 256]
    ref = st.ssf(q[pick])
    assert np.all(np.abs(alone[pick] - ref) <= 1e-5 * np.maximum(ref, w0["M"]))
    # in a mix of Delta V and S(q) requests on the other chains
    rng = np.random.default_rng(12)
    others, others_ssf = [], []
    for rep in range(40):
        k = 1 + rep % 2
        w, b = chains[k]
        s2, n2 = _segment(rng, w["M"], w["N"], w["side"], b, int(rng.integers(4, 12)))
        others.append(srv.submit(ids[k], s2, n2))
        if rep % 5 == 0:
            qk = orc.qvectors(w["side"], "max_int", "3 3 3")
            others_ssf.append(srv.submit_ssf(ids[k], qk))
        if rep == 20:
            mixed = srv.submit_ssf(ids[0], q)
    got, _ = srv.wait_ssf(mixed)
    assert np.array_equal(got, alone)
    assert all(np.isfinite(srv.wait(r)["dU"]) for r in others)
    assert all(np.all(np.isfinite(srv.wait_ssf(r)[0])) for r in others_ssf)
    # a committed move, then S(q) on the same chain: the S(q) sees the move
    sp, new = _segment(rng, w0["M"], w0["N"], L0, b0, 9)
    r_move = srv.submit(ids[0], sp, new, commit_mode=2)
    r_sq = srv.submit_ssf(ids[0], q)
    assert srv.wait(r_move)["accepted"]
    after, _ = srv.wait_ssf(r_sq)
    ctx.update_beads(sp, new)
    assert np.array_equal(after, ctx.ssf(q))
    assert not np.array_equal(after, alone)
    for (m, i), p in zip(sp, new):
        st.update_bead(int(m), int(i), p)
    ref2 = st.ssf(q[pick])
    assert np.all(np.abs(after[pick] - ref2) <= 1e-5 * np.maximum(ref2, w0["M"]))
    # errors
    with pytest.raises(pimc_amd.PimcError, match="lattice"):
        srv.submit_ssf(ids[0], [[0.3, 0.1, 0.2]])
    rid = srv.submit_ssf(ids[1], q[:3] * L0 / chains[1][0]["side"])
    with pytest.raises(pimc_amd.PimcError, match="wait_ssf"):
        srv.wait(rid)
    assert srv.wait_ssf(rid)[0].shape == (3,)
    srv.close()


def test_server_request_errors(orc):
    """Requests the server refuses with an error and nothing queued: an unknown
    chain (below zero, past the last one), a removed chain, a bead locator outside
    the chain, two beads on one slice; a chain added after the first requests is
    reachable at once (its extents published to the submit path)."""
    import pimc_amd
    w, b = walker(orc, 32, 16, 1.5, SEED0 + 7, kind=0)
    srv = pimc_amd.Server(1 << 27, streams=2, max_side=40.0)
    c0 = srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    sp = np.array([[1, 3], [2, 3]], dtype=np.int32)
    new = b[sp[:, 0], sp[:, 1]] + 0.05
    for chain in (-1, c0 + 1):
        with pytest.raises(pimc_amd.PimcError, match="unknown chain"):
            srv.submit(chain, sp, new)
    with pytest.raises(pimc_amd.PimcError, match="out of range"):
        srv.submit(c0, np.array([[w["M"], 0]], dtype=np.int32), new[:1])
    with pytest.raises(pimc_amd.PimcError, match="distinct slices"):
        srv.submit(c0, np.array([[1, 3], [1, 4]], dtype=np.int32), new)
    r0 = srv.wait(srv.submit(c0, sp, new))
    c1 = srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    r1 = srv.wait(srv.submit(c1, sp, new))
    assert np.array_equal(r1["raw"][:4], r0["raw"][:4])
    srv.remove_chain(c1)
    with pytest.raises(pimc_amd.PimcError, match="unknown chain"):
        srv.submit(c1, sp, new)
    assert srv.stats()["requests"] == 2
    srv.close()
