"""The batched and serving forms at their stated sizes.

Batched: the 512 walkers of the batched workload (N in {16..256}, M in
{64..512}, T in [1, 3] K, FP32-rounded Gaussian bridges), autotuned tiles, the
launch sequence replayed as a hipGraph.  Per-walker potentialAction() against the
oracle (src/action.cpp:456-472) at rel <= 1e-6 on two walkers of every (N, M)
class, every walker's graph replays bit-identical to the eager launch with the
same tiles, and the pair count of the whole set.

Serving: the 1024 chains of the serving workload (N in [32, 512], M in [64,
512]) resident in one arena.  Segment (Staging/Bisection) and CenterOfMass Delta
V requests (src/move.cpp:1124-1259, :1315-1408) on chains spanning the size
range against the oracle at |dU - dU_ref| <= 1e-6 (|A_old| + |A_new|), then the
same requests inside a mix of requests on every chain: bit-identical results;
an S(q) request longer than one slot's density-mode buffer equal to pimc_ssf."""
import collections

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAMBDA = 6.0596


def _oracle_state(orc, w, b):
    st = orc.State(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"])
    st.set_beads(b)
    return st


def test_batch_512_walkers(orc):
    import pimc_amd
    from pimc_amd.workloads import batched_walkers
    ws = batched_walkers()
    beads = [pimc_amd.gen_worldlines(w["seed"], w["side"], w["M"], w["N"], LAMBDA, w["tau"], w["kind"])
             .astype(np.float32).astype(np.float64) for w in ws]
    bt = pimc_amd.Batch(ws)
    for k, b in enumerate(beads):
        bt.upload(k, b)
    inf = bt.info()
    assert inf["total_pairs"] == sum(w["M"] * w["N"] * (w["N"] - 1) // 2 for w in ws)
    tune = bt.autotune(reps=1)
    print(f"512 walkers: {inf['total_pairs']} pairs, tiles {tune}", flush=True)
    eager = bt.potentialAction()
    bt.use_graph(True)
    for rep in range(3):
        g = bt.potentialAction()
        assert np.array_equal(g, eager), rep
    # two walkers of every (N, M) class against the oracle
    classes = collections.defaultdict(list)
    for k, w in enumerate(ws):
        classes[(w["N"], w["M"])].append(k)
    assert len(classes) == 20
    worst = 0.0
    for key in sorted(classes):
        for k in (classes[key][0], classes[key][-1]):
            ref = _oracle_state(orc, ws[k], beads[k]).total_action()
            err = abs(g[k] - ref) / abs(ref)
            worst = max(worst, err)
            assert err <= 1e-6, (k, key, g[k], ref)
    print(f"512 walkers: worst rel err over the 40 checked walkers {worst:.3e}", flush=True)


def _segment(rng, w, b, seg):
    M, N, L = w["M"], w["N"], w["side"]
    start, p = int(rng.integers(M)), int(rng.integers(N))
    sp = np.array([[(start + k) % M, p] for k in range(1, seg)], dtype=np.int32)
    new = b[sp[:, 0], sp[:, 1]] + rng.normal(scale=np.sqrt(LAMBDA * w["tau"]), size=(seg - 1, 3))
    new -= L * np.floor(new / L + 0.5)
    return sp, new


def _com(rng, w, b):
    M, N, L = w["M"], w["N"], w["side"]
    p = int(rng.integers(N))
    sp = np.array([[m, p] for m in range(M)], dtype=np.int32)
    new = b[:, p] + rng.normal(scale=0.1, size=3)
    new -= L * np.floor(new / L + 0.5)
    return sp, new


def _oracle_dv(orc, w, b, sp, new):
    st = _oracle_state(orc, w, b)
    old = st.action_list(sp)
    for (m, i), q in zip(sp, new):
        st.update_bead(int(m), int(i), q)
    return old, st.action_list(sp)


def test_server_1024_chains(orc):
    import pimc_amd
    from pimc_amd.workloads import box_side, served_chains
    ws = served_chains()
    srv = pimc_amd.Server(12 << 30, streams=4, max_side=box_side(512) + 1.0)
    beads = {}
    by_size = sorted(range(len(ws)), key=lambda k: ws[k]["M"] * ws[k]["N"])
    picks = [by_size[0], by_size[1], by_size[len(ws) // 2], by_size[-2], by_size[-1],
             min(range(len(ws)), key=lambda k: ws[k]["N"]), max(range(len(ws)), key=lambda k: ws[k]["M"])]
    for k, w in enumerate(ws):
        b = pimc_amd.gen_worldlines(w["seed"], w["side"], w["M"], w["N"], LAMBDA, w["tau"], w["kind"])
        cid = srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
        assert cid == k
        if k in picks:
            beads[k] = b
    assert all(srv.chain_state(k) == 1 for k in range(len(ws)))
    print(f"1024 chains resident, {srv.stats()['free_bytes'] / 2**30:.2f} GiB of the arena free", flush=True)
    rng = np.random.default_rng(11)
    reqs = []
    for k in picks:
        w, b = ws[k], beads[k]
        reqs.append((k,) + _segment(rng, w, b, int(rng.integers(4, 65))))
        reqs.append((k,) + _com(rng, w, b))
    alone = []
    worst = 0.0
    for k, sp, new in reqs:
        r = srv.wait(srv.submit(k, sp, new))
        alone.append(r)
        old, newa = _oracle_dv(orc, ws[k], beads[k], sp, new)
        scale = abs(old) + abs(newa)
        err = abs(r["dU"] - (newa - old)) / scale
        worst = max(worst, err)
        assert err <= 1e-6, (k, len(sp), r["dU"], newa - old)
        assert abs(r["A_old"] - old) <= 1e-6 * scale and abs(r["A_new"] - newa) <= 1e-6 * scale
    print(f"served dV against the oracle: worst |dU - ref| / (|A_old| + |A_new|) {worst:.3e}", flush=True)
    # the same requests inside a mix spread over every chain (evaluate only)
    mix = []
    for rep in range(3):
        for c in range(len(ws)):
            w = ws[c]
            seg = int(rng.integers(4, 65))
            start, p = int(rng.integers(w["M"])), int(rng.integers(w["N"]))
            sp = np.array([[(start + t) % w["M"], p] for t in range(1, min(seg, w["M"]))], dtype=np.int32)
            new = rng.uniform(-0.5 * w["side"], 0.5 * w["side"], size=(len(sp), 3))
            mix.append(srv.submit(c, sp, new))
        if rep == 1:
            again = [srv.submit(k, sp, new) for k, sp, new in reqs]
    res = [srv.wait(r) for r in mix]
    assert all(np.isfinite(r["dU"]) for r in res)
    for a, r in zip(alone, again):
        m = srv.wait(r)
        assert np.array_equal(m["raw"][:4], a["raw"][:4])
    st = srv.stats()
    assert st["requests"] == 2 * len(reqs) + len(mix)
    # S(q) on the chain with the most slices, with more q than one slot's
    # rho buffer holds at that M: the same as pimc_ssf on a context
    from pimc_amd.workloads import lattice_q
    k = max(picks, key=lambda c: ws[c]["M"])
    w = ws[k]
    q = lattice_q(w["side"], nmax=10)                    # the 9261 q of the cube |n_d| <= 10
    assert q.shape[0] * w["M"] * 16 > 32 << 20
    sq, lat = srv.wait_ssf(srv.submit_ssf(k, q))
    ctx = pimc_amd.Context(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"])
    ctx.upload_path(beads[k])
    assert np.array_equal(sq, ctx.ssf(q))
    print(f"S(q) request on chain {k} (N={w['N']}, M={w['M']}): {q.shape[0]} q in {lat:.0f} us", flush=True)
    srv.close()
