"""Short Delta V requests dispatched as AQL packets straight into the device's
queue, against the same requests launched through HIP (PIMC_DIRECT_DISPATCH=0):
bit-identical results, commits, cell tables and total actions in between, and
the oracle's values (tolerance |dU - dU_ref| <= 1e-6 (|A_old| + |A_new|) in the
mixed mode, 1e-12 in the table mode)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853


@pytest.fixture(scope="module")
def gpu():
    import pimc_amd
    return pimc_amd


def _context(gpu, L, M, N, tau, rc, mode, beads, packets):
    old = os.environ.get("PIMC_DIRECT_DISPATCH")
    try:
        if packets:
            os.environ.pop("PIMC_DIRECT_DISPATCH", None)
        else:
            os.environ["PIMC_DIRECT_DISPATCH"] = "0"
        ctx = gpu.Context(L, M, N, tau, LAMBDA, rc, v_mode=mode)
    finally:
        if old is None:
            os.environ.pop("PIMC_DIRECT_DISPATCH", None)
        else:
            os.environ["PIMC_DIRECT_DISPATCH"] = old
    ctx.upload_path(beads)
    return ctx


def _requests(rng, beads, L, M, N, tau, n_req, segs=(8, 16, 32)):
    out = []
    for k in range(n_req):
        seg = segs[k % len(segs)]
        start, p = int(rng.integers(M)), int(rng.integers(N))
        sp = np.array([[(start + j) % M, p] for j in range(1, seg)], dtype=np.int32)
        new = beads[sp[:, 0], sp[:, 1]] + rng.normal(scale=np.sqrt(LAMBDA * tau) * 2, size=(seg - 1, 3))
        new -= L * np.floor(new / L + 0.5)
        out.append((sp, new, bool(k % 3 == 2)))
    return out


@pytest.mark.parametrize("mode", [0, 2])
def test_packets_match_hip_launches(orc, gpu, mode):
    """400 Bisection requests of 7/15/31 beads on N=256 x 167 (every third one
    committed), a total action after each commit: packets and HIP launches agree
    bit for bit (each of the 64 argument slots is rewritten through the BAR six
    times, with other requests' contents), the packet context sent every request
    as a packet, and the paths and cell tables end identical."""
    N, M, tau = 256, 167, 0.004
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(SEED0 + 21, L, M, N, LAMBDA, tau, 0)
    rc = L / 2 if mode == 0 else L
    a = _context(gpu, L, M, N, tau, rc, mode, beads, packets=True)
    b = _context(gpu, L, M, N, tau, rc, mode, beads, packets=False)
    reqs = _requests(np.random.default_rng(5), beads, L, M, N, tau, 400)
    for sp, new, commit in reqs:
        oa = a.delta_action(sp, new, commit=commit)
        ob = b.delta_action(sp, new, commit=commit)
        assert np.array_equal(oa, ob), (sp[0], commit, oa, ob)
        if commit:
            assert a.potentialAction() == b.potentialAction()
    assert a.kernel_time("delta_packets")[1] == len(reqs)
    assert b.kernel_time("delta_packets")[1] == 0
    assert np.array_equal(a.download_path(), b.download_path())
    ca, cb = a.cell_lists(), b.cell_lists()
    for key in ("dims", "grid", "bead_label", "num_labels"):
        assert np.array_equal(ca[key], cb[key]), key
    a.close()
    b.close()


@pytest.mark.parametrize("seg", [8, 16, 32])
@pytest.mark.parametrize("mode", [0, 2])
def test_packets_against_oracle(orc, gpu, seg, mode):
    """Packet-dispatched Delta V of Bisection segments on N=256 x 167 against the
    oracle's bead actions, then committed: the downloaded path is the oracle's
    and the next total action matches it."""
    N, M, tau = 256, 167, 0.004
    L = orc.box_side(N, DENSITY)
    for seed in range(3):
        beads = orc.gen_worldlines(SEED0 + 40 + seed, L, M, N, LAMBDA, tau, 0)
        ref = orc.State(L, M, N, tau, LAMBDA, L)
        ref.set_beads(beads)
        ctx = _context(gpu, L, M, N, tau, L, mode, beads, packets=True)
        sp, new, _ = _requests(np.random.default_rng(seed), beads, L, M, N, tau, 1, segs=(seg,))[0]
        a_old = ref.action_list(sp)
        for (m, i), q in zip(sp, new):
            ref.update_bead(int(m), int(i), q)
        a_new = ref.action_list(sp)
        out = ctx.delta_action(sp, new, commit=True)
        assert ctx.kernel_time("delta_packets")[1] == 1
        scale = abs(a_old) + abs(a_new)
        tol = 1e-6 if mode == 0 else 1e-12
        assert abs(out[0] - a_old) <= tol * scale
        assert abs(out[1] - a_new) <= tol * scale
        assert abs(out[2] - (a_new - a_old)) <= tol * scale, (out, a_old, a_new)
        assert np.array_equal(ctx.download_path(), ref.beads())
        t = ref.total_action()
        assert abs(ctx.potentialAction() - t) <= (1e-6 if mode == 0 else 1e-12) * abs(t)
        ctx.close()


def test_packets_with_timing_and_many_contexts(orc, gpu):
    """Timed Delta V calls go through HIP (their events live on the stream);
    several contexts alive at once share the device queue, each with its own
    argument ring, and closing one leaves the others working."""
    N, M, tau = 128, 64, 0.004
    L = orc.box_side(N, DENSITY)
    live = []
    for k in range(4):
        beads = orc.gen_worldlines(SEED0 + 60 + k, L, M, N, LAMBDA, tau, 0)
        live.append((_context(gpu, L, M, N, tau, L, 2, beads, packets=True),
                     _context(gpu, L, M, N, tau, L, 2, beads, packets=False),
                     _requests(np.random.default_rng(k), beads, L, M, N, tau, 30)))
    for r in range(30):
        if r == 20:
            a, b, _ = live.pop(0)
            a.close()
            b.close()
        for a, b, reqs in live:
            sp, new, commit = reqs[r]
            if r == 10:
                a.set_timing(True, kinds=["delta_action"])
            if r == 12:
                a.set_timing(False)
            assert np.array_equal(a.delta_action(sp, new, commit=commit), b.delta_action(sp, new, commit=commit))
    assert len(live) == 3
    for a, b, _ in live:
        assert a.kernel_time("delta_packets")[1] == 30 - 2   # the two timed calls went through HIP
        assert np.array_equal(a.download_path(), b.download_path())
        a.close()
        b.close()


def _server(gpu, budget, packets, **kw):
    old = os.environ.get("PIMC_DIRECT_DISPATCH")
    try:
        if packets:
            os.environ.pop("PIMC_DIRECT_DISPATCH", None)
        else:
            os.environ["PIMC_DIRECT_DISPATCH"] = "0"
        return gpu.Server(budget, **kw)
    finally:
        if old is None:
            os.environ.pop("PIMC_DIRECT_DISPATCH", None)
        else:
            os.environ["PIMC_DIRECT_DISPATCH"] = old


def test_server_packets_match_hip_launches(orc, gpu):
    """The serving form with its Delta V batches as AQL packets (descriptors in
    device memory) against a server launching every batch through HIP: the same
    requests -- one at a time, then 60 at once in mixed batches, commits and
    Metropolis decisions among them, batches of more than 64 KB of descriptors
    going through HIP in the packet server too -- give bit-identical results and
    leave bit-identical chains."""
    specs = [(32, 40), (64, 32), (128, 24), (200, 16), (256, 64), (96, 48)] + [(64, 64)] * 18
    chains = []
    for k, (N, M) in enumerate(specs):
        L = orc.box_side(N, DENSITY)
        tau = 1.0 / (M * 1.5)
        b = orc.gen_worldlines(SEED0 + 80 + k, L, M, N, LAMBDA, tau, 0)
        chains.append((L, M, N, tau, b))
    srv = {p: _server(gpu, 1 << 28, p, streams=4, max_side=60.0) for p in (True, False)}
    ids = {p: [s.add_chain(L, M, N, tau, LAMBDA, L, b) for (L, M, N, tau, b) in chains] for p, s in srv.items()}
    rng = np.random.default_rng(3)

    def request(k, nb):
        L, M, N, tau, b = chains[k]
        start, p = int(rng.integers(M)), int(rng.integers(N))
        sp = np.array([[(start + j) % M, p] for j in range(1, nb + 1)], dtype=np.int32)
        new = b[sp[:, 0], sp[:, 1]] + rng.normal(scale=0.3, size=(nb, 3))
        new -= L * np.floor(new / L + 0.5)
        mode = int(rng.integers(3))
        return k, sp, new, mode, float(rng.uniform())

    # one at a time: every batch holds one request
    for _ in range(30):
        k, sp, new, mode, u = request(int(rng.integers(len(chains))), int(rng.integers(3, 16)))
        out = {p: s.wait(s.submit(ids[p][k], sp, new, commit_mode=mode, u=u)) for p, s in srv.items()}
        assert np.array_equal(out[True]["raw"][:4], out[False]["raw"][:4]), (k, mode)
    # sixty at once, segments up to the longest slice count (large batches among them)
    reqs = [request(int(rng.integers(len(chains))), int(rng.integers(4, 16))) for _ in range(60)]
    for p, s in srv.items():
        rid = [s.submit(ids[p][k], sp, new, commit_mode=mode, u=u) for (k, sp, new, mode, u) in reqs]
        reqs_out = [s.wait(r)["raw"][:4].copy() for r in rid]
        srv[p] = (s, reqs_out)
    for a, b in zip(srv[True][1], srv[False][1]):
        assert np.array_equal(a, b)
    # long segments on every chain at once (a batch holds one request per chain):
    # 24 requests of up to 63 beads pass 64 KB of descriptors
    big = []
    for rep in range(4):
        for k in range(len(chains)):
            M = chains[k][1]
            big.append(request(k, min(M - 1, 63)))
    res = {}
    for p in (True, False):
        s = srv[p][0]
        rid = [s.submit(ids[p][k], sp, new, commit_mode=0) for (k, sp, new, mode, u) in big]
        res[p] = [s.wait(r)["raw"][:3].copy() for r in rid]
    for a, b in zip(res[True], res[False]):
        assert np.array_equal(a, b)
    for k in range(len(chains)):
        assert np.array_equal(srv[True][0].download(ids[True][k]), srv[False][0].download(ids[False][k]))
    dp, dh = srv[True][0].dispatch_stats(), srv[False][0].dispatch_stats()
    assert dp["packet_batches"] >= 30 and dh["packet_batches"] == 0 and dh["hip_batches"] >= 30
    print("dispatch", dp, dh)
    for p in (True, False):
        srv[p][0].close()
