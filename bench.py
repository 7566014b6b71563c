"""Benchmark of the plain one-walker drop-in on the single-walker workload:
3D He-4, N=256 at 0.02198 A^-3, T=1.5 K, tau=0.004 (M=167), Aziz HFDHE2,
primitive action, FP32 pair terms with FP64 sums.

One step = one pass of the hot path over one seeded configuration:
  the full potential action (all 5,450,880 i<j slice pairs) and the Delta V of
  three Bisection segments of 8/16/32 slices at random start slice and particle
  (old + new partner pair evaluations, 3570 / 7650 / 15810 pairs).
value = bead-pair evaluations per second over the timed steps.  F(q,tau) for
all 5377 lattice q with |q| <= 3 A^-1 and all 167 lags is timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
F_PAIR = 50          # algorithmic FP32 flops per bead-pair evaluation
# Delta V reads each partner once for the old and the new position (two pair
# evaluations): FP64 x, y, z, its on-byte and its cell number
B_PARTNER = 3 * 8 + 1 + 4
TIMING_STRIDE = 8    # plain leg: the pair kernel of every 8th timed step carries HIP events
F_QB = 9             # per (q, bead) phasor
F_QMT = 4            # per (q, m, tau) lag product


def pct(a, p):
    return float(np.percentile(np.asarray(a), p)) if len(a) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--isf-reps", type=int, default=5)
    ap.add_argument("--legs", default="plain,large,gsf,virial,chain,batched,served",
                    help="comma list of plain, large, gsf, virial, chain, batched, served (plain always runs)")
    ap.add_argument("--served-seconds", type=float, default=2.0)
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="also run a plain copy kernel of this many bytes (HBM counter scaling)")
    ap.add_argument("--pair-only", type=int, nargs=2, metavar=("N", "M"), default=None,
                    help="profiling driver: only the full potential action of N beads x M slices, 5 calls "
                         "(mixed mode, rc = L, or L/2 from N = 2048), no bench line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU sharding is a design note "
                                   "(--gpus must be 1)"}))
        sys.exit(2)

    import pimc_amd
    from pimc_amd import worldlines
    from pimc_amd.workloads import DENSITY, LAMBDA, SEED0, box_side, lattice_q
    if args.pair_only:
        N, M = args.pair_only
        L = box_side(N, DENSITY)
        b = worldlines.generate(SEED0, L, M, N, LAMBDA, 0.004, 1)
        ctx = pimc_amd.Context(L, M, N, 0.004, LAMBDA, L / 2 if N >= 2048 else L, v_mode=pimc_amd.V_FP32)
        ctx.upload_path(b)
        for _ in range(5):
            a = ctx.potentialAction()
        print(json.dumps({"pair_only": [N, M], "action": a}), flush=True)
        return

    N, T, tau, M = 256, 1.5, 0.004, 167      # M as the workload states (the reference would force 166)
    L = box_side(N, DENSITY)
    beads = worldlines.generate(SEED0, L, M, N, LAMBDA, tau, 0)
    ctx = pimc_amd.Context(L, M, N, tau, LAMBDA, L, v_mode=pimc_amd.V_FP32)
    ctx.upload_path(beads)
    q = lattice_q(L, qmag_max=3.0)

    rng = np.random.default_rng(SEED0)
    segs = (8, 16, 32)

    def make_request(seg):
        start, p = int(rng.integers(M)), int(rng.integers(N))
        sp = np.array([[(start + k) % M, p] for k in range(1, seg)], dtype=np.int32)
        new = beads[sp[:, 0], sp[:, 1]] + rng.normal(scale=math.sqrt(LAMBDA * tau), size=(seg - 1, 3))
        new -= L * np.floor(new / L + 0.5)
        return sp, new

    pairs_full = sum(n * (n - 1) // 2 for n in [N] * M)
    pairs_dv = {s: 2 * (s - 1) * (N - 1) for s in segs}
    pairs_step = pairs_full + sum(pairs_dv.values())

    # requests are drawn (and converted for the library) before the timed
    # region, so it holds only the hot path
    n_req = args.warmup + args.steps
    reqs = [[make_request(s) for s in segs] for _ in range(n_req)]
    preps = [[ctx.prepare_delta(sp, new) for sp, new in r] for r in reqs]

    def step(k, lat, timed=False):
        if timed:
            ctx.set_timing(True, kinds=["pair_action"])
        ctx.potentialAction()
        if timed:
            ctx.set_timing(False)
        for s, prep in zip(segs, preps[k]):
            t0 = time.perf_counter()
            ctx.delta_action_prepared(prep)
            lat[s].append((time.perf_counter() - t0) * 1e6)

    lat = {s: [] for s in segs}
    for k in range(args.warmup):
        step(k, lat)
    lat = {s: [] for s in segs}
    ctx.set_timing(True, kinds=["pair_action"])   # creates the event pool ahead of the timed region
    ctx.set_timing(False)
    # the all-pairs kernel's own duration is stamped by HIP events (from a
    # pre-created pool) around its dispatch on every TIMING_STRIDE-th step of
    # the timed region itself; the other steps and the Delta V launches carry no
    # event records (on every step the records cost ~4 us of host time a step)
    pimc_amd.device_synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k, lat, k % TIMING_STRIDE == 0)
    pimc_amd.device_synchronize()
    elapsed = time.perf_counter() - t0
    pa_ms, pa_n = ctx.kernel_time("pair_action")
    # the Delta V kernel's duration, from a second pass over the same requests
    ctx.set_timing(True, kinds=["delta_action"])
    for k in range(args.steps):
        for prep in preps[args.warmup + k]:
            ctx.delta_action_prepared(prep)
    dv_ms, dv_n = ctx.kernel_time("delta_action")
    ctx.set_timing(False)

    # F(q, tau), timed beside the step, into a page-locked result array
    # allocated once (and once into a fresh pageable array)
    fqt = pimc_amd.pinned_empty((q.shape[0], M))
    ctx.isf(q, out=fqt)
    pimc_amd.device_synchronize()
    t1 = time.perf_counter()
    for _ in range(args.isf_reps):
        ctx.isf(q, out=fqt)
    isf_s = (time.perf_counter() - t1) / args.isf_reps
    t1 = time.perf_counter()
    ctx.isf(q)
    isf_pageable_s = time.perf_counter() - t1
    # the kernels' durations, from a timed pass after the wall-clock one
    ctx.set_timing(True, kinds=["rho", "isf"])
    for _ in range(args.isf_reps):
        ctx.isf(q, out=fqt)
    rho_ms, rho_n = ctx.kernel_time("rho")
    isf_ms, isf_n = ctx.kernel_time("isf")
    ctx.set_timing(False)

    value = pairs_step * args.steps / elapsed
    pa_avg_s = pa_ms / max(pa_n, 1) / 1e3
    achieved = pairs_full * F_PAIR / pa_avg_s / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pair_action_n256_m167_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

    line = {
        "metric": "bead-pair evals/s + p50/p99 dV latency, served vs plain GPU at <=1e-6 rel err",
        "value": value,
        "unit": "bead-pair evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 transcendentals (exp2/rcp/rsq) + fp64 arithmetic and sums",
        "data": "synthetic seeded worldlines (MTRand seed 139853: uniform slice 0 + Gaussian steps)",
        "config": {"workload": "single walker: He-4 N=256, rho=0.02198 A^-3, T=1.5 K (M tau = 1/T), tau=0.004, M=167, "
                               "Aziz 1979, primitive action, full action + Bisection dV 8/16/32",
                   "N": N, "M": M, "form": "plain one-walker drop-in (C-ABI)"},
        "roofline": {"bound": "valu", "kernel": "k_pair_action<mixed>",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "avg_launch_us": pa_avg_s * 1e6, "flop_per_pair": F_PAIR, "pairs_per_launch": pairs_full,
                     "timed_launches": pa_n,
                     "timing": f"HIP events around the pair kernel of every {TIMING_STRIDE}th step of the timed region"},
        "delta_v": {str(s): {"pairs": pairs_dv[s], "p50_us": pct(lat[s], 50), "p99_us": pct(lat[s], 99)}
                    for s in segs},
        "delta_v_pair_kernel_avg_us": dv_ms / max(dv_n, 1) * 1e3,
        # latency-bound: the partner bytes the kernel must read against its duration
        "delta_v_hbm_gb_per_s_algorithmic": (sum(pairs_dv.values()) / 2 * B_PARTNER)
                                            / (dv_ms / max(dv_n, 1) * 1e-3 * len(segs)) / 1e9,
        "isf": {"n_q": int(q.shape[0]), "lags": M, "ms": isf_s * 1e3, "ms_pageable_result": isf_pageable_s * 1e3,
                "rho_kernel_ms": rho_ms / max(rho_n, 1), "lag_kernel_ms": isf_ms / max(isf_n, 1),
                "q_bead_phasors_per_s": q.shape[0] * M * N / isf_s,
                "tflops_algorithmic": (q.shape[0] * M * N * F_QB + q.shape[0] * M * M * F_QMT) / isf_s / 1e12},
        "served": None,
    }

    legs = set(args.legs.split(","))
    dv_pairs_total = sum(pairs_dv[s] * len(lat[s]) for s in segs)
    dv_us_total = sum(sum(lat[s]) for s in segs)
    plain_dv_rate = dv_pairs_total / (dv_us_total * 1e-6)
    line["plain_delta_v_pair_evals_per_s"] = plain_dv_rate
    # the same calls from a C++ caller (no Python in the loop), as the reference's moves make them
    cxx = cxx_client(N, M, tau)
    if cxx:
        line["delta_v_cxx"] = cxx
        plain_dv_rate = sum(pairs_dv[s] for s in segs) / (sum(cxx[f"dv{s}_mean_us"] for s in segs) * 1e-6)
        line["plain_delta_v_pair_evals_per_s_cxx"] = plain_dv_rate
    if "gsf" in legs:
        line["gsf"] = gsf_leg(pimc_amd, beads, L, M, N, tau)
    if "virial" in legs:
        line["virial"] = virial_leg(pimc_amd, beads, L, M, N, tau, not args.no_cpu_baseline)
    if "chain" in legs:
        line["chain"] = chain_leg(pimc_amd, beads, L, M, N, tau, not args.no_cpu_baseline)
    if "large" in legs:
        line["large"] = large_leg(pimc_amd, not args.no_cpu_baseline)
    plain_refs = {"step": value, "pair_kernel": pairs_full / pa_avg_s}
    if cxx:
        plain_refs["cxx_total_action"] = pairs_full / (cxx["total_action_p50_us"] * 1e-6)
    if "batched" in legs:
        line["batched"] = batched_leg(pimc_amd, plain_refs)
    if "served" in legs:
        line["served"] = served_leg(pimc_amd, plain_dv_rate, args.served_seconds)
    if args.probe_copy:
        line["copy_probe"] = {"bytes": args.probe_copy,
                              "avg_ms": pimc_amd.api.copy_probe(args.probe_copy, 5)}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(beads, L, M, N, tau, LAMBDA, pairs_full, pairs_dv)
    print(json.dumps(line))


def cxx_client(N, M, tau, requests=2000, action="primitive"):
    """pimc_amd/dv_client: p50/p99 of pimc_delta_action and pimc_total_action from C++."""
    import subprocess
    exe = os.path.join(ROOT, "pimc_amd", "dv_client")
    if not os.path.exists(exe):
        return None
    cmd = [exe, str(N), str(M), str(tau), "139853", str(requests)] + (["gsf"] if action == "gsf" else [])
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def gsf_leg(pimc_amd, beads, L, M, N, tau, reps=50):
    """The same walker under the GSF action (VFactor {2/3, 4/3}, gradVFactor
    {0, 2/9}): the full action adds gradVSquared per slice (N(N-1) gradV pairs per
    slice), every shift-1 bead action the three-body gradVnnSquared."""
    from pimc_amd.workloads import LAMBDA
    ctx = pimc_amd.Context(L, M, N, tau, LAMBDA, L, v_mode=pimc_amd.V_FP32)
    ctx.upload_path(beads)
    ctx.set_action_factors((2.0 / 3.0, 4.0 / 3.0), (0.0, 2.0 / 9.0), 2)
    for _ in range(5):
        ctx.potentialAction()
    ctx.set_timing(True)
    pimc_amd.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.potentialAction()
    pimc_amd.device_synchronize()
    wall = (time.perf_counter() - t0) / reps
    pa_ms, pa_n = ctx.kernel_time("pair_action")
    g2_ms, g2_n = ctx.kernel_time("gradv2")
    ctx.set_timing(False)
    grad_pairs = M * N * (N - 1)
    out = {"action": "gsf", "total_action_wall_us": wall * 1e6,
           "pair_kernel_us": pa_ms / max(pa_n, 1) * 1e3, "gradv2_kernel_us": g2_ms / max(g2_n, 1) * 1e3,
           "gradv2_pair_evals": grad_pairs,
           "gradv2_pair_evals_per_s": grad_pairs / (g2_ms / max(g2_n, 1) * 1e-3) if g2_n else None}
    ctx.close()
    cxx = cxx_client(N, M, tau, requests=500, action="gsf")
    if cxx:
        out["cxx"] = cxx
    return out


def virial_leg(pimc_amd, beads, L, M, N, tau, cpu, window=5, reps=20):
    """VirialEnergyEstimator::accumulate on the same walker under the GSF action
    (virial window 5): V(slice) from the pair kernel plus the virial kernel's
    N(N-1) ordered pairs per slice (gradV on every slice, the T-matrix on the
    odd slices where gradVFactor = 2/9)."""
    from pimc_amd.workloads import LAMBDA
    ctx = pimc_amd.Context(L, M, N, tau, LAMBDA, L, v_mode=pimc_amd.V_FP32)
    ctx.upload_path(beads)
    ctx.set_action_factors((2.0 / 3.0, 4.0 / 3.0), (0.0, 2.0 / 9.0), 2)
    for _ in range(3):
        ctx.virial_energy(window)
    ctx.set_timing(True)
    pimc_amd.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        e = ctx.virial_energy(window)
    wall = (time.perf_counter() - t0) / reps
    v_ms, v_n = ctx.kernel_time("virial")
    pa_ms, pa_n = ctx.kernel_time("pair_action")
    ctx.set_timing(False)
    ctx.close()
    pairs = M * N * (N - 1)
    k_s = v_ms / max(v_n, 1) * 1e-3
    out = {"action": "gsf", "window": window, "estimator_wall_us": wall * 1e6,
           "virial_kernel_us": k_s * 1e6, "pair_kernel_us": pa_ms / max(pa_n, 1) * 1e3,
           "ordered_pairs_per_launch": pairs, "virial_pair_evals_per_s": pairs / k_s if v_n else None,
           "E": e["E"], "E_th": e["E_th"]}
    if cpu:
        # the oracle's accumulate on the first 24 slices of the same worldlines
        # (periodic over those 24), one call, single thread
        from oracle import pyoracle as orc
        Ms = 24
        st = orc.State(L, Ms, N, tau, LAMBDA, L)
        st.set_beads(np.ascontiguousarray(beads[:Ms]))
        st.set_action(1, (2.0 / 3.0, 4.0 / 3.0))
        st.set_correction((0.0, 2.0 / 9.0), 2)
        st.virial_energy(window)            # builds the derivative tables
        t0 = time.perf_counter()
        st.virial_energy(window)
        t_cpu = time.perf_counter() - t0
        out["cpu_port"] = {"sample": f"{Ms} slices of N={N}, one accumulate", "seconds": t_cpu, "cores": 1,
                           "virial_pair_evals_per_s": Ms * N * (N - 1) / t_cpu,
                           "estimator_us_scaled_to_M": t_cpu * M / Ms * 1e6}
    return out


def chain_leg(pimc_amd, beads, L, M, N, tau, cpu, moves=3000):
    """One Markov chain on the same walker: Staging (Mbar 16), Bisection (4
    levels) and CenterOfMass (0.1 A) moves in rotation, primitive action.
    Host-driven (proposals on the host, one device Delta V pass per move) against
    device-resident (pimc_chain_run_device: the whole run in one launch)."""
    from pimc_amd.workloads import LAMBDA
    kinds = [k % 3 for k in range(moves)]
    ctx = pimc_amd.Context(L, M, N, tau, LAMBDA, L, v_mode=pimc_amd.V_FP32)
    ctx.upload_path(beads)
    ch = pimc_amd.Chain(ctx, 4357, 16, 4, True, 1, 0.1)
    host_moves = min(600, moves)
    names = ("staging", "bisection", "com")
    for k in range(30):
        ch.move(names[k % 3])
    pimc_amd.device_synchronize()
    t0 = time.perf_counter()
    for k in range(host_moves):
        ch.move(names[k % 3])
    t_host = time.perf_counter() - t0
    ch.run_device(kinds[:60])
    ctx.set_timing(True)
    t0 = time.perf_counter()
    acc, info, near = ch.run_device(kinds)
    t_dev = time.perf_counter() - t0
    k_ms, k_n = ctx.kernel_time("chain_run")
    ctx.set_timing(False)
    ctx.close()
    # an ensemble of independent chains of the same size, one workgroup each, one launch
    from pimc_amd.workloads import SEED0
    n_ens, ens_moves = 256, 600
    ectx, echains = [], []
    for c in range(n_ens):
        e = pimc_amd.Context(L, M, N, tau, LAMBDA, L, v_mode=pimc_amd.V_FP32)
        e.upload_path(pimc_amd.gen_worldlines(SEED0 + 100 + c, L, M, N, LAMBDA, tau, 1))
        ectx.append(e)
        echains.append(pimc_amd.Chain(e, 5000 + c, 16, 4, True, 1, 0.1))
    pimc_amd.run_chains_device(echains, kinds[:30])
    t0 = time.perf_counter()
    e_acc, _, e_near = pimc_amd.run_chains_device(echains, kinds[:ens_moves])
    t_ens = time.perf_counter() - t0
    for ch in echains:
        ch.close()
    for e in ectx:
        e.close()
    out = {"moves": moves, "mix": "staging / bisection / centre of mass in rotation",
           "ensemble": {"chains": n_ens, "moves_per_chain": ens_moves, "seconds": t_ens,
                        "moves_per_s": n_ens * ens_moves / t_ens, "accept_frac": float(np.mean(e_acc)),
                        "near_threshold_decisions": int(np.sum(e_near))},
           "host_driven_moves_per_s": host_moves / t_host,
           "device_resident_moves_per_s": moves / t_dev, "device_kernel_ms": k_ms / max(k_n, 1),
           "device_us_per_move": t_dev / moves * 1e6, "accept_frac": float(np.mean(acc)),
           "near_threshold_decisions": near}
    if cpu:
        from oracle import pyoracle as orc
        st = orc.State(L, M, N, tau, LAMBDA, L)
        st.set_beads(beads)
        mv = orc.Moves(st, 16, 4, True, 1, 0.1)
        rng = orc.MTRand(4357)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 5.0:
            getattr(mv, names[n % 3])(rng)
            n += 1
        out["cpu_port"] = {"moves_per_s": n / (time.perf_counter() - t0), "moves": n, "cores": 1,
                           "sample": "the oracle's moves on the same walker for 5 s, single thread"}
    return out


def large_leg(pimc_amd, cpu, reps=10, sq_reps=3):
    """The large system: N=2048 at 0.02198 A^-3, T=1.0 K, tau=0.004 (M=250), rc =
    L/2 (2x2x2 LookupTable cells), Gaussian-bridge worldlines.  The full action
    (524 032 000 pair evaluations), S(q) and F(q,tau) on the 100 401 lattice q
    with |q| <= 4 A^-1 (all 250 lags), each with the roofline of its dominant
    kernel from the kernels' own event timing."""
    from pimc_amd.workloads import DENSITY, LAMBDA, SEED0, box_side, lattice_q
    N, M, tau = 2048, 250, 0.004
    L = box_side(N, DENSITY)
    b = pimc_amd.gen_worldlines(SEED0, L, M, N, LAMBDA, tau, 1)
    ctx = pimc_amd.Context(L, M, N, tau, LAMBDA, L / 2, v_mode=pimc_amd.V_FP32)
    ctx.upload_path(b)
    q = lattice_q(L, qmag_max=4.0)
    nq = q.shape[0]
    pairs = M * N * (N - 1) // 2
    for _ in range(2):
        ctx.potentialAction()
    ctx.set_timing(True, kinds=["pair_action"])
    pimc_amd.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        a = ctx.potentialAction()
    wall = (time.perf_counter() - t0) / reps
    pa_ms, pa_n = ctx.kernel_time("pair_action")
    ctx.set_timing(False)
    pa_s = pa_ms / max(pa_n, 1) * 1e-3
    # the same call after every path change (a one-bead update), which also
    # reruns the per-slice spatial ordering pass
    one = np.array([[0, 0]], dtype=np.int32)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.update_beads(one, b[0, 0][None, :])
        a_changed = ctx.potentialAction()
    wall_changed = (time.perf_counter() - t0) / reps
    assert a_changed == a
    ach = pairs * F_PAIR / pa_s / 1e12
    ctx.ssf(q)
    t0 = time.perf_counter()
    for _ in range(sq_reps):
        ctx.ssf(q)
    ssf_wall = (time.perf_counter() - t0) / sq_reps
    # F(q, tau) is 200 MB: into a page-locked result array (allocated once, as a
    # caller keeping its estimator buffers would), and into a fresh pageable one
    fqt = pimc_amd.pinned_empty((nq, M))
    ctx.isf(q, out=fqt)
    t0 = time.perf_counter()
    for _ in range(sq_reps):
        ctx.isf(q, out=fqt)
    isf_wall = (time.perf_counter() - t0) / sq_reps
    t0 = time.perf_counter()
    f_pageable = ctx.isf(q)
    isf_wall_pageable = time.perf_counter() - t0
    assert np.array_equal(f_pageable, fqt)
    del f_pageable
    # the kernels' durations, from timed passes after the wall-clock ones
    ctx.set_timing(True, kinds=["rho", "ssf"])
    for _ in range(sq_reps):
        ctx.ssf(q)
    rho_ms, rho_n = ctx.kernel_time("rho")
    ctx.set_timing(False)
    ctx.set_timing(True, kinds=["rho", "isf"])
    for _ in range(sq_reps):
        ctx.isf(q, out=fqt)
    rho2_ms, rho2_n = ctx.kernel_time("rho")
    lag_ms, lag_n = ctx.kernel_time("isf")
    ctx.set_timing(False)
    del fqt
    ctx.close()
    rho_s = (rho_ms + rho2_ms) / max(rho_n + rho2_n, 1) * 1e-3
    lag_s = lag_ms / max(lag_n, 1) * 1e-3
    rho_flop = nq * M * N * F_QB + nq * M * F_QMT
    out = {"workload": "He-4 N=2048, rho=0.02198 A^-3, T=1.0 K, tau=0.004, M=250, rc=L/2, Aziz 1979, "
                       "Gaussian-bridge worldlines; full action + S(q) + F(q,tau) on 100401 lattice q (|q|<=4 A^-1)",
           "action": a, "pairs_per_launch": pairs, "total_action_wall_ms": wall * 1e3,
           "total_action_wall_ms_after_path_change": wall_changed * 1e3,
           "pair_evals_per_s": pairs / pa_s,
           "roofline_pair": {"bound": "valu", "kernel": "k_pair_window (spatial order, far-pair path)", "achieved": ach,
                             "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": ach / PEAK_FP32_TFLOPS,
                             "avg_launch_us": pa_s * 1e6, "flop_per_pair": F_PAIR},
           "n_q": int(nq), "ssf_wall_ms": ssf_wall * 1e3, "isf_wall_ms": isf_wall * 1e3,
           "isf_wall_ms_pageable_result": isf_wall_pageable * 1e3,
           "roofline_rho": {"bound": "valu", "kernel": "k_rho_lattice", "achieved": rho_flop / rho_s / 1e12,
                            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                            "frac": rho_flop / rho_s / 1e12 / PEAK_FP32_TFLOPS, "avg_launch_ms": rho_s * 1e3,
                            "flop_per_q_bead": F_QB, "q_bead_phasors_per_s": nq * M * N / rho_s},
           "lag_kernel": {"kernel": "k_isf_from_rho", "avg_launch_ms": lag_s * 1e3,
                          "tflops_algorithmic": nq * M * M * F_QMT / lag_s / 1e12}}
    if cpu:
        # the oracle on a bounded sample of the same worldlines: the pair sums of
        # 4 slices, and the direct S(q) pair sum of one q over 4 slices
        from oracle import pyoracle as orc
        Ms = 4
        st = orc.State(L, Ms, N, tau, LAMBDA, L / 2)
        st.set_beads(np.ascontiguousarray(b[:Ms]))
        t0 = time.perf_counter()
        st.total_action()
        t_a = time.perf_counter() - t0
        qk = q[[int(np.argmax(np.abs(q[:, 2])))]]
        t0 = time.perf_counter()
        st.ssf(qk)
        t_s = time.perf_counter() - t0
        out["cpu_port"] = {"sample": f"{Ms} slices of N=2048: the total action's pair sums, and the direct S(q) "
                                     f"pair sum of one q", "cores": 1,
                           "pair_evals_per_s": Ms * N * (N - 1) // 2 / t_a,
                           "ssf_pair_terms_per_s": Ms * N * (N - 1) // 2 / t_s,
                           "full_action_s_scaled": t_a * M / Ms}
    return out


def batched_leg(pimc_amd, plain):
    """512 walkers, N in {16..256}, M in {64..512}, T in [1,3] K, Gaussian-bridge
    worldlines rounded to FP32; autotuned tiles; hipGraph replays."""
    from pimc_amd.workloads import LAMBDA, batched_walkers
    ws = batched_walkers()
    beads = [pimc_amd.api.gen_worldlines(w["seed"], w["side"], w["M"], w["N"], LAMBDA, w["tau"], w["kind"])
             .astype(np.float32).astype(np.float64) for w in ws]
    bt = pimc_amd.Batch(ws)
    for k, b in enumerate(beads):
        bt.upload(k, b)
    bt.use_graph(True)
    tune = bt.autotune(reps=2)
    bt.potentialAction()
    reps = 10
    t0 = time.perf_counter()
    ms = []
    for _ in range(reps):
        bt.potentialAction()
        ms.append(bt.info()["last_ms"])
    wall = (time.perf_counter() - t0) / reps
    info = bt.info()
    dev_ms = float(np.mean(ms))
    rate = info["total_pairs"] / (dev_ms * 1e-3)
    return {"workload": "512 walkers, N in {16,32,64,128,256}, M in {64,128,256,512}, T in [1,3] K, "
                        "tau = 1/(M T), persistent launch, hipGraph replay",
            "pairs_per_launch": info["total_pairs"], "device_ms": dev_ms, "wall_ms": wall * 1e3,
            "pair_evals_per_s": rate, "tflops_algorithmic": rate * F_PAIR / 1e12,
            "frac_fp32_peak": rate * F_PAIR / 1e12 / PEAK_FP32_TFLOPS,
            "tiles": tune,
            "vs_plain": {k: rate / v for k, v in plain.items()},
            "vs_plain_note": "step: the plain one-walker step from Python (full action + 3 dV); pair_kernel: the "
                             "plain all-pairs kernel alone (events); cxx_total_action: pimc_total_action p50 wall "
                             "from a C++ caller"}


def served_leg(pimc_amd, plain_dv_rate, seconds):
    """1024 resident chains, N in [32,512], M in [64,512]; Poisson arrivals of
    Staging-style segment moves (4..64 slices) and 10% CenterOfMass moves
    (sigma 0.1 A); device-side Metropolis and commit."""
    from pimc_amd.workloads import LAMBDA, box_side, served_chains
    srv = pimc_amd.Server(12 << 30, streams=4, max_side=box_side(512) + 1.0)
    for w in served_chains():
        b = pimc_amd.api.gen_worldlines(w["seed"], w["side"], w["M"], w["N"], LAMBDA, w["tau"], w["kind"])
        srv.add_chain(w["side"], w["M"], w["N"], w["tau"], LAMBDA, w["rc"], b)
    threads = 4   # client threads submitting (the chains split among them)
    p_ssf = 0.01  # occasional S(q) requests: 257 lattice q (|n| <= 4) of the chain's box
    out = {"workload": "1024 chains, N in [32,512], M in [64,512], segment moves of 4..64 slices + 10% CoM + "
                       "1% S(q) requests (257 q), Poisson arrivals", "client_threads": threads, "rates": {}}
    for rate in (2e3, 2e4, 2e5):
        st = srv.run_synthetic(1024, rate=rate, duration_s=seconds, threads=threads, p_ssf=p_ssf, ssf_nmax=4)
        d = {k: float(v) for k, v in st.items()}
        d["vs_plain_delta_v_cxx"] = d["pair_evals_per_s"] / plain_dv_rate
        d["hbm_gb_per_s_algorithmic"] = d["pair_evals_per_s"] / 2 * B_PARTNER / 1e9
        out["rates"][f"{int(rate)}"] = d
    # closed loop (every chain always has a request in flight): the server's
    # throughput limit, with more client threads since submitting is the limit there
    closed_threads = 8
    st = srv.run_synthetic(1024, rate=0.0, duration_s=seconds, threads=closed_threads)
    out["closed_loop"] = {k: float(v) for k, v in st.items()}
    out["closed_loop"]["client_threads"] = closed_threads
    out["closed_loop"]["vs_plain_delta_v_cxx"] = out["closed_loop"]["pair_evals_per_s"] / plain_dv_rate
    out["closed_loop"]["hbm_gb_per_s_algorithmic"] = out["closed_loop"]["pair_evals_per_s"] / 2 * B_PARTNER / 1e9
    top = out["rates"]["200000"]
    out["pair_evals_per_s"] = top["pair_evals_per_s"]
    out["vs_plain_delta_v"] = top["pair_evals_per_s"] / plain_dv_rate
    out["vs_plain_note"] = ("served dV pair evaluations/s at the 200k req/s open-loop rate over the plain drop-in's "
                            "dV pair evaluations/s (pimc_delta_action from a C++ caller, segments 8/16/32)")
    srv.close()
    return out


def cpu_baseline(beads, L, M, N, tau, lam, pairs_full, pairs_dv):
    """The oracle (FP64 restatement of the reference loops), 1 core, bounded sample."""
    from oracle import pyoracle as orc
    st = orc.State(L, M, N, tau, lam, L)
    st.set_beads(beads)
    st.total_action()                     # warm-up (table pages in)
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < 10.0 and reps < 20:
        st.total_action()
        reps += 1
        t = time.perf_counter() - t0
    pairs = pairs_full * reps
    lat = []
    rng = np.random.default_rng(7)
    dv_pairs = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 5.0:
        s = int(rng.choice([8, 16, 32]))
        start, p = int(rng.integers(M)), int(rng.integers(N))
        sp = np.array([[(start + k) % M, p] for k in range(1, s)], dtype=np.int32)
        t1 = time.perf_counter()
        st.action_list(sp)
        st.action_list(sp)
        lat.append((time.perf_counter() - t1) * 1e6)
        dv_pairs += pairs_dv[s]
    dv_t = time.perf_counter() - t0
    # every host core this process may use (at most 16, a GPU's share of its box):
    # the slices' pair sums over that many threads, the same sum as the serial one
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    st.total_action(threads=cores)
    reps_mt, t_mt = 0, 0.0
    t0 = time.perf_counter()
    while t_mt < 5.0 and reps_mt < 50:
        st.total_action(threads=cores)
        reps_mt += 1
        t_mt = time.perf_counter() - t0
    return {"value": (pairs + dv_pairs) / (t + dv_t), "unit": "bead-pair evaluations/s", "cores": 1,
            "all_cores": {"cores": cores, "full_action_pairs_per_s": pairs_full * reps_mt / t_mt,
                          "sample": f"{reps_mt} full potential actions, slices over {cores} threads"},
            "kind": "port",
            "sample": f"{reps} full potential actions of the N=256 M=167 configuration ({t:.1f} s) + "
                      f"{len(lat)} Bisection dV (old+new Vnn lists, {dv_t:.1f} s), single thread",
            "full_action_pairs_per_s": pairs / t, "dv_p50_us": pct(lat, 50), "dv_p99_us": pct(lat, 99)}


if __name__ == "__main__":
    main()
