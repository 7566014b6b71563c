"""S(q) and F(q,tau) on the GPU at the full sizes of the two estimator workloads,
value by value against FP64 density modes (tests/_oracle_util.lattice_rho,
itself pinned to the oracle's direct pair sums in test_oracle_estimators.py):

  N=256,  M=167, tau=0.004:  5 377 lattice q (|q| <= 3 A^-1), all 167 lags;
  N=2048, M=250, tau=0.004:  100 401 lattice q (|q| <= 4 A^-1), all 250 lags,
                             Gaussian-bridge worldlines;

plus the oracle's direct minimum-image pair sum (src/estimator.cpp:3705-3737) on
the largest-|n| q at N=2048, and the exact identities S(0) = N M, F(q,0) = S(q),
F(q,tau) = F(q,M-tau).

Tolerance (the north star's 1e-5 on S(q) and F(q,tau)): with values summed
over the M slices (the estimator before its 1/M norm),
    |S_gpu(q) - S_ref(q)|         <= 1e-5 * max(S_ref(q), M)
    |F_gpu(q,tau) - F_ref(q,tau)| <= 1e-5 * max(S_ref(q), M)
i.e. relative to S(q) itself, with the ideal-gas value S = 1 per slice as the
floor where the structure factor of the worldlines is smaller than that; F(q,tau)
is measured on the scale of its own lag-0 value S(q).  The pure relative error
|S_gpu - S_ref| / S_ref is held to 1e-5 as well."""
import math

import numpy as np
import pytest

from tests._oracle_util import isf_from_rho, lattice_rho

pytestmark = pytest.mark.gpu

DENSITY, LAMBDA, SEED0 = 0.02198, 6.0596, 139853
TOL = 1e-5


def _case(orc, N, M, kind, qmax):
    import pimc_amd
    from pimc_amd.workloads import lattice_q
    L = orc.box_side(N, DENSITY)
    beads = orc.gen_worldlines(SEED0, L, M, N, LAMBDA, 0.004, kind)
    q = lattice_q(L, qmag_max=qmax)
    nvec = np.rint(q * L / (2.0 * math.pi)).astype(np.int64)
    rc = 0.5 * L if N == 2048 else L
    ctx = pimc_amd.Context(L, M, N, 0.004, LAMBDA, rc)
    ctx.upload_path(beads)
    print(f"N={N} M={M}: {q.shape[0]} q, |n_d| <= {np.abs(nvec).max()}", flush=True)
    s_gpu = ctx.ssf(q)
    f_gpu = ctx.isf(q)
    extra = {}
    if N == 256:
        # 7.2 MB into page-locked memory: two halves, each copied out while the
        # other computes
        f_pin = pimc_amd.pinned_empty((q.shape[0], M))
        ctx.isf(q, out=f_pin)
        extra["chunks"] = ctx.kernel_time("isf_chunks")[1]
        extra["f_pin"] = [np.array(f_pin)]
    if N == 2048:
        # into page-locked memory: the chunked form (modes and lags per chunk of
        # q, each chunk's rows copied out while the next computes), twice in a
        # row; then with the rho kernel timed, which keeps one launch per kernel
        f_pin = [pimc_amd.pinned_empty((q.shape[0], M)) for _ in range(3)]
        ctx.isf(q, out=f_pin[0])
        ctx.isf(q, out=f_pin[1])
        extra["chunks"] = ctx.kernel_time("isf_chunks")[1]
        ctx.set_timing(True, kinds=["rho"])
        ctx.isf(q, out=f_pin[2])
        ctx.set_timing(False)
        extra["chunks_timed"] = ctx.kernel_time("isf_chunks")[1] - extra["chunks"]
        extra["f_pin"] = [np.array(f) for f in f_pin]
    ctx.close()
    rho = lattice_rho(beads, nvec, L)
    f_ref = isf_from_rho(rho, N)
    return dict(L=L, beads=beads, q=q, n=nvec, s_gpu=s_gpu, f_gpu=f_gpu, s_ref=f_ref[:, 0].copy(), f_ref=f_ref,
                N=N, M=M, **extra)


@pytest.fixture(scope="module")
def walker(orc):
    return _case(orc, 256, 167, 0, 3.0)


@pytest.fixture(scope="module")
def large(orc):
    return _case(orc, 2048, 250, 1, 4.0)


def _check_values(c):
    M = c["M"]
    scale = np.maximum(c["s_ref"], M)
    es = np.abs(c["s_gpu"] - c["s_ref"]) / scale
    ef = np.abs(c["f_gpu"] - c["f_ref"]) / scale[:, None]
    rel = np.abs(c["s_gpu"] - c["s_ref"]) / c["s_ref"]
    qi, ti = np.unravel_index(np.argmax(ef), ef.shape)
    print(f"N={c['N']} M={M}: S err/max(S,M) {es.max():.3e} (pure rel {rel.max():.3e} at S/M = "
          f"{c['s_ref'][np.argmax(rel)] / M:.3g}); F err/max(S,M) {ef.max():.3e} at n={c['n'][qi]}, tau={ti}",
          flush=True)
    assert es.max() <= TOL
    assert rel.max() <= TOL
    assert ef.max() <= TOL


def _check_identities(c):
    N, M, s, f = c["N"], c["M"], c["s_gpu"], c["f_gpu"]
    i0 = int(np.where(np.all(c["n"] == 0, axis=1))[0][0])
    assert s[i0] == pytest.approx(N * M, rel=1e-12)
    assert np.allclose(f[i0], N * M, rtol=1e-12)
    assert np.allclose(f[:, 0], s, rtol=1e-12, atol=1e-9)
    assert np.allclose(f[:, 1:], f[:, :0:-1], rtol=1e-10, atol=1e-7 * M)


def test_isf_n256_m167_all_q_all_lags(walker):
    assert walker["q"].shape[0] == 5377
    _check_identities(walker)
    _check_values(walker)


def test_isf_n256_m167_halves_into_pinned(walker):
    """F(q,tau) of the 5377 q into page-locked memory runs as two halves of the
    q list: bit-identical to the one-launch result."""
    assert walker["chunks"] == 2
    assert np.array_equal(walker["f_pin"][0], walker["f_gpu"])


def test_ssf_isf_n2048_m250_all_q_all_lags(large):
    assert large["q"].shape[0] == 100401
    # the largest |n_d| the q set reaches (|q| <= 4 A^-1: n_d up to 28 on an axis)
    assert np.abs(large["n"]).max() == 28
    _check_identities(large)
    _check_values(large)


def test_ssf_n2048_m250_oracle_pair_sum(orc, large):
    """The oracle's direct pair sum over all 250 slices of N=2048 on the q with
    the largest |n| along an axis, (0, 0, 28) 2 pi / L."""
    c = large
    k = int(np.where(np.all(c["n"] == [0, 0, 28], axis=1))[0][0])
    st = orc.State(c["L"], c["M"], c["N"], 0.004, LAMBDA, 0.5 * c["L"])
    st.set_beads(c["beads"])
    s_orc = st.ssf(c["q"][[k]])[0]
    err = abs(c["s_gpu"][k] - s_orc) / max(s_orc, c["M"])
    print(f"S(0,0,28): gpu {c['s_gpu'][k]:.10g} oracle {s_orc:.10g} fp64 phasor {c['s_ref'][k]:.10g}, err {err:.2e}")
    assert abs(c["s_ref"][k] - s_orc) <= 1e-10 * s_orc
    assert err <= TOL


def test_isf_n2048_m250_chunked_into_pinned(large):
    """F(q,tau) of the 100 401 q into page-locked memory takes the chunked form
    (3 chunks of at most 34 000 q, the last one ragged): bit-identical to the one-launch
    result, on two calls in a row, and not taken while the rho kernel is timed."""
    assert large["chunks"] == 2 * 3
    assert large["chunks_timed"] == 0
    for f in large["f_pin"]:
        assert np.array_equal(f, large["f_gpu"])
