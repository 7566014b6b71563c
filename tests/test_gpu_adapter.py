"""The C++ adapter (include/pimc_gpu.hpp), driven by tests/cpp/adapter_check the
way pimc's own callers drive LocalAction / PotentialBase, against the oracle
(table mode: rel <= 1e-12)."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_adapter_matches_oracle(orc):
    exe = os.path.join(ROOT, "tests", "cpp", "adapter_check")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe, "32", "16"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = json.loads(r.stdout)
    N, M, L = g["N"], g["M"], g["L"]
    b = orc.gen_worldlines(139853, L, M, N, 6.0596, 0.004, 1)
    st = orc.State(L, M, N, 0.004, 6.0596, L)
    st.set_beads(b)
    assert g["total"] == pytest.approx(st.total_action(), rel=1e-12)
    assert g["v3"][1] == pytest.approx(st.V_slice(3)[1], rel=1e-12)
    assert g["d3"] == pytest.approx(st.V_slice(3)[1], rel=1e-12)
    seg = [(2, 5), (3, 5), (4, 5), (5, 5)]
    assert g["list"] == pytest.approx(st.action_list(seg), rel=1e-12)
    assert g["range"] == pytest.approx(st.action_range((2, 5), (5, 5)), rel=1e-12)
    old = st.action_list(seg)
    for (m, i) in seg:
        st.update_bead(m, i, b[m, i] + np.array([0.3, -0.2, 0.1]))
    new = st.action_list(seg)
    assert g["dv"][0] == pytest.approx(old, rel=1e-12)
    assert g["dv"][1] == pytest.approx(new, rel=1e-12)
    st.set_beads(b)
    st.set_action(2, (1.0, 1.0))
    assert g["shifted"] == pytest.approx(st.total_action(), rel=1e-12)
    st.set_action(1, (1.0, 1.0))
    e = st.energy(mu=0.5)
    cols = ("K", "V", "V_ext", "V_int", "E", "E_mu", "K/N", "V/N", "E/N")
    for k, name in enumerate(cols):
        assert g["E"][k] == pytest.approx(e[name], rel=1e-10, abs=1e-12), name
    assert g["E"][5] == pytest.approx(g["E"][4] - 0.5 * N, rel=1e-14)
    # V(3) read through the cache after bead (3, 5) moved, and put back
    st.update_bead(3, 5, b[3, 5] + np.array([0.3, -0.2, 0.1]))
    assert g["v3_after"] == pytest.approx(st.V_slice(3)[1], rel=1e-12)
    st.set_beads(b)
    # GSF factors: the adapter's per-slice tau and lambda derivatives give the
    # estimator's thermodynamic V (src/estimator.cpp:985-1003) and the library's energy
    st.set_action(1, (2.0 / 3.0, 4.0 / 3.0))
    st.set_correction((0.0, 2.0 / 9.0), 2)
    eg = st.energy()
    tail = N * N / L ** 3 * g["tailV"]
    assert g["gsf_t1"] != 0.0
    assert g["gsf_t2"] - g["gsf_t1"] + tail == pytest.approx(eg["V"], rel=1e-12)
    for k, name in enumerate(cols[:5]):
        assert g["gsf_E"][k] == pytest.approx(eg[name], rel=1e-10), name
    az = orc.Aziz(1979, max_sep=np.sqrt(3) * L / 2, rc=L)
    assert g["V_rm"] == az.V([2.9673, 0.0, 0.0])
    assert g["tailV"] == az.tailV
