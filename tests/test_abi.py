"""CPU-side checks of the drop-in boundary: libpimc_gpu.so loads and exports
every entry point declared in include/pimc_gpu.h, and the Python mirror binds
each of them.  No compute call is made (there is no GPU in this container)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pimc_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pimc_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "pimc_ctx_create" in syms and "pimc_delta_action" in syms and "pimc_isf" in syms
    assert len(syms) >= 25


def test_library_exports_every_symbol():
    import ctypes
    path = os.path.join(ROOT, "pimc_amd", "libpimc_gpu.so")
    if not os.path.exists(path):
        pytest.fail("libpimc_gpu.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    from pimc_amd import api
    assert set(declared_symbols()) == set(api.SIGNATURES)
    api.lib()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from pimc_amd import api
    monkeypatch.setattr(api, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(api, "_lib", None)
    with pytest.raises(api.PimcError):
        api.lib()


def test_package_exports_the_mirror_classes():
    import pimc_amd
    for name in ("Context", "Aziz", "Batch", "Server", "Chain", "PimcError", "gen_worldlines",
                 "device_synchronize", "copy_probe", "V_FP32", "V_FP64", "V_TABLE"):
        assert hasattr(pimc_amd, name), name
    assert set(pimc_amd.__all__) <= set(dir(pimc_amd))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message before any
    device work."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_argument_checks_without_a_device():
    """Entry points that refuse bad arguments before touching the device do so
    here too: page-locked allocation of zero bytes or into a null pointer, and
    the server's dispatch counters of a null server."""
    import ctypes
    from pimc_amd import api
    E_ARG = -1   # PIMC_E_ARG (include/pimc_gpu.h)
    lib = api.lib()
    p = ctypes.c_void_p()
    assert lib.pimc_host_alloc(0, ctypes.byref(p)) == E_ARG and not p.value
    assert lib.pimc_host_alloc(64, None) == E_ARG
    assert lib.pimc_host_free(None) == 0
    a = ctypes.c_int64(-1)
    assert lib.pimc_server_dispatch_stats(None, ctypes.byref(a), None) == E_ARG and a.value == -1
