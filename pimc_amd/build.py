"""Builds libpimc_gpu.so (gfx950) in-tree with hipcc.

Every .hip file under csrc/ is compiled to an object under _build/ and linked
into pimc_amd/libpimc_gpu.so.  Packed FP32 instructions are disabled for all
device code (-packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpimc_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PIMC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    # keep the compiler's warnings visible, minus the known host-pass notice
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=True):
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(os.path.join(SRC, f) for f in os.listdir(SRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, warn in results:
        if warn and verbose:
            print(warn, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhsa-runtime64", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_dv_client()
    _build_adapter_check()
    return LIB


def _build_dv_client():
    """pimc_amd/dv_client: a C++ caller timing pimc_delta_action / pimc_total_action."""
    src = os.path.join(SRC, "dv_client.cpp")
    out = os.path.join(HERE, "dv_client")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), src, "-o", out,
           "-L", HERE, "-lpimc_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"dv_client build failed:\n{r.stderr}")


def _build_adapter_check():
    """tests/cpp/adapter_check: the C++ adapter (include/pimc_gpu.hpp) as a caller uses it."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "adapter_check.cpp")
    if not os.path.exists(src):
        return
    out = os.path.join(root, "tests", "cpp", "adapter_check")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src, "-o", out,
           "-L", HERE, "-lpimc_gpu", "-Wl,-rpath,$ORIGIN/../../pimc_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"adapter_check build failed:\n{r.stderr}")


if __name__ == "__main__":
    print(build())
