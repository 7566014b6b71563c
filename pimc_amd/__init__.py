"""pimc_amd: MI355X-native (gfx950 HIP) all-bead-pairs hot path of pimc.

The product is libpimc_gpu.so (C-ABI: include/pimc_gpu.h); pimc_amd.api is
its Python mirror.  Nothing in this package computes the physics on the host.
"""
from .api import (V_FP32, V_FP64, V_TABLE, Aziz, Batch, Chain, Context, PimcError,  # noqa: F401
                  Server, StateFile, copy_probe, device_synchronize, gen_worldlines, lib, pinned_empty,
                  run_chains_device)

__all__ = ["Context", "Aziz", "Batch", "Server", "Chain", "StateFile", "PimcError", "V_FP32", "V_FP64", "V_TABLE", "lib",
           "device_synchronize", "gen_worldlines", "copy_probe", "run_chains_device", "pinned_empty"]
