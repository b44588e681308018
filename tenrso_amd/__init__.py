"""tenrso_amd -- MI355X (gfx950) implementation of tenrso's CP decomposition hot path.

Modules
  kernels  mttkrp / khatri_rao / cp_reconstruct (tenrso-kernels API)
  decomp   cp_als / cp_als_constrained / cp_als_with_init (tenrso-decomp CP API)
  synth    deterministic synthetic workloads
All compute goes through the HIP library libtenrso_hip.so (C ABI in
include/tenrso_hip.h); importing this package does not load it.
"""
from . import synth  # noqa: F401

__version__ = "0.1.0"
