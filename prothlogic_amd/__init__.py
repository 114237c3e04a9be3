"""prothlogic_amd -- MI355X-native (gfx950) Proth modular-squaring engine.

The hot path (x <- x^2 mod k*2^n+1, n-1 times, residue resident in HBM) runs
in HIP kernels behind the C ABI of include/prothcore.h; `core` is its ctypes
binding, `workloads` the seeded benchmark candidates.
"""
from . import workloads  # noqa: F401

__all__ = ["core", "workloads"]
