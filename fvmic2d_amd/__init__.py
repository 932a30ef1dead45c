"""fvmic2d_amd: the marker hot path of fvmic2d as HIP kernels for MI355X (gfx950).

The product is libfvmic2d_hip.so (C ABI in include/fvm_abi.h); ``lib`` binds it
with ctypes and ``host.MarkerPath`` mirrors the reference's function names.
"""
from . import abi  # noqa: F401

__all__ = ["abi", "lib", "host", "workloads", "build"]
