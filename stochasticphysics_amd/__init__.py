"""MI355X (gfx950) ensemble SDE engine for the StochasticPhysics particle simulator.

The compute path is libstophys_gpu.so (hand-written HIP kernels behind the C ABI in
include/stophys_gpu.h); this package binds it with ctypes and builds workloads.
"""
from . import capi, configs
from .engine import Context, Ensemble, Settings, compute_params, host_box_muller, host_normals

__all__ = ["Context", "Ensemble", "Settings", "capi", "configs", "compute_params", "host_box_muller", "host_normals"]
