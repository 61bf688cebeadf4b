"""efs_amd: MI355X-native FDTD time-stepping engine.

The hot path is HIP for gfx950 in efs_amd/csrc, exposed through the C ABI in
include/efs_engine.h.  ``efs_amd.engine.HIPEngine`` is the Python host
binding with the reference WebGPUEngine's method names; efs_amd/js holds
the Node (N-API) binding.
"""
__version__ = "0.1.0"
