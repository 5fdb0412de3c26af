"""MI355X-native single-partition MPM substep of iceMPM_2d_sp.

The product is the HIP library lib/libicempm.so (C-ABI in include/icempm.h);
this package is its host-side mirror: the reference's GPU_Implementation5 /
GPU_Partition / HostSideSOA surface (gpu.py), SimParams (params.py), ctypes
bindings (abi.py) and synthetic scenes (scenes.py).
"""
from . import abi
from .params import SimParams, sea_ice_params

__all__ = ["abi", "SimParams", "sea_ice_params"]
